// plotTracker (reference include/plotter/plotTracker.h:15-160): each tracked
// object's histories written to a shared-memory file (int32 sample count,
// int32 id, then six float32 histories of that many samples) and handed to the
// case's plotting script, ../../include/plotter/plotTracker.py, as the
// reference does.  A missing or failing script is reported and the simulation
// goes on.
#pragma once
#include "../types.h"

inline void plotTracker(const TrackerStruct &Tracker)
{
    if (!(TRACK_WALL_FORCE || TRACK_ROTOR_FORCE || TRACK_OPEN_BOUNDARIES || Tracker.TRACK_CUSTOM_VARIABLES)) return;
    const int count = std::min(Tracker.iterationsFinished, ITERATION_COUNT);
    if (count <= 0) return;
    const std::string script = "../../include/plotter/plotTracker.py";
    auto run = [&](const char *kind, const std::string &path) {
        const std::string command = "python3 '" + script + "' " + kind + " '" + path + "' --average-percent " +
                                    std::to_string((double)Tracker.averagePercent);
        if (std::system(command.c_str()) != 0)
            std::cerr << "Warning: tracker plot failed for " << kind << ". Continuing simulation.\n";
        std::remove(path.c_str());
    };
    auto object = [&](const char *kind, int id, std::array<const FloatArray2DTypeCPU *, 6> h) {
        char path[] = "/dev/shm/lbmTrackerXXXXXX";
        const int fd = mkstemp(path);
        if (fd < 0) { std::cerr << "Warning: cannot create a tracker file. Continuing simulation.\n"; return; }
        FILE *fp = fdopen(fd, "wb");
        if (!fp) { close(fd); std::remove(path); return; }
        const int32_t header[2] = {count, id};
        bool ok = std::fwrite(header, sizeof(int32_t), 2, fp) == 2;
        std::vector<float> row((size_t)count);
        for (const FloatArray2DTypeCPU *a : h) {
            for (int s = 0; s < count; ++s) row[(size_t)s] = (*a)(id, s);
            ok = ok && std::fwrite(row.data(), sizeof(float), row.size(), fp) == row.size();
        }
        ok = std::fclose(fp) == 0 && ok;
        if (!ok) { std::remove(path); std::cerr << "Warning: tracker file incomplete. Continuing simulation.\n"; return; }
        run(kind, path);
    };
    if (TRACK_OPEN_BOUNDARIES)
        for (int id = 0; id < Tracker.openBCCount; ++id)
            object("boundary", id, {&Tracker.normalVelocityArray, &Tracker.massFlowArray, &Tracker.momentumThrustArray,
                                    &Tracker.pressureArray, &Tracker.pressurePowerArray, &Tracker.normalKineticPowerArray});
    if (TRACK_WALL_FORCE)
        for (int id = 0; id < Tracker.wallCount; ++id)
            object("wall", id, {&Tracker.wallFxArray, &Tracker.wallFyArray, &Tracker.wallFzArray, &Tracker.wallTxArray,
                                &Tracker.wallTyArray, &Tracker.wallTzArray});
    if (TRACK_ROTOR_FORCE)
        for (int id = 0; id < Tracker.rotorCount; ++id)
            object("rotor", id, {&Tracker.rotorFxArray, &Tracker.rotorFyArray, &Tracker.rotorFzArray, &Tracker.rotorTxArray,
                                 &Tracker.rotorTyArray, &Tracker.rotorTzArray});
    if (Tracker.TRACK_CUSTOM_VARIABLES && Tracker.customArray.getSize<0>() > 0) {
        // int32 count, int32 slots; per slot: uint32 + name bytes, uint32 + unit bytes, count float32 values
        char path[] = "/dev/shm/lbmTrackerCustomXXXXXX";
        const int fd = mkstemp(path);
        if (fd < 0) return;
        FILE *fp = fdopen(fd, "wb");
        if (!fp) { close(fd); std::remove(path); return; }
        const int32_t slots = (int32_t)Tracker.customArray.getSize<0>();
        const int32_t header[2] = {count, slots};
        bool ok = std::fwrite(header, sizeof(int32_t), 2, fp) == 2;
        std::vector<float> row((size_t)count);
        for (int32_t v = 0; v < slots; ++v) {
            const std::string name = (size_t)v < Tracker.customNames.size() ? Tracker.customNames[(size_t)v] : "";
            const std::string unit = (size_t)v < Tracker.customUnits.size() ? Tracker.customUnits[(size_t)v] : "";
            for (const std::string *t : {&name, &unit}) {
                const uint32_t n = (uint32_t)t->size();
                ok = ok && std::fwrite(&n, sizeof(n), 1, fp) == 1 && std::fwrite(t->data(), 1, n, fp) == n;
            }
            for (int s = 0; s < count; ++s) row[(size_t)s] = Tracker.customArray(v, s);
            ok = ok && std::fwrite(row.data(), sizeof(float), row.size(), fp) == row.size();
        }
        ok = std::fclose(fp) == 0 && ok;
        if (ok) run("custom", path);
        else std::remove(path);
    }
}
