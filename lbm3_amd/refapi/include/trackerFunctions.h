// Tracker (reference include/trackerFunctions.h:9-520): wall, rotor and
// open-boundary families in physical units per tracker period, kept as
// histories indexed by iteration, plus the case's custom variables.  The sums
// come from the library's FP64 reductions (lbm3_update_tracker), which also
// reset the accumulators as the reference does at the end of updateTracker.
#pragma once
#include "types.h"
#include "gridBuilderFunctions.h"

inline void initializeTracker(TrackerStruct &Tracker, std::vector<GridStruct> &grids)
{
    std::cout << "Initializing Tracker ... " << std::flush;
    for (GridStruct &G : grids) {
        Tracker.openBCCount = std::max(Tracker.openBCCount, (int)G.openBCs.size());
        Tracker.rotorCount = std::max(Tracker.rotorCount, (int)G.rotors.size());
        // wall count: the largest wall id of a wall-adjacent cell, plus one
        long long W = 0;
        lbm3facade::check(lbm3_download_array(G.sim, G.level, "wallData", nullptr, 0, &W), "wallData");
        std::vector<uint32_t> wd((size_t)W);
        if (W > 0) lbm3facade::check(lbm3_download_array(G.sim, G.level, "wallData", wd.data(), W, &W), "wallData");
        for (uint32_t d : wd) Tracker.wallCount = std::max(Tracker.wallCount, (int)((d >> 27) & 31u) + 1);
    }
    auto hist = [](FloatArray2DTypeCPU &a, int rows) { a.setSizes(rows, ITERATION_COUNT); a.setValue(0.f); };
    if (TRACK_OPEN_BOUNDARIES) {
        for (FloatArray2DTypeCPU *a : {&Tracker.normalVelocityArray, &Tracker.massFlowArray, &Tracker.momentumThrustArray,
                                       &Tracker.pressureArray, &Tracker.pressurePowerArray, &Tracker.normalKineticPowerArray})
            hist(*a, Tracker.openBCCount);
        for (FloatArrayTypeCPU *a : {&Tracker.normalVelocity, &Tracker.massFlow, &Tracker.momentumThrust, &Tracker.pressure,
                                     &Tracker.pressurePower, &Tracker.normalKineticPower})
            a->setSize(Tracker.openBCCount);
    }
    if (TRACK_WALL_FORCE) {
        for (FloatArray2DTypeCPU *a : {&Tracker.wallFxArray, &Tracker.wallFyArray, &Tracker.wallFzArray, &Tracker.wallTxArray,
                                       &Tracker.wallTyArray, &Tracker.wallTzArray})
            hist(*a, Tracker.wallCount);
        for (FloatArrayTypeCPU *a : {&Tracker.wallFx, &Tracker.wallFy, &Tracker.wallFz, &Tracker.wallTx, &Tracker.wallTy, &Tracker.wallTz})
            a->setSize(Tracker.wallCount);
    }
    if (TRACK_ROTOR_FORCE) {
        for (FloatArray2DTypeCPU *a : {&Tracker.rotorFxArray, &Tracker.rotorFyArray, &Tracker.rotorFzArray, &Tracker.rotorTxArray,
                                       &Tracker.rotorTyArray, &Tracker.rotorTzArray})
            hist(*a, Tracker.rotorCount);
        for (FloatArrayTypeCPU *a : {&Tracker.rotorFx, &Tracker.rotorFy, &Tracker.rotorFz, &Tracker.rotorTx, &Tracker.rotorTy, &Tracker.rotorTz})
            a->setSize(Tracker.rotorCount);
    }
    if (Tracker.TRACK_CUSTOM_VARIABLES)
        hist(Tracker.customArray, (int)std::max(Tracker.customNames.size(), Tracker.customUnits.size()));
    std::cout << "Done" << std::endl << std::endl;
}

inline void updateTracker(TrackerStruct &Tracker, std::vector<GridStruct> &grids)
{
    Tracker.iterationsFinished = grids[0].Info.iterationsFinished;
    const int at = grids[0].Info.iterationsFinished - 1;
    const int nO = TRACK_OPEN_BOUNDARIES ? Tracker.openBCCount : 0;
    const int nW = TRACK_WALL_FORCE ? Tracker.wallCount : 0;
    const int nR = TRACK_ROTOR_FORCE ? Tracker.rotorCount : 0;
    std::vector<double> walls(6 * (size_t)std::max(nW, 1)), rotors(6 * (size_t)std::max(nR, 1)), open(7 * (size_t)std::max(nO, 1));
    lbm3facade::check(lbm3_update_tracker(grids[0].sim, walls.data(), nW, rotors.data(), nR, open.data(), nO), "updateTracker");
    const bool keep = at >= 0 && at < ITERATION_COUNT;
    for (int b = 0; b < nO; ++b) {
        // out7 = {area, normal velocity, mass flow, momentum thrust, pressure, pressure power, normal kinetic power}
        FloatArrayTypeCPU *now[6] = {&Tracker.normalVelocity, &Tracker.massFlow, &Tracker.momentumThrust, &Tracker.pressure,
                                     &Tracker.pressurePower, &Tracker.normalKineticPower};
        FloatArray2DTypeCPU *h[6] = {&Tracker.normalVelocityArray, &Tracker.massFlowArray, &Tracker.momentumThrustArray,
                                     &Tracker.pressureArray, &Tracker.pressurePowerArray, &Tracker.normalKineticPowerArray};
        for (int q = 0; q < 6; ++q) {
            const float v = (float)open[7 * (size_t)b + 1 + q];
            (*now[q])[b] = v;
            if (keep) (*h[q])(b, at) = v;
        }
    }
    auto store = [&](int n, const std::vector<double> &src, std::array<FloatArrayTypeCPU *, 6> now,
                     std::array<FloatArray2DTypeCPU *, 6> h) {
        for (int b = 0; b < n; ++b)
            for (int q = 0; q < 6; ++q) {
                const float v = (float)src[6 * (size_t)b + q];
                (*now[q])[b] = v;
                if (keep) (*h[q])(b, at) = v;
            }
    };
    store(nW, walls, {&Tracker.wallFx, &Tracker.wallFy, &Tracker.wallFz, &Tracker.wallTx, &Tracker.wallTy, &Tracker.wallTz},
          {&Tracker.wallFxArray, &Tracker.wallFyArray, &Tracker.wallFzArray, &Tracker.wallTxArray, &Tracker.wallTyArray,
           &Tracker.wallTzArray});
    store(nR, rotors, {&Tracker.rotorFx, &Tracker.rotorFy, &Tracker.rotorFz, &Tracker.rotorTx, &Tracker.rotorTy, &Tracker.rotorTz},
          {&Tracker.rotorFxArray, &Tracker.rotorFyArray, &Tracker.rotorFzArray, &Tracker.rotorTxArray, &Tracker.rotorTyArray,
           &Tracker.rotorTzArray});
    lbm3facade::sync_step_state(grids);   // iterationsSinceTrackerReset is 0 again
}

inline void trackCustomVariables(TrackerStruct &Tracker, std::initializer_list<float> values)
{
    if (!Tracker.TRACK_CUSTOM_VARIABLES) return;
    const long long count = (long long)std::max({values.size(), Tracker.customNames.size(), Tracker.customUnits.size()});
    if (count > Tracker.customArray.getSize<0>()) {
        FloatArray2DTypeCPU grown;
        grown.setSizes(count, ITERATION_COUNT);
        grown.setValue(0.f);
        for (long long v = 0; v < Tracker.customArray.getSize<0>(); ++v)
            for (int s = 0; s < ITERATION_COUNT; ++s) grown(v, s) = Tracker.customArray(v, s);
        Tracker.customArray = std::move(grown);
    }
    const int at = Tracker.iterationsFinished - 1;
    if (at < 0 || at >= ITERATION_COUNT) return;
    int v = 0;
    for (const float x : values) Tracker.customArray(v++, at) = x;
}
