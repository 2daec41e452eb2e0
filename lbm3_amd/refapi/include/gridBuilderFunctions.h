// buildGrids (reference include/gridBuilderFunctions.h:1577-1670) on the
// MI355X library.  A case file includes this header after defining its four
// device hooks; the kernels below call them, compiled into the case's own
// translation unit as the reference inlines them into its builder and step
// kernels, and the library reaches them through lbm3_hooks callbacks:
//   getRefinementModifier  markerFunctions.h:258-276
//   getInitialCondition    applyInitialCondition.h:22-35
//   getOpenBC              gridBuilderFunctions.h:1232-1300,1486-1503; updateGrid.h:256
//   getLocalBC             updateGrid.h:86-95
// The case's compile-time constants (GRID_LEVEL_COUNT, WALL_REFINEMENT_COUNT,
// RES_GLOBAL, DT_PHYS_GLOBAL, NU_PHYS, RHO_PHYS, TRACK_*) become the
// simulation's lbm3_case.
#pragma once
#include "types.h"
#include "cellFunctions.h"
#include "STLFunctions.h"

namespace lbm3facade {

// the case's grids while a simulation built from them lives: the hooks see the
// Info fields the case sets on the host (iRegulatorInlet, iRegulatorOutlet,
// useRotors) as the reference's kernels do
inline std::vector<GridStruct> *&hook_grids()
{
    static std::vector<GridStruct> *g = nullptr;
    return g;
}

inline InfoStruct hook_info(const lbm3_info *info)
{
    InfoStruct I = info_from(*info);
    std::vector<GridStruct> *g = hook_grids();
    const int L = (int)info->gridID;
    if (g && L >= 0 && L < (int)g->size() && (*g)[L].sim) {
        I.iRegulatorInlet = (*g)[L].Info.iRegulatorInlet;
        I.iRegulatorOutlet = (*g)[L].Info.iRegulatorOutlet;
        I.useRotors = (*g)[L].Info.useRotors;
    }
    return I;
}

__global__ void k_hook_refinement(InfoStruct Info, const int *ci, const int *cj, const int *ck, int n, uint8_t *out)
{
    const int c = blockIdx.x * 256 + threadIdx.x;
    if (c >= n) return;
    const int i = ci[c], j = cj[c], k = ck[c];
    bool fromFalse = false, fromTrue = true;
    getRefinementModifier(i, j, k, fromFalse, Info);
    getRefinementModifier(i, j, k, fromTrue, Info);
    out[c] = (uint8_t)((fromFalse ? 1u : 0u) | (fromTrue ? 2u : 0u));
}

__global__ void k_hook_initial(InfoStruct Info, const int *ci, const int *cj, const int *ck, int n, float *out4)
{
    const int c = blockIdx.x * 256 + threadIdx.x;
    if (c >= n) return;
    const int i = ci[c], j = cj[c], k = ck[c];
    BCStruct BC;
    getInitialCondition(BC, i, j, k, Info);
    out4[4 * (size_t)c] = BC.dRho;
    out4[4 * (size_t)c + 1] = BC.ux;
    out4[4 * (size_t)c + 2] = BC.uy;
    out4[4 * (size_t)c + 3] = BC.uz;
}

__global__ void k_hook_open(InfoStruct Info, const int *ci, const int *cj, const int *ck, int n, lbm3_bc *out)
{
    const int c = blockIdx.x * 256 + threadIdx.x;
    if (c >= n) return;
    const int i = ci[c], j = cj[c], k = ck[c];
    BCStruct BC;
    getOpenBC(BC, i, j, k, Info);
    out[c] = bc_to(BC);
}

__global__ void k_hook_local(InfoStruct Info, const int *ci, const int *cj, const int *ck, const lbm3_bc *in, int n, lbm3_bc *out)
{
    const int c = blockIdx.x * 256 + threadIdx.x;
    if (c >= n) return;
    const int i = ci[c], j = cj[c], k = ck[c];
    BCStruct BC = bc_from(in[c]);
    getLocalBC(BC, i, j, k, Info);
    out[c] = bc_to(BC);
}

inline int launched() { return hipGetLastError() == hipSuccess ? 0 : 1; }
inline dim3 grid_for(int n) { return dim3((unsigned)((n + 255) / 256)); }

inline int cb_refinement(void *, const lbm3_info *info, const int *i, const int *j, const int *k, int n, uint8_t *out, void *stream)
{
    if (n > 0) hipLaunchKernelGGL(k_hook_refinement, grid_for(n), dim3(256), 0, (hipStream_t)stream, hook_info(info), i, j, k, n, out);
    return launched();
}
inline int cb_initial(void *, const lbm3_info *info, const int *i, const int *j, const int *k, int n, float *out4, void *stream)
{
    if (n > 0) hipLaunchKernelGGL(k_hook_initial, grid_for(n), dim3(256), 0, (hipStream_t)stream, hook_info(info), i, j, k, n, out4);
    return launched();
}
inline int cb_open(void *, const lbm3_info *info, const int *i, const int *j, const int *k, int n, lbm3_bc *out, void *stream)
{
    if (n > 0) hipLaunchKernelGGL(k_hook_open, grid_for(n), dim3(256), 0, (hipStream_t)stream, hook_info(info), i, j, k, n, out);
    return launched();
}
inline int cb_local(void *, const lbm3_info *info, const int *i, const int *j, const int *k, const lbm3_bc *in, int n, lbm3_bc *out,
                    void *stream)
{
    if (n > 0)
        hipLaunchKernelGGL(k_hook_local, grid_for(n), dim3(256), 0, (hipStream_t)stream, hook_info(info), i, j, k, in, n, out);
    return launched();
}

// storage of the populations: FP32 as the reference; a case may choose the
// 16-bit stores by defining LBM3_POPULATION_STORAGE before this header
#ifndef LBM3_POPULATION_STORAGE
#define LBM3_POPULATION_STORAGE LBM3_STORE_F32
#endif

template <class T> inline void borrow_array(lbm3_sim *sim, int level, const char *name, TNL::Containers::Vector<T, TNL::Devices::Cuda, size_t> &a)
{
    void *p = nullptr;
    long long n = 0;
    check(lbm3_device_array(sim, level, name, &p, &n), name);
    a.borrow(static_cast<T *>(p), n);
}

// the step state of every level after the library moved it on
inline void sync_step_state(std::vector<GridStruct> &grids)
{
    for (GridStruct &G : grids) {
        if (!G.sim) continue;
        lbm3_info I;
        check(lbm3_get_info(G.sim, G.level, &I), "lbm3_get_info");
        G.Info.iterationsFinished = I.iterationsFinished;
        G.Info.iterationsSinceTrackerReset = I.iterationsSinceTrackerReset;
        G.Info.esotwistFlipper = I.esotwistFlipper != 0;
        for (size_t r = 0; r < G.rotors.size(); ++r) {
            double ri[12];
            if (lbm3_rotor_info(G.sim, G.level, (int)r, ri) == 0) G.rotors[r].Info.radiansElapsed = ri[10];
        }
    }
}

// the grid arrays and Info of a built level, as views of the simulation's buffers
inline void attach_level(GridStruct &G, lbm3_sim *sim, int level)
{
    G.sim = sim;
    G.level = level;
    lbm3_info I;
    check(lbm3_get_info(sim, level, &I), "lbm3_get_info");
    const bool useRotors = G.Info.useRotors;
    G.Info = info_from(I);
    G.Info.useRotors = useRotors;
    G.sentUseRotors = useRotors;
    int tile = 256, bytes = 4;
    check(lbm3_population_layout(sim, &tile, &bytes), "lbm3_population_layout");
    if (bytes != 4) throw std::runtime_error("the case-facing population view reads FP32 storage");
    G.fArray.sim = sim;
    G.fArray.level = level;
    G.fArray.tileShift = 0;
    while ((1 << G.fArray.tileShift) < tile) ++G.fArray.tileShift;
    borrow_array(sim, level, "shifter", G.IJKNBR.shifterArray);
    borrow_array(sim, level, "rowI", G.IJKNBR.iArray);
    borrow_array(sim, level, "rowJ", G.IJKNBR.jArray);
    borrow_array(sim, level, "rowK", G.IJKNBR.kArray);
    borrow_array(sim, level, "rowJPlus", G.IJKNBR.jPlusArray);
    borrow_array(sim, level, "rowKPlus", G.IJKNBR.kPlusArray);
    borrow_array(sim, level, "rowJKPlus", G.IJKNBR.jkPlusArray);
    borrow_array(sim, level, "wallMap", G.Wall.wallMapArray);
    borrow_array(sim, level, "wallIndex", G.Wall.indexArray);
    borrow_array(sim, level, "wallData", G.Wall.wallDataArray);
    borrow_array(sim, level, "linkLength", G.Wall.linkLengthArray);
    G.Wall.wallAdjacentCount = (int)G.Wall.indexArray.getSize();
    borrow_array(sim, level, "periodicIndex", G.PeriodicBC.indexArray);
    borrow_array(sim, level, "periodicSource", G.PeriodicBC.sourceArray);
    G.PeriodicBC.periodicBCCount = (int)G.PeriodicBC.indexArray.getSize();
    // open boundaries: the library keeps every BCID's cells in one list, BCID order
    long long nIds = 0;
    check(lbm3_download_array(sim, level, "openCount", nullptr, 0, &nIds), "openCount");
    std::vector<int> counts((size_t)nIds);
    if (nIds > 0) check(lbm3_download_array(sim, level, "openCount", counts.data(), nIds, &nIds), "openCount");
    IntArrayType all, src;
    borrow_array(sim, level, "openIndex", all);
    borrow_array(sim, level, "openSource", src);
    G.openBCs.assign((size_t)nIds, OpenBCArrayStruct{});
    long long off = 0;
    for (long long b = 0; b < nIds; ++b) {
        OpenBCArrayStruct &O = G.openBCs[(size_t)b];
        O.openBCID = (int)b;
        O.openBCCount = counts[(size_t)b];
        O.indexArray.borrow(all.getData() + off, counts[(size_t)b]);
        O.sourceArray.borrow(src.getData() + off, counts[(size_t)b]);
        off += counts[(size_t)b];
    }
    // rotors of the level
    G.rotors.clear();
    for (int r = 0; r < LBM3_MAX_ROTORS; ++r) {
        double ri[12];
        if (lbm3_rotor_info(sim, level, r, ri) != 0) break;
        RotorStruct R;
        R.Info.rotorID = (float)r;
        R.Info.res = (float)ri[0];
        R.Info.Bounds.xMin = (float)ri[1]; R.Info.Bounds.yMin = (float)ri[2]; R.Info.Bounds.zMin = (float)ri[3];
        R.Info.Bounds.xMax = (float)ri[4]; R.Info.Bounds.yMax = (float)ri[5]; R.Info.Bounds.zMax = (float)ri[6];
        R.Info.cellCountX = (int)ri[7]; R.Info.cellCountY = (int)ri[8]; R.Info.cellCountZ = (int)ri[9];
        R.Info.radiansElapsed = ri[10];
        G.rotors.push_back(R);
    }
}

} // namespace lbm3facade

inline long long buildGrids(std::vector<GridStruct> &grids, std::vector<STLStruct> &gridStaticSTLs,
                            std::vector<STLStruct> &rotorSTLs, BoundsStruct DomainBounds)
{
    using namespace lbm3facade;
    if ((int)grids.size() < GRID_LEVEL_COUNT) grids.resize(GRID_LEVEL_COUNT);
    lbm3_case c{};
    c.gridLevelCount = GRID_LEVEL_COUNT;
    c.wallRefinementCount = WALL_REFINEMENT_COUNT;
    c.resGlobal = RES_GLOBAL;
    c.dtPhysGlobal = DT_PHYS_GLOBAL;
    c.nuPhys = NU_PHYS;
    c.rhoPhys = RHO_PHYS;
    c.trackWallForce = TRACK_WALL_FORCE;
    c.trackRotorForce = TRACK_ROTOR_FORCE;
    c.trackOpenBoundaries = TRACK_OPEN_BOUNDARIES;
    c.openDefault = bc_to(BCStruct{});
    lbm3_sim *sim = nullptr;
    check(lbm3_sim_create(&c, LBM3_POPULATION_STORAGE, &sim), "lbm3_sim_create");
    lbm3_hooks h{};
    h.refinement = cb_refinement;
    h.initial_condition = cb_initial;
    h.open_bc = cb_open;
    h.local_bc = cb_local;
    check(lbm3_set_hooks(sim, &h), "lbm3_set_hooks");
    hook_grids() = &grids;
    for (GridStruct &G : grids) G.sim = nullptr;
    // the static bodies, one wall id each in order (STLs as readSTL left them)
    std::vector<float> tris;
    std::vector<int> counts;
    for (const STLStruct &S : gridStaticSTLs) {
        tris.insert(tris.end(), S.triangles.begin(), S.triangles.end());
        counts.push_back(S.triangleCount);
    }
    if (tris.empty()) tris.assign(9, 0.f);
    if (counts.empty()) counts.push_back(0);
    long long fluidUpdatesPerIteration = 0;
    check(lbm3_build_grids(sim, tris.data(), counts.data(), (int)gridStaticSTLs.size(),
                           bounds_to(DomainBounds), &fluidUpdatesPerIteration), "buildGrids");
    // rotors on the levels whose Info.useRotors the case left on (:1611-1619)
    if (!rotorSTLs.empty()) {
        std::vector<float> rt;
        std::vector<int> rc;
        for (const STLStruct &S : rotorSTLs) {
            rt.insert(rt.end(), S.triangles.begin(), S.triangles.end());
            rc.push_back(S.triangleCount);
        }
        int mask = 0;
        for (int L = 0; L < GRID_LEVEL_COUNT; ++L)
            if (grids[L].Info.useRotors) mask |= 1 << L;
        check(lbm3_build_rotors(sim, rt.data(), rc.data(), (int)rc.size(), mask), "buildRotors");
    }
    for (int L = 0; L < GRID_LEVEL_COUNT; ++L) attach_level(grids[L], sim, L);
    std::cout << "Total fluid cell updates per iteration: " << fluidUpdatesPerIteration << std::endl;
    std::cout << "Applying initial condition ... " << std::flush;
    for (int L = 0; L < GRID_LEVEL_COUNT; ++L) check(lbm3_apply_initial_condition(sim, L), "applyInitialCondition");
    check(lbm3_synchronize(sim), "applyInitialCondition");
    std::cout << "Done" << std::endl;
    return fluidUpdatesPerIteration;
}
