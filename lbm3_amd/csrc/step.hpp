// Launch wrappers for the per-step kernels (step.hip).
#pragma once
#include "device.hpp"

#define LBM3_MAX_BODIES 31

namespace lbm3 {

struct InitDesc {
    int kind;                 // 0 uniform, 1 Taylor-Green, 2 per-cell fields
    float dRho, ux, uy, uz;
    float u0, length;
    int pressure;
    const float *fields;      // kind 2: [N][4] (dRho, ux, uy, uz), device
    lbm3_profile profile;     // kind 0: velocity scaled by the profile factor of the cell
};

// one level's part of a section cut (exportSectionCutPlotGeneral)
struct CutLevel {
    int plane;                // 0 XY (normal z), 1 ZY (normal x), 2 ZX (normal y)
    int lower;                // normal index of the level's plane
    float alpha;              // weight of the next plane
    int interpolate;          // alpha > 0
    int down, up;             // level cells per pixel / pixels per level cell, along each in-plane axis
    int pixH, pixV;           // image size
    int startH, startV;       // first pixel of the crop, in image-level cells
    int gridID;
};
struct CutUnits {
    int plane;
    float rhoPhys;
    float res[LBM3_MAX_LEVELS], dt[LBM3_MAX_LEVELS];
};
template <class S> void launch_section_cut(const S *f, const Topology &t, const int *wallMap, const Codec &cd, bool flip,
                                           const CutLevel &d, float *img, hipStream_t s);
void launch_cut_clear(float *img, long long pixels, hipStream_t s);
template <class S> void launch_vtk_fields(const S *f, const Topology &t, const int *wallMap, const uint32_t *wallData,
                                          const Codec &cd, bool flip, const LevelGeom &g, float rhoPhys, float *out,
                                          hipStream_t s);
void launch_cut_finish(float *img, long long pixels, const CutUnits &u, hipStream_t s);

// the radial profile of every open rule, and none for the default record
struct OpenProfiles {
    lbm3_profile p[LBM3_MAX_RULES + 1];
};

// bulk cells (all but wall-adjacent and solid) and wall-adjacent cells; together one collide-stream step
template <class S> void launch_collide_stream(const StepArgs<S> &a, bool flip, hipStream_t s);
template <class S> void launch_open_cells(const StepArgs<S> &a, bool flip, hipStream_t s);
template <class S> void launch_rotor_cells(const StepArgs<S> &a, bool flip, hipStream_t s);
template <class S> void launch_collide_wall(const StepArgs<S> &a, bool flip, hipStream_t s);

// one time step of many single-level domains in four launches (bulk, wall,
// open boundary, periodic copy); starts holds four block-prefix arrays of n + 1
template <class S> struct MultiDesc {
    StepArgs<S> step;
    OpenArgs<S> open;             // at the parity after this step
    const int *perIndex, *perSource;
    int nPeriodic;
    int flip;                     // esotwist parity before this step
    int k17;                      // collision variant of the bulk cells
};
template <class S>
void launch_multi_step(const MultiDesc<S> *D, const int *starts, int n, const int (&blocks)[4], hipStream_t s);
template <class S> void launch_open(const OpenArgs<S> &a, bool flip, hipStream_t s);
template <class S> void launch_periodic(S *f, const Topology &t, int count, const int *index, const int *source, bool flip, hipStream_t s);
template <class S> void launch_init(S *f, const Topology &t, const InitDesc &d, const LevelGeom &g, const Codec &cd, bool flip,
                                    hipStream_t s);
template <class S> void launch_open_ref_init(const OpenArgs<S> &a, int X, int Y, int Z, bool flip, hipStream_t s);
template <class S> void launch_macros(const S *f, const Topology &t, typename StoreTraits<S>::compute *out, const Codec &cd,
                                      bool flip, hipStream_t s);
// FP64 sums over the counted cells of a level: the kinetic energy into *out,
// or dRho into out2[0] and the cell count into out2[1] (launch_mass), each
// divided by `divisor` and added to what the output holds when `accumulate`;
// partial holds one double per 256 cells (two for the mass); returns the blocks
template <class S> int launch_kinetic_energy(const S *f, const Topology &t, const int *wallMap, const uint32_t *wallData,
                                             double *partial, double *out, const Codec &cd, bool flip, hipStream_t s,
                                             double divisor = 1.0, bool accumulate = false);
template <class S> int launch_mass(const S *f, const Topology &t, const int *wallMap, const uint32_t *wallData,
                                   double *partial, double *out2, const Codec &cd, bool flip, hipStream_t s,
                                   double divisor = 1.0, bool accumulate = false);
void launch_wall_force(const double *gWall, int W, const uint32_t *wallData, const int *wallIndex, const Topology &t,
                       const LevelGeom &g, const double origin[3], int nBodies, double *partial, hipStream_t s);
template <class T> void launch_collide_batch(T *f, int n, const Relax<T> &R, hipStream_t s);
void launch_open_nb(const Topology &t, int count, const int *cell, int4 *out, hipStream_t s);
int launch_open_stats(const double *dRhoCum, const double *uNormalCum, const uint32_t *meta, int start, int count,
                      double invSpan, double *partial, hipStream_t s);
void launch_open_scale(const Topology &t, const LevelGeom &g, int count, const int *cell, const uint32_t *meta,
                       const OpenProfiles &P, float *scale, hipStream_t s);
void launch_fast_topo(const Topology &t, const int *wallMap, const int *openOf, const uint8_t *rotorCand, int2 *cellRow,
                      int4 *rows, hipStream_t s);
void launch_scatter_index(const int *cell, int count, int *openOf, hipStream_t s);
void launch_scatter_entries(const int *cell, const int *list, int count, int *openOf, hipStream_t s);
void launch_bulk_mask(const int2 *cellRow, int N, uint64_t *skip, int *tileRow, uint64_t *open, uint64_t *rotor,
                      hipStream_t s);
void launch_group_bits(const int *groups, int count, int G, uint64_t *skip, uint64_t *open, hipStream_t s);
void launch_mark_groups(const int *groups, int count, int G, int N, int2 *cellRow, hipStream_t s);
void launch_rotor_reach(const Topology &t, const LevelGeom &g, const RotorReachSet &R, uint8_t *cand, hipStream_t s);
// mode 0: listed cells, 1: signed cells (-c-1), 2: 2x2x2 stencils from their first cell, 3: cells next to a wall
void launch_open_conflicts(const int *cells, int count, int mode, const int *openOf, const int *wallMap,
                           const uint8_t *cand, const FastTopo &t, int *flag, hipStream_t s);
void launch_copy(const void *a, void *b, long long bytes, hipStream_t s);
void launch_relayout(const void *src, void *dst, int N, int elemBytes, bool toPlanes, hipStream_t s);

} // namespace lbm3
