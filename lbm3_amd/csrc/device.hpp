// Device-side views of one grid level and the per-step argument blocks.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include "lattice.hpp"
#include "cumulant.hpp"
#include "boundary.hpp"
#include "storage.hpp"
#include "../../include/lbm3mi_types.h"

namespace lbm3 {

// Compressed IJK/NBR (reference types.h:162-164): one shifter per cell plus six
// per-row ints.  shifter >= 0: the cell starts row `shifter`; < 0: offset back
// to the row head.  i+ is cell+1 (wrapping to 0), j+/k+/jk+ come from the row.
struct Topology {
    int N;
    const int *shifter;
    const int *ri, *rj, *rk, *rjp, *rkp, *rjkp;
};

__device__ __forceinline__ void decode(const Topology &t, int cell, int (&nb)[8], int *ijk)
{
    const int shift = t.shifter[cell];
    int row = shift, back = 0;
    if (shift < 0) { row = t.shifter[cell + shift]; back = -shift; }
    if (ijk) { ijk[0] = t.ri[row] + back; ijk[1] = t.rj[row]; ijk[2] = t.rk[row]; }
    const int N = t.N;
    nb[0] = cell;
    nb[2] = t.rjp[row] + back;
    nb[3] = t.rkp[row] + back;
    nb[6] = t.rjkp[row] + back;
    nb[1] = cell + 1 >= N ? 0 : cell + 1;
    nb[4] = nb[2] + 1 >= N ? 0 : nb[2] + 1;
    nb[5] = nb[3] + 1 >= N ? 0 : nb[3] + 1;
    nb[7] = nb[6] + 1 >= N ? 0 : nb[6] + 1;
}

// Hot-loop form of the compressed topology, derived once per level: one 8-byte
// record per cell {wallMap, row} and one 16-byte record per row {first cell of
// the row, first cell of the j+ row, of the k+ row, of the jk+ row}.  The
// neighbour cells of a cell are then two dependent loads away instead of four.
// The bulk kernel reads neither per-cell record: a bit per cell says whether
// another kernel (or none) updates it, and the row of every 256th cell, with
// the rows' first cells, gives the row of the others.
struct FastTopo {
    int N;
    const int2 *cellRow;   // [N]
    const int4 *rows;      // [nRows]
    const uint64_t *skip;  // [N / 64 rounded up to 4] bit c % 64 of word c / 64: not a bulk cell; null: read cellRow
    const int *tileRow;    // [N / 256 rounded up] row of cell 256 t
    int nRows;
    const uint64_t *openBits;  // as skip: open cells the bulk kernel completes and steps itself; null: none
    const uint64_t *rotorBits; // as skip: cells the rotors can reach, stepped by the bulk kernel; null: none
};

__device__ __forceinline__ void decode_fast(const FastTopo &t, int cell, int row, int (&nb)[8])
{
    const int4 r = t.rows[row];
    const int back = cell - r.x;
    const int N = t.N;
    nb[0] = cell;
    nb[2] = r.y + back;
    nb[3] = r.z + back;
    nb[6] = r.w + back;
    nb[1] = cell + 1 >= N ? 0 : cell + 1;
    nb[4] = nb[2] + 1 >= N ? 0 : nb[2] + 1;
    nb[5] = nb[3] + 1 >= N ? 0 : nb[3] + 1;
    nb[7] = nb[6] + 1 >= N ? 0 : nb[6] + 1;
}

// getLocalBC resolved per (level, wallID): rules depend only on those two keys,
// except the rotating-wall velocity which needs the cell position.
struct LocalBC {
    float gx, gy, gz, limiter, nuMult;
    int useIBB;
    int velMode;            // 0: wall moves with the fluid's own velocity, 1: fixed, 2: rotation,
                            // 3: per wall-adjacent cell (the level's wallVel array)
    float ux, uy, uz;
    float omega[3], origin[3];
    Relax<float> rf;        // relaxation constants for the level's viscosity
    Relax<double> rd;
};

template <class T> __host__ __device__ __forceinline__ const Relax<T> &relax_of(const LocalBC &b);
template <> __host__ __device__ __forceinline__ const Relax<float> &relax_of<float>(const LocalBC &b) { return b.rf; }
template <> __host__ __device__ __forceinline__ const Relax<double> &relax_of<double>(const LocalBC &b) { return b.rd; }

inline void set_relax(LocalBC &b, double nu)
{
    b.rf = make_relax<float>(float(nu), b.nuMult, b.limiter, b.gx, b.gy, b.gz);
    b.rd = make_relax<double>(nu, double(b.nuMult), double(b.limiter), double(b.gx), double(b.gy), double(b.gz));
}

struct LevelGeom {          // what the hooks need about the level
    float res, ox, oy, oz, dtPhys;
    int cellCountX, cellCountY, cellCountZ;
};

// Open-boundary cells updated by their own kernel (k_open_groups): each step
// it completes the populations the previous step left unknown (the deferred
// boundary condition), then collides and streams the cell.  The kernel takes
// whole aligned groups of LBM3_OPEN_GROUP cells around the open cells (the
// other cells of a group are plain fluid cells, record -5), so that a face
// cell at the end of a row loads and stores full lines together with its row
// neighbours.  count == 0: the level's open cells go through the bulk kernel
// and the separate pass.
template <class T> struct FusedOpen {
    const int *groups;            // [count] groups, sorted by the boundary condition of their first open cell
    int count;
    int pending;                  // the previous step's boundary condition is still to apply
    const uint32_t *meta;         // per open-list entry: normal index, record, track flag
    const OpenBCRecord *records;
    T *dRhoRef, *uNormalRef;
    double *dRhoCum, *uNormalCum;
    int iter;                     // iteration the pending boundary condition belongs to
    const float *scale;           // per open-list entry: velocity profile factor, or null (none)
    int groupCells;               // cells per group (the width the groups were built with)
};

// Per-step values of a step replayed from a captured HIP graph, whose kernel
// arguments are fixed at capture: the iteration the open-boundary condition
// belongs to, whether a deferred condition is pending, and each rotor's cosine
// and sine.  A null pointer in the argument block: the arguments hold them.
struct StepParam {
    int iter;                     // iteration of a condition the open-cell kernel applies (pending from the last step)
    int pending;
    int iterNow;                  // iteration of this step's separate boundary pass
    int pad;
    double cs[2 * LBM3_MAX_ROTORS];
};

// One rotor of a level (RotorStruct, types.h:206-219) as the kernels read it;
// the angle enters as the step's cosine and sine (StepArgs::rotorCS).
struct RotorDev {
    float res;                    // interpolation cell: half the level spacing
    float xMin, yMin, zMin, xMax, yMax, zMax;
    int cx, cy, cz;               // interpolation cells (multiples of 4)
    float ox, oy, oz;             // a point of the rotation axis
    int axis;                     // LBM3_AXIS_*
    float radiansPerSecond;
    int nonZero;
    const int *rotorMap;          // [cx/4 * cy/4 * cz/4] compact block index or -1
    const uint32_t *interp;       // [nonZero * 64] packed corner counts
    double *g;                    // [3][nonZero * 27] forcing per force cell (FP64)
};

// what a rotor can reach: its interpolation bounds, axis and a point of it
struct RotorReach {
    double lo[3], hi[3], o[3];
    int axis;
};
struct RotorReachSet {
    RotorReach r[LBM3_MAX_ROTORS];
    int n;
};

template <class S> struct StepArgs {
    S *f;
    Topology topo;
    FastTopo fast;
    const int *wallMap;
    const int *wallIndex;         // [W] cell of each wall-adjacent entry
    const uint32_t *wallData;
    const uint32_t *linkLength;   // [7][W]
    int W;
    double *gWall;                // [3][W] or null (no force tracking)
    float nu;
    LocalBC bulk;                 // cells that are not next to a wall
    const LocalBC *wallBC;        // [32] by wallID
    const float *wallVel;         // [3][W] wall velocity per wall-adjacent cell (velMode 3) or null
    LevelGeom geom;
    Codec cd;                     // 16-bit population encoding
    FusedOpen<typename StoreTraits<S>::compute> open;
    // rotors of the level: every kernel that collides a cell the rotors can
    // reach applies their forcing (nRotors == 0: none)
    const RotorDev *rotors;
    int nRotors;
    int trackRotor;               // TRACK_ROTOR_FORCE
    const int *rotorCells;        // [rotorCellCount] cells with their own kernel (record -4)
    int rotorCellCount;
    double rotorCS[2 * LBM3_MAX_ROTORS];   // cos, sin of each rotor's angle this step, in the compute type
    const StepParam *param;       // this step's values in device memory (graph replay) or null
    int xcdRemap;                 // bulk block order (xcd_block)
    int bulkLds;                  // dynamic LDS bytes per bulk block (host side: caps resident bulk blocks per CU)
    int laneRot;                  // host side: the bulk kernel with open cells inside takes its block's lanes rotated
    int wallBlocks;               // blocks of wall-adjacent cells ahead of the bulk blocks in one launch (0: none)
};

// Open boundary cells of one level, all BCIDs concatenated (BCID order, then cell order)
template <class S> struct OpenArgs {
    S *f;
    Topology topo;
    int count;
    const int *cell;
    const int4 *nb;               // [2*count] the 8 neighbour cells of each open cell (setup-time decode)
    const uint32_t *meta;         // open_meta: normal index, track flow, record
    const OpenBCRecord *records;
    typename StoreTraits<S>::compute *dRhoRef, *uNormalRef;  // reference state, compute precision
    double *dRhoCum, *uNormalCum; // flow accumulators or null
    int iterationsFinished;
    Codec cd;
    const float *scale;           // per entry: velocity profile factor of its rule, or null (none)
    const StepParam *param;       // this step's iteration in device memory (graph replay) or null
    const int *list;              // the entries to complete ([count]), or null: entries 0..count-1
};

// Blocks are dispatched round-robin over the 8 XCDs.  With bit 0 of the block
// order each XCD gets a contiguous range of blocks (its own L2 then sees whole
// neighbouring rows); with the tiled population layout the plain order, in
// which the whole chip works on one narrow window of tiles at a time, measured
// 4% faster on the 256x128x128 sphere bulk kernel (162 vs 169 us).  Bit 1
// sweeps the tiles backwards on odd esotwist parity, so that each step starts
// on the lines the previous step wrote last, still in the 256 MB last-level
// cache: the FP32 sphere bulk kernel (452 MB of populations) 186 -> 158 us.
// Default: plain order, reversed every other step.
#ifndef LBM3_XCD_REMAP
#define LBM3_XCD_REMAP 2
#endif

// Dynamic LDS per bulk block, unused by the kernel: a cap on resident bulk
// blocks per CU (160 KiB / 40 KiB = 4, i.e. 16 of the 32 wave slots a CU has
// at this kernel's register use).  Measured on the 256x128x128 sphere: the
// bulk kernel is 1.5% faster than at full occupancy, and the free slots let
// the wall-adjacent kernel on the second stream run beside it instead of
// after it (the whole step 7% faster).  The default of Sim::Tune::bulkLds.
#ifndef LBM3_BULK_LDS
#define LBM3_BULK_LDS 40960
#endif

// The tile a bulk block takes.  remap bit 0: blocks dealt so that each XCD
// sweeps one contiguous part (else round-robin over the XCDs); bit 1: on odd
// esotwist parity the tiles in reverse order, so that a step starts where the
// previous one ended, on the lines the last-level cache still holds
__device__ __forceinline__ int xcd_block(int b, int n, int remap, bool flip)
{
    int t = b;
    if (remap & 1) {
        const int per = n / 8;
        if (b < per * 8) t = (b % 8) * per + b / 8;
    }
    if ((remap & 2) && flip) t = n - 1 - t;
    return t;
}

} // namespace lbm3
