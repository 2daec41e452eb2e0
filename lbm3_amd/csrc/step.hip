// Per-step kernels of one grid level for gfx950.
//
//  k_collide_stream  fused read -> cumulant collision -> esotwist write of the
//                    cells that are neither wall-adjacent nor open boundary
//                    cells (reference updateGrid.h:51-178)
//  k_collide_wall    the same with interpolated bounceback and FP64 momentum-
//                    exchange force on wall-adjacent cells
//                    (interpolatedBouncebackFunctions.h:43-157)
//  k_open_groups     the same for aligned groups of cells around the open-
//                    boundary cells, which first complete the populations
//                    the previous step left unknown
//  k_open            moment-based open boundaries as a pass of their own after
//                    streaming, all BCIDs in one launch (updateGrid.h:183-355)
//  k_periodic        periodic ghost copy (updateGrid.h:357-382)
//  k_init, k_open_ref_init   initial condition (applyInitialCondition.h:8-84)
//  k_macros, k_cell_sums, k_wall_force        observables, FP64 sums
//
// Populations are stored in tiles of LBM3_TILE cells, each tile holding its 27
// slot runs one after another (storage.hpp pop_index); each thread owns one
// cell, keeps its 27 populations in registers and writes back into the same 27
// slots it read (esotwist), so the kernel is race free without any
// synchronisation.  The esotwist parity is a template parameter, which makes
// every slot run's offset within a tile a compile-time constant.
#include "step.hpp"
#include "hooks.hpp"
#include <type_traits>

namespace lbm3 {

template <class S, bool FLIP>
__device__ __forceinline__ void load_pre(const S *f, int N, const int (&nb)[8], typename StoreTraits<S>::compute (&v)[27],
                                         const Codec &cd)
{
#pragma unroll
    for (int q = 0; q < 27; ++q) v[q] = StoreTraits<S>::load(f + pop_index(pre_plane(q, FLIP), nb[pre_nb(q)]), q, cd);
}

template <class S, bool FLIP>
__device__ __forceinline__ void store_pre(S *f, int N, const int (&nb)[8], const typename StoreTraits<S>::compute (&v)[27],
                                          const Codec &cd)
{
#pragma unroll
    for (int q = 0; q < 27; ++q) StoreTraits<S>::store(f + pop_index(pre_plane(q, FLIP), nb[pre_nb(q)]), v[q], q, cd);
}

// interpolated bounceback (Zhao & Yong 2017, single node, l = gamma) on the
// post-collision populations of a wall-adjacent cell; opposite directions are
// handled as pairs so both use the unmodified post-collision values.  `pre`
// holds the cell's pre-collision populations (the values its slots held before
// this step), `codes` the cell's seven words of 8-bit link lengths in link order.
template <class T>
__device__ __forceinline__ void bounce_back(T (&f)[27], const T (&pre)[27], uint32_t wd, bool useIBB,
                                            const uint32_t (&codes)[7], const T (&wu)[3], double (&g3)[3])
{
    T dRho, ux_, uy_, uz_;
    moments(f, dRho, ux_, uy_, uz_);
    int linkIndex = 0;
    auto gamma_next = [&]() -> T {
        if (!useIBB) return T(0.5);
        const uint32_t c8 = (codes[linkIndex >> 2] >> (8 * (linkIndex & 3))) & 255u;
        ++linkIndex;
        T g = T(c8) / T(254);
        return g < T(0.01f) ? T(0.01f) : (g > T(1) ? T(1) : g);
    };
#pragma unroll
    for (int p = 1; p < 27; p += 2) {
        const int da = p, db = p + 1;   // opposite pair
        const bool la = (wd >> da) & 1u, lb = (wd >> db) & 1u;
        if (!la && !lb) continue;
        const T fa = f[da], fb = f[db];
        T ra = fa, rb = fb;
        if (la) {
            const T g = gamma_next();
            const T inv1 = T(1) / (T(1) + g);
            const T eu = T(cx(db)) * T(wu[0]) + T(cy(db)) * T(wu[1]) + T(cz(db)) * T(wu[2]);
            const T wall = T(6) * weight<T>(db) * (T(1) + dRho) * eu;
            ra = ((T(1) - g) * inv1) * pre[da] + (g * inv1) * fb + (g * inv1) * fa + inv1 * wall;
            g3[0] += double(-fa * (T(cx(da)) - T(wu[0])) + ra * (T(cx(db)) - T(wu[0])));
            g3[1] += double(-fa * (T(cy(da)) - T(wu[1])) + ra * (T(cy(db)) - T(wu[1])));
            g3[2] += double(-fa * (T(cz(da)) - T(wu[2])) + ra * (T(cz(db)) - T(wu[2])));
        }
        if (lb) {
            const T g = gamma_next();
            const T inv1 = T(1) / (T(1) + g);
            const T eu = T(cx(da)) * T(wu[0]) + T(cy(da)) * T(wu[1]) + T(cz(da)) * T(wu[2]);
            const T wall = T(6) * weight<T>(da) * (T(1) + dRho) * eu;
            rb = ((T(1) - g) * inv1) * pre[db] + (g * inv1) * fa + (g * inv1) * fb + inv1 * wall;
            g3[0] += double(-fb * (T(cx(db)) - T(wu[0])) + rb * (T(cx(da)) - T(wu[0])));
            g3[1] += double(-fb * (T(cy(db)) - T(wu[1])) + rb * (T(cy(da)) - T(wu[1])));
            g3[2] += double(-fb * (T(cz(db)) - T(wu[2])) + rb * (T(cz(da)) - T(wu[2])));
        }
        f[da] = ra;
        f[db] = rb;
    }
}

// Moment-based open boundary of one cell with normal NI (updateGrid.h:183-355):
// f holds the known populations (the unknown ones are overwritten), the
// unknowns are solved for, and the reference state (EMA) and flow accumulators
// of open-list entry x are updated.  `iter` is the iteration the boundary
// condition belongs to.
template <int NI, class T>
__device__ __forceinline__ void open_fill(T (&f)[27], int x, uint32_t meta, const OpenBCRecord *records, T *dRhoRef,
                                          T *uNormalRef, double *dRhoCum, double *uNormalCum, int iter, const float *scale)
{
    constexpr int nx = LBM3_BT_NORMAL[NI][0], ny = LBM3_BT_NORMAL[NI][1], nz = LBM3_BT_NORMAL[NI][2];
    // open_cell leaves the known populations as full (unshifted) values: keep
    // the stored ones, so the cell collides exactly what the separate pass would
    // leave in memory (it writes back only the unknown ones)
    T known[27];
#pragma unroll
    for (int q = 0; q < 27; ++q) {
        known[q] = f[q];
        if (unknown_dir(nx, ny, nz, q)) f[q] = T(0);
    }
    OpenBCRecord R = records[open_meta_record(meta)];
    if (scale) {                   // the rule's velocity profile at this cell
        const float sc = scale[x];
        R.ux *= sc; R.uy *= sc; R.uz *= sc;
    }
    const T dRef = dRhoRef[x], uRef = uNormalRef[x];
    T dRho, un;
    open_cell<NI>(f, R, dRef, uRef, dRho, un);
#pragma unroll
    for (int q = 0; q < 27; ++q)
        if (!unknown_dir(nx, ny, nz, q)) f[q] = known[q];
    if (R.impedance) {
        // the impedance scheme keeps the completed state for the next step
        dRhoRef[x] = dRho;
        uNormalRef[x] = un;
    } else {
        // exponential moving average of the reference state
        const T beta = T(R.beta);
        T alpha = beta / (T(iter) + beta + T(10));
        int every = 1;
        while (alpha < T(1.e-5f)) { alpha *= T(3); every *= 3; }
        if (iter % every == 0) {
            dRhoRef[x] = fma(alpha, dRho - dRef, dRef);
            uNormalRef[x] = fma(alpha, un - uRef, uRef);
        }
    }
    if (dRhoCum && (meta & kOpenTrackBit)) {
        dRhoCum[x] += double(dRho);
        uNormalCum[x] += double(un);
    }
}

template <class T>
__device__ __forceinline__ void open_dispatch(T (&f)[27], int x, uint32_t meta, const OpenBCRecord *records, T *dRhoRef,
                                              T *uNormalRef, double *dRhoCum, double *uNormalCum, int iter,
                                              const float *scale)
{
    switch (meta & 31u) {
#define OPEN_CASE(n) case n: open_fill<n, T>(f, x, meta, records, dRhoRef, uNormalRef, dRhoCum, uNormalCum, iter, scale); break;
        OPEN_CASE(0) OPEN_CASE(1) OPEN_CASE(2) OPEN_CASE(3) OPEN_CASE(4) OPEN_CASE(5) OPEN_CASE(6)
        OPEN_CASE(7) OPEN_CASE(8) OPEN_CASE(9) OPEN_CASE(10) OPEN_CASE(11) OPEN_CASE(12) OPEN_CASE(13)
        OPEN_CASE(14) OPEN_CASE(15) OPEN_CASE(16) OPEN_CASE(17) OPEN_CASE(18) OPEN_CASE(19) OPEN_CASE(20)
        OPEN_CASE(21) OPEN_CASE(22) OPEN_CASE(23) OPEN_CASE(24) OPEN_CASE(25)
#undef OPEN_CASE
    default: break;
    }
}

// neighbour addressing: a wave-uniform base (array + plane * TILE) plus a byte
// offset per lane, 32-bit where the level's array is below 4 GiB
template <class S, class Off> __device__ __forceinline__ S *at(S *base, Off off)
{
    return reinterpret_cast<S *>(reinterpret_cast<char *>(base) + off);
}

// ---------------------------------------------------------------------------
// rotors (rotorFunctions.h): immersed forcing toward rigid rotation inside a
// voxelized rotor, evaluated in the rotor's frame
// ---------------------------------------------------------------------------

// getRotorFraction (rotorFunctions.h:155-208, projectXYZIntoRotorFrame :8-35,
// interpolateRotorCube :135-153): interpolation cell of the point in the rotor
// frame and its solid fraction; false outside every marked block
template <class T>
__device__ __forceinline__ bool rotor_locate(const RotorDev &R, T c, T s, T x, T y, T z, int &rm, int (&ijk)[3], T (&w)[3],
                                             T &fr)
{
    T xr = x, yr = y, zr = z;
    const T dx = x - T(R.ox), dy = y - T(R.oy), dz = z - T(R.oz);
    if (R.axis == LBM3_AXIS_X) { yr = T(R.oy) + c * dy + s * dz; zr = T(R.oz) - s * dy + c * dz; }
    else if (R.axis == LBM3_AXIS_Y) { xr = T(R.ox) + c * dx - s * dz; zr = T(R.oz) + s * dx + c * dz; }
    else if (R.axis == LBM3_AXIS_Z) { xr = T(R.ox) + c * dx + s * dy; yr = T(R.oy) - s * dx + c * dy; }
    if (xr < T(R.xMin) || xr >= T(R.xMax) || yr < T(R.yMin) || yr >= T(R.yMax) || zr < T(R.zMin) || zr >= T(R.zMax))
        return false;
    const T xRel = xr - T(R.xMin), yRel = yr - T(R.yMin), zRel = zr - T(R.zMin);
    const T res = T(R.res);
    const int ii = int(xRel / res), ji = int(yRel / res), ki = int(zRel / res);
    if (ii >= R.cx || ji >= R.cy || ki >= R.cz) return false;
    const int bx = R.cx / 4, by = R.cy / 4;
    rm = R.rotorMap[(ki / 4) * bx * by + (ji / 4) * bx + ii / 4];
    if (rm < 0) return false;
    const uint32_t packed = R.interp[rm * 64 + (ki % 4) * 16 + (ji % 4) * 4 + (ii % 4)];
    w[0] = xRel / res - T(ii);
    w[1] = yRel / res - T(ji);
    w[2] = zRel / res - T(ki);
    T v[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) v[q] = T((packed >> (4u * q)) & 0xFu);
    const T v00 = v[0] + w[0] * (v[1] - v[0]), v10 = v[2] + w[0] * (v[3] - v[2]);
    const T v01 = v[4] + w[0] * (v[5] - v[4]), v11 = v[6] + w[0] * (v[7] - v[6]);
    const T v0 = v00 + w[1] * (v10 - v00), v1 = v01 + w[1] * (v11 - v01);
    fr = (v0 + w[2] * (v1 - v0)) * T(0.125f);
    ijk[0] = ii; ijk[1] = ji; ijk[2] = ki;
    return true;
}

// the rotor loop of the cell update (updateGrid.h:98-129) with processRotor
// (rotorFunctions.h:210-304): each rotor adds its smoothstep share of the
// forcing rho (u_target - u_pre) to the cell's force, the cumulative share is
// capped at 1 and scales the collision limiter down; the forcing is tracked,
// turned into the rotor frame, per force cell in FP64
template <class S, class T>
__device__ __forceinline__ void rotors_apply_at(const StepArgs<S> &a, const int (&ijk)[3], T dRho, T ux, T uy, T uz,
                                                Relax<T> &R, bool track)
{
    const T x = T(float(ijk[0]) * a.geom.res + a.geom.ox), y = T(float(ijk[1]) * a.geom.res + a.geom.oy),
            z = T(float(ijk[2]) * a.geom.res + a.geom.oz);
    const T rho = dRho + T(1), rhoInv = T(1) / rho;
    const T uxPre = (ux * rho + R.gx) * rhoInv, uyPre = (uy * rho + R.gy) * rhoInv, uzPre = (uz * rho + R.gz) * rhoInv;
    T cum = T(0);
    for (int r = 0; r < a.nRotors; ++r) {
        const RotorDev &Rt = a.rotors[r];
        const double *cs = a.param ? a.param->cs : a.rotorCS;
        const T c = T(cs[2 * r]), s = T(cs[2 * r + 1]);
        int rm, cI[3];
        T w[3], fr;
        if (!rotor_locate<T>(Rt, c, s, x, y, z, rm, cI, w, fr) || fr == T(0)) continue;
        const T dx = x - T(Rt.ox), dy = y - T(Rt.oy), dz = z - T(Rt.oz);
        const T scale = T(Rt.radiansPerSecond) * T(a.geom.dtPhys) / T(a.geom.res);
        T ut0 = T(0), ut1 = T(0), ut2 = T(0);
        if (Rt.axis == LBM3_AXIS_X) { ut1 = -scale * dz; ut2 = scale * dy; }
        else if (Rt.axis == LBM3_AXIS_Y) { ut0 = scale * dz; ut2 = -scale * dx; }
        else if (Rt.axis == LBM3_AXIS_Z) { ut0 = -scale * dy; ut1 = scale * dx; }
        T gx = rho * (ut0 - uxPre), gy = rho * (ut1 - uyPre), gz = rho * (ut2 - uzPre);
        T ff = fr * fr * (T(3) - T(2) * fr);
        if (cum + ff > T(1)) { ff = T(1) - cum; cum = T(1); }
        else cum += ff;
        gx *= ff; gy *= ff; gz *= ff;
        R.gx += gx; R.gy += gy; R.gz += gz;
        if (track) {
            if (Rt.axis == LBM3_AXIS_X) { const T t = gy; gy = c * t + s * gz; gz = -s * t + c * gz; }
            else if (Rt.axis == LBM3_AXIS_Y) { const T t = gx; gx = c * t - s * gz; gz = s * t + c * gz; }
            else if (Rt.axis == LBM3_AXIS_Z) { const T t = gx; gx = c * t + s * gy; gy = -s * t + c * gy; }
            const int iF = min(2, max(0, int(T(0.75f) * (T(cI[0] % 4) + w[0]))));
            const int jF = min(2, max(0, int(T(0.75f) * (T(cI[1] % 4) + w[1]))));
            const int kF = min(2, max(0, int(T(0.75f) * (T(cI[2] % 4) + w[2]))));
            const size_t fi = (size_t)rm * 27 + (size_t)(kF * 9 + jF * 3 + iF), nf = (size_t)Rt.nonZero * 27;
            atomicAdd(Rt.g + fi, double(gx));
            atomicAdd(Rt.g + nf + fi, double(gy));
            atomicAdd(Rt.g + 2 * nf + fi, double(gz));
        }
        if (cum >= T(1)) break;
    }
    if (cum > T(0)) {
        R.lambda *= T(1) - (cum < T(0) ? T(0) : (cum > T(1) ? T(1) : cum));
        R.k17 = R.k17 && R.lambda > T(0);
    }
}

template <class S, class T>
__device__ __forceinline__ void rotors_apply(const StepArgs<S> &a, int cell, T dRho, T ux, T uy, T uz, Relax<T> &R,
                                             bool track)
{
    int nbx[8], ijk[3];
    decode(a.topo, cell, nbx, ijk);
    rotors_apply_at<S, T>(a, ijk, dRho, ux, uy, uz, R, track);
}

// grid coordinates of a cell from its row (FastTopo and Topology number rows alike)
__device__ __forceinline__ void cell_ijk(const Topology &t, const FastTopo &ft, int cell, int row, int (&ijk)[3])
{
    ijk[0] = t.ri[row] + (cell - ft.rows[row].x);
    ijk[1] = t.rj[row];
    ijk[2] = t.rk[row];
}

// the collision of a bulk cell; ROTOR: the level's rotor cells are the bulk
// kernel's too (FastTopo::rotorBits): a wave holding one takes the collision
// with run-time K15/K17 choice, and its rotor lanes add the rotors' forcing
// and scaled limiter first, as k_rotor_cells does
template <class S, int BULK, bool ROTOR>
__device__ __forceinline__ void bulk_collide(const StepArgs<S> &a, int cell, int row,
                                             typename StoreTraits<S>::compute (&f)[27])
{
    using T = typename StoreTraits<S>::compute;
    if constexpr (ROTOR) {
        const uint64_t m = a.fast.rotorBits[__builtin_amdgcn_readfirstlane(cell >> 6)];
        if (m) {
            Relax<T> R = relax_of<T>(a.bulk);
            if ((m >> (cell & 63)) & 1ull) {
                T dRho, ux, uy, uz;
                moments(f, dRho, ux, uy, uz);
                int ijk[3];
                cell_ijk(a.topo, a.fast, cell, row, ijk);
                rotors_apply_at<S, T>(a, ijk, dRho, ux, uy, uz, R, a.trackRotor && a.wallMap[cell] != -2);
            }
            collide<T, 0>(f, R);
            return;
        }
    }
    collide<T, BULK>(f, relax_of<T>(a.bulk));
}

// Cells are independent under esotwist streaming (each one reads and writes
// only its own 27 slots), so the level is updated by kernels over disjoint cell
// sets: one over all cells that skips the special ones, one over the short list
// of wall-adjacent cells and one over the open-boundary cells.  The bulk kernel
// then carries neither bounceback nor boundary-condition code, which keeps its
// register budget and instruction stream to the collision itself, and no wave
// of it diverges on the odd boundary cell at the start or end of a row.
//
// BULK: collision variant of the cells that are not next to a wall, fixed for
// the level on the host (1: K17, 2: K15).  Topology record >= 0: wall-adjacent,
// -3: solid, -4: a cell the rotors can reach, -5: a fluid cell in a group of
// open cells, <= -8: open cell (each with its own kernel).
// 16-bit storage in the bulk kernel: 0 global loads and stores, 1 raw buffer
// loads and stores, 2 buffer loads and global stores.  Measured on the
// 256x128x128 sphere (FP16 bulk kernel with the open cells inside, after the
// stores' plain conversion): 103.2 us, 98.0 us, 99.8 us (with the fused
// conversion the order was reversed: 93.4, 97.6, 94.4 us, open cells in groups).
#ifndef LBM3_BUF16
#define LBM3_BUF16 1
#endif

// OPEN: the level's open cells are in the bulk kernel too (FastTopo::openBits):
// such a cell first completes the populations the previous step left unknown,
// as k_open_groups does, then collides and streams like its neighbours
// the bit of `cell` in a mask of one bit per cell, 64 to a word: a wave of
// consecutive 64-aligned cells reads one word (a scalar load); with its
// block's lanes rotated (ROT) a wave may span two words, and each lane reads
// its own
template <bool ROT>
__device__ __forceinline__ bool mask_bit(const uint64_t *m, int cell)
{
    const uint64_t w = ROT ? m[cell >> 6] : m[__builtin_amdgcn_readfirstlane(cell >> 6)];
    return (w >> (cell & 63)) & 1ull;
}

template <class S, bool OPEN, bool ROT = false>
__device__ __forceinline__ void open_in_bulk(const StepArgs<S> &a, int cell, typename StoreTraits<S>::compute (&f)[27])
{
    if constexpr (OPEN) {
        using T = typename StoreTraits<S>::compute;
        if (!mask_bit<ROT>(a.fast.openBits, cell)) return;
        if (!(a.param ? a.param->pending : a.open.pending)) return;
        const int x = -(a.fast.cellRow[cell].x + 8);
        open_dispatch<T>(f, x, a.open.meta[x], a.open.records, a.open.dRhoRef, a.open.uNormalRef, a.open.dRhoCum,
                         a.open.uNormalCum, a.param ? a.param->iter : a.open.iter, a.open.scale);
    }
}

template <class S, bool FLIP, int BULK, class Off, bool OPEN = false, bool ROTOR = false, bool ROT = false>
__device__ __forceinline__ void bulk_cell(const StepArgs<S> &a, int cell)
{
    using T = typename StoreTraits<S>::compute;
    int row;
    if (a.fast.skip) {
        // a wave's 64 cells are 64-aligned and a block's 256: one mask word per
        // wave and one row per block, both scalar loads
        if (mask_bit<ROT>(a.fast.skip, cell)) return;   // handled by another kernel, or solid
        int r = a.fast.tileRow[__builtin_amdgcn_readfirstlane(cell >> 8)];
        while (r + 1 < a.fast.nRows && a.fast.rows[r + 1].x <= cell) ++r;
        row = r;
    } else {
        const long long rec = __builtin_nontemporal_load(reinterpret_cast<const long long *>(a.fast.cellRow) + cell);
        const int wm = int(rec);
        row = int(rec >> 32);
        if (wm >= 0 || wm <= -3) return;   // handled by another kernel, or solid
    }
    int nb[8];
    decode_fast(a.fast, cell, row, nb);
    // byte offset of each neighbour's slot run in its tile; the plane adds a
    // wave-uniform plane * TILE * sizeof(S)
    Off off[8];
#pragma unroll
    for (int k = 0; k < 8; ++k)
        off[k] = (Off(uint32_t(nb[k]) / LBM3_TILE) * Off(27 * LBM3_TILE) + Off(uint32_t(nb[k]) % LBM3_TILE)) * Off(sizeof(S));
    T f[27];
    if constexpr (sizeof(Off) == 4 && (sizeof(S) >= 4 || LBM3_BUF16 > 0)) {
        // raw buffer accesses: the plane offset is a scalar operand
        const __amdgpu_buffer_rsrc_t r = pop_rsrc(a.f, 27LL * padded_cells(a.fast.N) * (long long)sizeof(S));
#pragma unroll
        for (int q = 0; q < 27; ++q)
            f[q] = StoreTraits<S>::bload(r, off[pre_nb(q)], pre_plane(q, FLIP) * LBM3_TILE * int(sizeof(S)), q, a.cd);
        open_in_bulk<S, OPEN, ROT>(a, cell, f);
        bulk_collide<S, BULK, ROTOR>(a, cell, row, f);
        if constexpr (sizeof(S) >= 4 || LBM3_BUF16 == 1) {
#pragma unroll
            for (int q = 0; q < 27; ++q)
                StoreTraits<S>::bstore(r, off[post_nb(q)], post_plane(q, FLIP) * LBM3_TILE * int(sizeof(S)), f[q], q, a.cd);
        } else {
#pragma unroll
            for (int q = 0; q < 27; ++q)
                StoreTraits<S>::store(at(a.f + post_plane(q, FLIP) * LBM3_TILE, off[post_nb(q)]), f[q], q, a.cd);
        }
    } else {
#pragma unroll
        for (int q = 0; q < 27; ++q)
            f[q] = StoreTraits<S>::load(at(a.f + pre_plane(q, FLIP) * LBM3_TILE, off[pre_nb(q)]), q, a.cd);
        open_in_bulk<S, OPEN, ROT>(a, cell, f);
        bulk_collide<S, BULK, ROTOR>(a, cell, row, f);
#pragma unroll
        for (int q = 0; q < 27; ++q)
            StoreTraits<S>::store(at(a.f + post_plane(q, FLIP) * LBM3_TILE, off[post_nb(q)]), f[q], q, a.cd);
    }
}

// WIDE: the level's population array exceeds 4 GiB (64-bit lane offsets);
// OPEN: the open cells are stepped here too; ROTOR: the rotor cells (neither
// with WIDE).  ROT: lane t of a block takes cell (t + 32) mod 256 of its tile,
// so that where rows are whole tiles the last 32 and the first 32 cells of a
// row -- both x-face cells -- share one wave, which then runs the open
// condition once for the two instead of two waves once each; every access of
// that wave still fills whole 128-byte lines
template <class S, bool FLIP, int BULK, bool WIDE, bool OPEN = false, bool ROTOR = false, bool ROT = false>
__global__ __launch_bounds__(256) void k_collide_stream(StepArgs<S> a)
{
    const int lane = ROT ? int((threadIdx.x + 32) & 255) : int(threadIdx.x);
    const int cell = xcd_block(blockIdx.x, gridDim.x, a.xcdRemap, FLIP) * 256 + lane;
    if (cell >= a.fast.N) return;
    if (WIDE) bulk_cell<S, FLIP, BULK, uint64_t>(a, cell);
    else bulk_cell<S, FLIP, BULK, uint32_t, OPEN, ROTOR, ROT>(a, cell);
}

// Open-boundary cells (never wall-adjacent when this kernel is used): load,
// complete the populations the previous step left unknown -- the moment-based
// boundary condition of updateGrid.h:183-355, deferred from the end of that
// step to this load -- then collide and stream like a bulk cell.  The kernel
// runs over aligned groups of G cells that hold open cells and updates every
// fluid cell of the group (record -5 or <= -8); a wave covers 64 / G groups,
// sorted so that they share one boundary condition and normal, so the
// boundary-condition code runs once per wave for all its open lanes.
template <class S, bool FLIP, int BULK, int G>
__global__ __launch_bounds__(256) void k_open_groups(StepArgs<S> a)
{
    using T = typename StoreTraits<S>::compute;
    const int gi = blockIdx.x * (256 / G) + threadIdx.x / G;
    if (gi >= a.open.count) return;
    const int cell = a.open.groups[gi] * G + int(threadIdx.x % G);
    if (cell >= a.fast.N) return;
    const long long rec = __builtin_nontemporal_load(reinterpret_cast<const long long *>(a.fast.cellRow) + cell);
    const int code = int(rec), row = int(rec >> 32);
    if (code != -5 && code > -8) return;   // wall-adjacent, solid or rotor cell: other kernels
    int nb[8];
    decode_fast(a.fast, cell, row, nb);
    uint32_t off[8];
#pragma unroll
    for (int k = 0; k < 8; ++k)
        off[k] = ((uint32_t(nb[k]) / LBM3_TILE) * uint32_t(27 * LBM3_TILE) + uint32_t(nb[k]) % LBM3_TILE) * uint32_t(sizeof(S));
    const __amdgpu_buffer_rsrc_t r = pop_rsrc(a.f, 27LL * padded_cells(a.fast.N) * (long long)sizeof(S));
    T f[27];
#pragma unroll
    for (int q = 0; q < 27; ++q)
        f[q] = StoreTraits<S>::bload(r, off[pre_nb(q)], pre_plane(q, FLIP) * LBM3_TILE * int(sizeof(S)), q, a.cd);
    if (code <= -8 && (a.param ? a.param->pending : a.open.pending)) {
        const int x = -(code + 8);
        open_dispatch<T>(f, x, a.open.meta[x], a.open.records, a.open.dRhoRef, a.open.uNormalRef, a.open.dRhoCum,
                         a.open.uNormalCum, a.param ? a.param->iter : a.open.iter, a.open.scale);
    }
    collide<T, BULK>(f, relax_of<T>(a.bulk));
#pragma unroll
    for (int q = 0; q < 27; ++q)
        StoreTraits<S>::bstore(r, off[post_nb(q)], post_plane(q, FLIP) * LBM3_TILE * int(sizeof(S)), f[q], q, a.cd);
}

// the plain fluid cells of the open groups: record -5 in place of -1 / -2
__global__ __launch_bounds__(256) void k_mark_groups(const int *groups, int count, int G, int N, int2 *cellRow)
{
    const long long t = blockIdx.x * 256LL + threadIdx.x;
    if (t >= (long long)count * G) return;
    const int cell = groups[t / G] * G + int(t % G);
    if (cell >= N) return;
    const int code = cellRow[cell].x;
    if (code == -1 || code == -2) cellRow[cell].x = -5;
}

// the bulk kernel's view of the cell records: per 64 cells a mask of the cells
// it skips (record >= 0 or <= -3, and the padding past N), per 256 cells the
// row of the first one; with `open` non-null the open cells (record <= -8) are
// the bulk kernel's too, marked in `open`
__global__ __launch_bounds__(256) void k_bulk_mask(const int2 *cellRow, int N, uint64_t *skip, int *tileRow,
                                                   uint64_t *open, uint64_t *rotor)
{
    const int cell = blockIdx.x * 256 + threadIdx.x;
    bool sk = true, op = false, rt = false;
    int row = 0;
    if (cell < N) {
        const int2 r = cellRow[cell];
        op = open && r.x <= -8;
        rt = rotor && r.x == -4;
        sk = (r.x >= 0 || r.x <= -3) && !op && !rt;
        row = r.y;
    }
    const unsigned long long m = __ballot(sk), o = __ballot(op), q = __ballot(rt);
    if ((threadIdx.x & 63) == 0) {
        skip[cell >> 6] = m;
        if (open) open[cell >> 6] = o;
        if (rotor) rotor[cell >> 6] = q;
    }
    if (threadIdx.x == 0) tileRow[blockIdx.x] = row;
}

void launch_bulk_mask(const int2 *cellRow, int N, uint64_t *skip, int *tileRow, uint64_t *open, uint64_t *rotor,
                      hipStream_t s)
{
    if (N <= 0) return;
    k_bulk_mask<<<(N + 255) / 256, 256, 0, s>>>(cellRow, N, skip, tileRow, open, rotor);
}

// the cells of the open groups leave the bulk kernel's view: skip bits set,
// open bits cleared (a group of 8 or 16 aligned cells lies in one 64-bit word)
__global__ __launch_bounds__(256) void k_group_bits(const int *groups, int count, int G, uint64_t *skip, uint64_t *open)
{
    const int t = blockIdx.x * 256 + threadIdx.x;
    if (t >= count) return;
    const long long first = (long long)groups[t] * G;
    const unsigned long long bits = (G >= 64 ? ~0ull : ((1ull << G) - 1ull)) << (first & 63);
    atomicOr(reinterpret_cast<unsigned long long *>(skip + (first >> 6)), bits);
    atomicAnd(reinterpret_cast<unsigned long long *>(open + (first >> 6)), ~bits);
}

void launch_group_bits(const int *groups, int count, int G, uint64_t *skip, uint64_t *open, hipStream_t s)
{
    if (count <= 0) return;
    k_group_bits<<<(count + 255) / 256, 256, 0, s>>>(groups, count, G, skip, open);
}

void launch_mark_groups(const int *groups, int count, int G, int N, int2 *cellRow, hipStream_t s)
{
    if (count <= 0) return;
    hipLaunchKernelGGL(k_mark_groups, dim3((unsigned)(((long long)count * G + 255) / 256)), dim3(256), 0, s, groups, count,
                       G, N, cellRow);
}

// cells the rotors can reach (never wall-adjacent or open; record -4): the
// bulk update with the rotors' forcing and the scaled limiter per cell
template <class S, bool FLIP>
__global__ __launch_bounds__(256) void k_rotor_cells(StepArgs<S> a)
{
    using T = typename StoreTraits<S>::compute;
    const int x = blockIdx.x * 256 + threadIdx.x;
    if (x >= a.rotorCellCount) return;
    const int cell = a.rotorCells[x];
    const int row = a.fast.cellRow[cell].y;
    int nb[8];
    decode_fast(a.fast, cell, row, nb);
    T f[27];
#pragma unroll
    for (int q = 0; q < 27; ++q) f[q] = StoreTraits<S>::load(a.f + pop_index(pre_plane(q, FLIP), nb[pre_nb(q)]), q, a.cd);
    T dRho, ux, uy, uz;
    moments(f, dRho, ux, uy, uz);
    Relax<T> R = relax_of<T>(a.bulk);
    const int wm = a.wallMap[cell];
    int ijk[3];
    cell_ijk(a.topo, a.fast, cell, row, ijk);
    rotors_apply_at<S, T>(a, ijk, dRho, ux, uy, uz, R, a.trackRotor && wm != -2);
    collide<T, 0>(f, R);
#pragma unroll
    for (int q = 0; q < 27; ++q)
        StoreTraits<S>::store(a.f + pop_index(post_plane(q, FLIP), nb[post_nb(q)]), f[q], q, a.cd);
}

// wall-adjacent cells: collision with the wall id's local BC, interpolated
// bounceback of the links that cross the wall, momentum-exchange force
template <class S, bool FLIP>
__device__ __forceinline__ void wall_cell(const StepArgs<S> &a, int wm)
{
    using T = typename StoreTraits<S>::compute;
    const int N = a.fast.N;
    const int cell = a.wallIndex[wm];
    const int row = a.fast.cellRow[cell].y;
    int nb[8];
    decode_fast(a.fast, cell, row, nb);
    T f[27];
    load_pre<S, FLIP>(a.f, N, nb, f, a.cd);
    const uint32_t wd = a.wallData[wm];
    uint32_t codes[7];
#pragma unroll
    for (int k = 0; k < 7; ++k) codes[k] = a.linkLength[(size_t)k * (size_t)a.W + wm];
    const LocalBC &b = a.wallBC[(wd >> 27) & 31u];
    // wall velocity in the compute type: a wall without a velocity rule moves
    // with the fluid's own (pre-collision) velocity
    T wu[3] = {T(b.ux), T(b.uy), T(b.uz)};
    const int velMode = b.velMode;
    if (velMode == 0) {
        T d, u0, u1, u2;
        moments(f, d, u0, u1, u2);
        wu[0] = u0; wu[1] = u1; wu[2] = u2;
    } else if (velMode == 2) {
        int nbx[8], ijk[3];
        decode(a.topo, cell, nbx, ijk);
        const float x = float(ijk[0]) * a.geom.res + a.geom.ox - b.origin[0];
        const float y = float(ijk[1]) * a.geom.res + a.geom.oy - b.origin[1];
        const float z = float(ijk[2]) * a.geom.res + a.geom.oz - b.origin[2];
        const float s = a.geom.dtPhys / a.geom.res;
        wu[0] = T((b.omega[1] * z - b.omega[2] * y) * s);
        wu[1] = T((b.omega[2] * x - b.omega[0] * z) * s);
        wu[2] = T((b.omega[0] * y - b.omega[1] * x) * s);
    } else if (velMode == 3) {
        wu[0] = T(a.wallVel[wm]);
        wu[1] = T(a.wallVel[(size_t)a.W + wm]);
        wu[2] = T(a.wallVel[2 * (size_t)a.W + wm]);
    }
    Relax<T> R = relax_of<T>(b);
    if (a.nRotors > 0) {
        T d, u0, u1, u2;
        moments(f, d, u0, u1, u2);
        rotors_apply<S, T>(a, cell, d, u0, u1, u2, R, a.trackRotor && !(wd & 1u));
    }
    T pre[27];
#pragma unroll
    for (int q = 0; q < 27; ++q) pre[q] = f[q];
    collide<T, 0>(f, R);
    double g3[3] = {0.0, 0.0, 0.0};
    bounce_back<T>(f, pre, wd, b.useIBB != 0, codes, wu, g3);
    // a reflected population is read next step as the opposite direction
#pragma unroll
    for (int d = 0; d < 27; ++d) {
        if (d != 0 && ((wd >> d) & 1u))
            StoreTraits<S>::store(a.f + pop_index(pre_plane(inv(d), !FLIP), nb[pre_nb(inv(d))]), f[d], inv(d), a.cd);
        else
            StoreTraits<S>::store(a.f + pop_index(post_plane(d, FLIP), nb[post_nb(d)]), f[d], d, a.cd);
    }
    if (a.gWall && !(wd & 1u)) {
        a.gWall[wm] += g3[0];
        a.gWall[(size_t)a.W + wm] += g3[1];
        a.gWall[2 * (size_t)a.W + wm] += g3[2];
    }
}

// neighbour cells of every open-boundary cell, decoded once at setup
__global__ __launch_bounds__(256) void k_open_nb(Topology t, int count, const int *cell, int4 *out)
{
    const int x = blockIdx.x * 256 + threadIdx.x;
    if (x >= count) return;
    int nb[8];
    decode(t, cell[x], nb, nullptr);
    out[2 * (size_t)x] = make_int4(nb[0], nb[1], nb[2], nb[3]);
    out[2 * (size_t)x + 1] = make_int4(nb[4], nb[5], nb[6], nb[7]);
}

template <class S, bool FLIP>
__global__ __launch_bounds__(256) void k_collide_wall(StepArgs<S> a)
{
    const int wm = blockIdx.x * 256 + threadIdx.x;
    if (wm >= a.W) return;
    wall_cell<S, FLIP>(a, wm);
}

// The wall path of the combined launch below as a call of its own, for 16-bit
// storage with the open cells in the bulk path: inlined there, the register
// allocation of the two paths together slowed the bulk blocks about 3.5x; as a
// callee, the wall path spills within its own frame.  The arguments stay in
// the kernel's constant argument segment (address space 4), read by scalar
// loads.
#ifndef LBM3_WALL_CALL
#define LBM3_WALL_CALL 1
#endif
template <class S, bool FLIP>
__device__ __noinline__ void wall_cell_call(const __attribute__((address_space(4))) StepArgs<S> *a, int wm)
{
    wall_cell<S, FLIP>(*(const StepArgs<S> *)a, wm);
}

// The bulk kernel with the wall-adjacent cells in blocks of their own ahead of
// the bulk blocks, in one launch: no second stream and no join per step.  The
// wall path needs about twice the bulk path's registers; the register budget is
// held to six waves per SIMD (the bulk path fits it, the wall path spills), so
// the bulk blocks keep their occupancy.  BF16 with the open cells inside: seven
// waves (72 registers, no spill in the bulk path), measured on the
// 256x128x128 sphere at 0.0988-0.0996 ms per step against 0.1012-0.1025 with
// six; FP16 measured the other way round (0.0976-0.0990 with six, and seven
// lost to the wall kernel on the side stream).
template <class S>
constexpr int walls_waves(bool open)
{
    return open && std::is_same<S, __hip_bfloat16>::value ? 7 : 6;
}

template <class S, bool FLIP, int BULK, bool OPEN>
__global__ __launch_bounds__(256, walls_waves<S>(OPEN)) void k_collide_stream_walls(StepArgs<S> a)
{
    const int b = blockIdx.x;
    if (b < a.wallBlocks) {
        const int wm = b * 256 + threadIdx.x;
        if (wm < a.W) {
            if constexpr (sizeof(S) == 2 && OPEN && LBM3_WALL_CALL)
                // the kernel's one argument starts the argument segment; its
                // address, not &a, which would copy the argument to the stack
                wall_cell_call<S, FLIP>(
                    (const __attribute__((address_space(4))) StepArgs<S> *)__builtin_amdgcn_kernarg_segment_ptr(), wm);
            else
                wall_cell<S, FLIP>(a, wm);
        }
        return;
    }
    const int cell = xcd_block(b - a.wallBlocks, gridDim.x - a.wallBlocks, a.xcdRemap, FLIP) * 256 + threadIdx.x;
    if (cell >= a.fast.N) return;
    bulk_cell<S, FLIP, BULK, uint32_t, OPEN>(a, cell);
}

// derive the hot-loop topology records from the compressed IJK/NBR arrays;
// with rotorCand, the fluid cells the rotors can reach get -4, and with openOf
// the open-boundary cells get -(8 + open-list index), in place of
// their wall-map code (they are never wall-adjacent when this is used)
__global__ __launch_bounds__(256) void k_fast_topo(Topology t, const int *wallMap, const int *openOf, const uint8_t *rotorCand,
                                                   int2 *cellRow, int4 *rows)
{
    const int cell = blockIdx.x * 256 + threadIdx.x;
    if (cell >= t.N) return;
    const int shift = t.shifter[cell];
    const int row = shift >= 0 ? shift : t.shifter[cell + shift];
    int code = wallMap[cell];
    if (rotorCand && rotorCand[cell] && code >= -2 && code < 0) code = -4;
    if (openOf && openOf[cell] >= 0) code = -(8 + openOf[cell]);
    cellRow[cell] = make_int2(code, row);
    if (shift >= 0) rows[row] = make_int4(cell, t.rjp[row], t.rkp[row], t.rjkp[row]);
}

// cells a later pass of the step reads or writes as a whole (periodic copies,
// interface stencils) must not carry a boundary condition that is still pending
__global__ __launch_bounds__(256) void k_open_conflicts(const int *cells, int count, int mode, const int *openOf,
                                                        const int *wallMap, const uint8_t *cand, FastTopo t, int *flag)
{
    const int x = blockIdx.x * 256 + threadIdx.x;
    if (x >= count) return;
    int c = cells[x];
    bool hit = false;
    if (mode == 0) hit = openOf[c] >= 0;                                  // listed cell
    else if (mode == 1) { if (c < 0) c = -c - 1; hit = openOf[c] >= 0; }  // signed (leftover neighbours)
    else if (mode == 2) {                                                  // 2x2x2 stencil from its first cell
        int nb[8];
        decode_fast(t, c, t.cellRow[c].y, nb);
        for (int k = 0; k < 8; ++k) hit = hit || openOf[nb[k]] >= 0;
    } else if (mode == 3) hit = wallMap[c] >= 0;                           // open cell next to a wall
    else hit = cand && cand[c];                                            // open cell a rotor can reach
    if (hit) atomicAdd(flag, 1);
}

// cells a rotor can reach at some angle: within its bounds' distance from the
// axis (plus two level spacings) and inside its axial extent (the same margin)
__global__ __launch_bounds__(256) void k_rotor_reach(Topology t, LevelGeom g, RotorReachSet R, uint8_t *cand)
{
    const int cell = blockIdx.x * 256 + threadIdx.x;
    if (cell >= t.N) return;
    int nb[8], ijk[3];
    decode(t, cell, nb, ijk);
    const double p[3] = {double(float(ijk[0]) * g.res + g.ox), double(float(ijk[1]) * g.res + g.oy),
                         double(float(ijk[2]) * g.res + g.oz)};
    const double m = 2.0 * double(g.res);
    uint8_t hit = 0;
    for (int r = 0; r < R.n && !hit; ++r) {
        const RotorReach &q = R.r[r];
        const double lo[3] = {q.lo[0] - m, q.lo[1] - m, q.lo[2] - m}, hi[3] = {q.hi[0] + m, q.hi[1] + m, q.hi[2] + m};
        if (q.axis < 0) {
            hit = p[0] >= lo[0] && p[0] <= hi[0] && p[1] >= lo[1] && p[1] <= hi[1] && p[2] >= lo[2] && p[2] <= hi[2];
            continue;
        }
        const int a = q.axis, b = (a + 1) % 3, c = (a + 2) % 3;
        if (p[a] < lo[a] || p[a] > hi[a]) continue;
        double rmax2 = 0.0;
        for (int k = 0; k < 4; ++k) {
            const double db = ((k & 1) ? q.hi[b] : q.lo[b]) - q.o[b], dc = ((k & 2) ? q.hi[c] : q.lo[c]) - q.o[c];
            rmax2 = fmax(rmax2, db * db + dc * dc);
        }
        const double db = p[b] - q.o[b], dc = p[c] - q.o[c];
        hit = sqrt(db * db + dc * dc) <= sqrt(rmax2) + m;
    }
    cand[cell] = hit;
}

void launch_rotor_reach(const Topology &t, const LevelGeom &g, const RotorReachSet &R, uint8_t *cand, hipStream_t s)
{
    if (t.N <= 0) return;
    hipLaunchKernelGGL(k_rotor_reach, dim3((unsigned)((t.N + 255) / 256)), dim3(256), 0, s, t, g, R, cand);
}

__global__ __launch_bounds__(256) void k_scatter_index(const int *cell, int count, int *openOf)
{
    const int x = blockIdx.x * 256 + threadIdx.x;
    if (x < count) openOf[cell[x]] = x;
}

template <int NI, class S, bool FLIP>
__device__ __forceinline__ void open_one(const OpenArgs<S> &a, int x, const int (&nb)[8], uint32_t meta)
{
    using T = typename StoreTraits<S>::compute;
    constexpr int nx = LBM3_BT_NORMAL[NI][0], ny = LBM3_BT_NORMAL[NI][1], nz = LBM3_BT_NORMAL[NI][2];
    T f[27];
#pragma unroll
    for (int q = 0; q < 27; ++q)
        f[q] = unknown_dir(nx, ny, nz, q) ? T(0)
                                          : StoreTraits<S>::load(a.f + pop_index(pre_plane(q, FLIP), nb[pre_nb(q)]), q, a.cd);
    open_fill<NI, T>(f, x, meta, a.records, a.dRhoRef, a.uNormalRef, a.dRhoCum, a.uNormalCum,
                     a.param ? a.param->iterNow : a.iterationsFinished, a.scale);
#pragma unroll
    for (int q = 0; q < 27; ++q)
        if (unknown_dir(nx, ny, nz, q))
            StoreTraits<S>::store(a.f + pop_index(pre_plane(q, FLIP), nb[pre_nb(q)]), f[q], q, a.cd);
}

template <class S, bool FLIP>
__device__ __forceinline__ void open_entry(const OpenArgs<S> &a, int x)
{
    const int4 n0 = a.nb[2 * (size_t)x], n1 = a.nb[2 * (size_t)x + 1];
    const int nb[8] = {n0.x, n0.y, n0.z, n0.w, n1.x, n1.y, n1.z, n1.w};
    const uint32_t meta = a.meta[x];
    switch (meta & 31u) {
#define OPEN_CASE(n) case n: open_one<n, S, FLIP>(a, x, nb, meta); break;
        OPEN_CASE(0) OPEN_CASE(1) OPEN_CASE(2) OPEN_CASE(3) OPEN_CASE(4) OPEN_CASE(5) OPEN_CASE(6)
        OPEN_CASE(7) OPEN_CASE(8) OPEN_CASE(9) OPEN_CASE(10) OPEN_CASE(11) OPEN_CASE(12) OPEN_CASE(13)
        OPEN_CASE(14) OPEN_CASE(15) OPEN_CASE(16) OPEN_CASE(17) OPEN_CASE(18) OPEN_CASE(19) OPEN_CASE(20)
        OPEN_CASE(21) OPEN_CASE(22) OPEN_CASE(23) OPEN_CASE(24) OPEN_CASE(25)
#undef OPEN_CASE
    default: break;
    }
}

template <class S, bool FLIP>
__global__ __launch_bounds__(256) void k_open(OpenArgs<S> a)
{
    const int t = blockIdx.x * 256 + threadIdx.x;
    if (t >= a.count) return;
    open_entry<S, FLIP>(a, a.list ? a.list[t] : t);
}

// the stored values are copied as they are: source and ghost read them as the same directions
template <class S, bool FLIP>
__global__ __launch_bounds__(256) void k_periodic(S *f, Topology topo, int count, const int *index, const int *source)
{
    const int x = blockIdx.x * 256 + threadIdx.x;
    if (x >= count) return;
    int ns[8], nd[8];
    decode(topo, source[x], ns, nullptr);
    decode(topo, index[x], nd, nullptr);
    S v[27];
#pragma unroll
    for (int q = 0; q < 27; ++q) v[q] = f[pop_index(pre_plane(q, FLIP), ns[pre_nb(q)])];
#pragma unroll
    for (int q = 0; q < 27; ++q) f[pop_index(pre_plane(q, FLIP), nd[pre_nb(q)])] = v[q];
}

template <class S, bool FLIP>
__global__ __launch_bounds__(256) void k_init(S *f, Topology topo, InitDesc d, LevelGeom g, Codec cd)
{
    using T = typename StoreTraits<S>::compute;
    const int cell = blockIdx.x * 256 + threadIdx.x;
    if (cell >= topo.N) return;
    int nb[8], ijk[3];
    decode(topo, cell, nb, ijk);
    float dRho = d.dRho, ux = d.ux, uy = d.uy, uz = d.uz;
    if (d.kind == 1) {
        const float X = (float(ijk[0]) * g.res + g.ox) / d.length;
        const float Y = (float(ijk[1]) * g.res + g.oy) / d.length;
        const float Z = (float(ijk[2]) * g.res + g.oz) / d.length;
        ux = d.u0 * sinf(X) * cosf(Y) * cosf(Z);
        uy = -d.u0 * cosf(X) * sinf(Y) * cosf(Z);
        uz = 0.f;
        dRho = d.pressure ? (3.f * d.u0 * d.u0 / 16.f) * (cosf(2.f * X) + cosf(2.f * Y)) * (2.f + cosf(2.f * Z)) : 0.f;
    } else if (d.kind == 2) {
        dRho = d.fields[4 * (size_t)cell]; ux = d.fields[4 * (size_t)cell + 1];
        uy = d.fields[4 * (size_t)cell + 2]; uz = d.fields[4 * (size_t)cell + 3];
    } else if (d.profile.radius > 0.f) {
        float x, y, z;
        cell_xyz(g, ijk, x, y, z);
        const float sc = profile_factor(d.profile, x, y, z);
        ux *= sc; uy *= sc; uz *= sc;
    }
    T v[27];
#pragma unroll
    for (int q = 0; q < 27; ++q) v[q] = feq_shifted<T>(q, T(dRho), T(ux), T(uy), T(uz));
    store_pre<S, FLIP>(f, topo.N, nb, v, cd);
}

template <class S, bool FLIP>
__global__ __launch_bounds__(256) void k_open_ref_init(OpenArgs<S> a, int cx_, int cy_, int cz_)
{
    using T = typename StoreTraits<S>::compute;
    const int x = blockIdx.x * 256 + threadIdx.x;
    if (x >= a.count) return;
    int nb[8], ijk[3];
    decode(a.topo, a.cell[x], nb, ijk);
    T v[27];
    load_pre<S, FLIP>(a.f, a.topo.N, nb, v, a.cd);
    T d, ux, uy, uz;
    moments(v, d, ux, uy, uz);
    const int nx = ijk[0] == 0 ? -1 : (ijk[0] == cx_ - 1 ? 1 : 0);
    const int ny = ijk[1] == 0 ? -1 : (ijk[1] == cy_ - 1 ? 1 : 0);
    const int nz = ijk[2] == 0 ? -1 : (ijk[2] == cz_ - 1 ? 1 : 0);
    a.dRhoRef[x] = d;
    a.uNormalRef[x] = T(nx) * ux + T(ny) * uy + T(nz) * uz;
}

template <class S, bool FLIP>
__global__ __launch_bounds__(256) void k_macros(const S *f, Topology topo, typename StoreTraits<S>::compute *out, Codec cd)
{
    using T = typename StoreTraits<S>::compute;
    const int cell = blockIdx.x * 256 + threadIdx.x;
    if (cell >= topo.N) return;
    int nb[8];
    decode(topo, cell, nb, nullptr);
    T v[27];
    load_pre<S, FLIP>(f, topo.N, nb, v, cd);
    T d, ux, uy, uz;
    moments(v, d, ux, uy, uz);
    out[4 * (size_t)cell] = d;
    out[4 * (size_t)cell + 1] = ux;
    out[4 * (size_t)cell + 2] = uy;
    out[4 * (size_t)cell + 3] = uz;
}

__device__ __forceinline__ double block_sum(double v, double *lds)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) lds[wid] = v;
    __syncthreads();
    double s = 0.0;
    if (threadIdx.x == 0)
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) s += lds[w];
    return s;
}

// per-block partial sums over the counted cells -- not solid, not a periodic
// ghost, not an interface duplicate (taylorGreenVortexMain.cu:141-145):
// MASS false: the kinetic energy 0.5 (1 + dRho) |u|^2 into partial[block];
// MASS true: dRho into partial[block] and the cell count into partial[gridDim.x + block]
template <class S, bool FLIP, bool MASS>
__global__ __launch_bounds__(256) void k_cell_sums(const S *f, Topology topo, const int *wallMap, const uint32_t *wallData,
                                                   double *partial, Codec cd)
{
    __shared__ double lds[4];
    using T = typename StoreTraits<S>::compute;
    const int cell = blockIdx.x * 256 + threadIdx.x;
    double e = 0.0, n = 0.0;
    if (cell < topo.N) {
        const int wm = wallMap[cell];
        const bool skip = wm == -3 || wm == -2 || (wm >= 0 && (wallData[wm] & 1u));
        if (!skip) {
            int nb[8];
            decode(topo, cell, nb, nullptr);
            T v[27];
            load_pre<S, FLIP>(f, topo.N, nb, v, cd);
            T d, ux, uy, uz;
            moments(v, d, ux, uy, uz);
            if (MASS) e = double(d);
            else e = 0.5 * (1.0 + double(d)) * (double(ux) * double(ux) + double(uy) * double(uy) + double(uz) * double(uz));
            n = 1.0;
        }
    }
    const double s = block_sum(e, lds);
    if (threadIdx.x == 0) partial[blockIdx.x] = s;
    if (MASS) {
        __syncthreads();
        const double c = block_sum(n, lds);
        if (threadIdx.x == 0) partial[gridDim.x + blockIdx.x] = c;
    }
}

// one block: out = sum of partial[0..n) / divisor, added to the value out holds
// when accumulate is set (levels weighted and summed on the device, in order)
__global__ __launch_bounds__(256) void k_sum_partials(const double *partial, int n, double *out, double divisor, int accumulate)
{
    __shared__ double lds[4];
    double v = 0.0;
    for (int i = threadIdx.x; i < n; i += 256) v += partial[i];
    const double s = block_sum(v, lds);
    if (threadIdx.x == 0) *out = accumulate ? *out + s / divisor : s / divisor;
}

// open-boundary flow statistics (trackerFunctions.h:145-170): per block, the
// sums over the tracked entries [start, start + count) of the open list of
// uNormal, rho uNormal, rho uNormal |uNormal|, dRho, dRho uNormal,
// rho uNormal^3 / 2 (time averages of the accumulators, lattice units) and the
// tracked-entry count; partial[block * 7 + k]
__global__ __launch_bounds__(256) void k_open_stats(const double *dRhoCum, const double *uNormalCum, const uint32_t *meta,
                                                    int start, int count, double invSpan, double *partial)
{
    __shared__ double lds[7 * 4];
    const int i = blockIdx.x * 256 + threadIdx.x;
    double v[7] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    if (i < count && (meta[start + i] & kOpenTrackBit)) {
        const double dRho = dRhoCum[start + i] * invSpan, un = uNormalCum[start + i] * invSpan;
        const double rho = 1.0 + dRho;
        v[0] = un;
        v[1] = rho * un;
        v[2] = rho * un * fabs(un);
        v[3] = dRho;
        v[4] = dRho * un;
        v[5] = 0.5 * rho * un * un * un;
        v[6] = 1.0;
    }
#pragma unroll
    for (int k = 0; k < 7; ++k) {
        const double sum = block_sum(v[k], lds + 4 * k);
        if (threadIdx.x == 0) partial[7 * (size_t)blockIdx.x + k] = sum;
    }
}

int launch_open_stats(const double *dRhoCum, const double *uNormalCum, const uint32_t *meta, int start, int count,
                      double invSpan, double *partial, hipStream_t s)
{
    const int blocks = (count + 255) / 256;
    if (blocks > 0) k_open_stats<<<blocks, 256, 0, s>>>(dRhoCum, uNormalCum, meta, start, count, invSpan, partial);
    return blocks;
}

// per-wall-body sums of the accumulated force and its torque about `origin`;
// out[body*6 + {Fx,Fy,Fz,Tx,Ty,Tz}] per block
__global__ __launch_bounds__(256) void k_wall_force(const double *gWall, int W, const uint32_t *wallData, const int *wallIndex,
                                                    Topology topo, LevelGeom g, double ox, double oy, double oz, int nBodies,
                                                    double *partial)
{
    // a fixed-order tree sum per body and component: the same bits on every
    // run, whatever the order the block's waves run in
    __shared__ double red[6][256];
    const int t = threadIdx.x;
    const int w = blockIdx.x * 256 + t;
    int body = -1;
    double v[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    if (w < W) {
        body = (wallData[w] >> 27) & 31u;
        if (body < nBodies) {
            int nb[8], ijk[3];
            decode(topo, wallIndex[w], nb, ijk);
            const double x = double(float(ijk[0]) * g.res + g.ox) - ox;
            const double y = double(float(ijk[1]) * g.res + g.oy) - oy;
            const double z = double(float(ijk[2]) * g.res + g.oz) - oz;
            v[0] = gWall[w];
            v[1] = gWall[(size_t)W + w];
            v[2] = gWall[2 * (size_t)W + w];
            v[3] = y * v[2] - z * v[1];
            v[4] = z * v[0] - x * v[2];
            v[5] = x * v[1] - y * v[0];
        } else {
            body = -1;
        }
    }
    double *out = partial + (size_t)blockIdx.x * nBodies * 6;
    for (int b = 0; b < nBodies; ++b) {
        if (!__syncthreads_or(body == b)) {      // no cell of this body in the block
            if (t < 6) out[b * 6 + t] = 0.0;
            continue;
        }
#pragma unroll
        for (int k = 0; k < 6; ++k) red[k][t] = body == b ? v[k] : 0.0;
        __syncthreads();
        for (int s = 128; s > 0; s >>= 1) {
            if (t < s) {
#pragma unroll
                for (int k = 0; k < 6; ++k) red[k][t] += red[k][t + s];
            }
            __syncthreads();
        }
        if (t < 6) out[b * 6 + t] = red[t][0];
    }
}

// unit-level entry: collide n cells stored cell-major (27 per cell)
template <class T>
__global__ __launch_bounds__(256) void k_collide_batch(T *f, int n, Relax<T> R)
{
    const int c = blockIdx.x * 256 + threadIdx.x;
    if (c >= n) return;
    T v[27];
#pragma unroll
    for (int q = 0; q < 27; ++q) v[q] = f[27 * (size_t)c + q];
    collide<T, 0>(v, R);
#pragma unroll
    for (int q = 0; q < 27; ++q) f[27 * (size_t)c + q] = v[q];
}

// ---------------------------------------------------------------------------
// many domains in one launch (continuous-batching server): block b belongs to
// the domain d with start[d] <= b < start[d + 1]; each domain carries its own
// arguments, esotwist parity and collision variant
// ---------------------------------------------------------------------------
// The domain is the same for every lane of a block; readfirstlane states it,
// so that the domain's descriptor is read by scalar loads and the buffer
// resource made from its population pointer sits in scalar registers.  Left
// to the compiler, the index counted as per-lane, and every raw buffer access
// became a loop over the wave's distinct resources (168 such loops in the
// bulk kernel).
__device__ __forceinline__ int domain_of(const int *start, int n, int b)
{
    int lo = 0, hi = n;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (start[mid] <= b) lo = mid;
        else hi = mid;
    }
    return __builtin_amdgcn_readfirstlane(lo);
}

template <class S>
__global__ __launch_bounds__(256) void k_multi_bulk(const MultiDesc<S> *__restrict__ D, const int *__restrict__ start, int n)
{
    const int d = domain_of(start, n, blockIdx.x);
    const MultiDesc<S> &m = D[d];
    const int cell = (blockIdx.x - start[d]) * 256 + threadIdx.x;
    if (cell >= m.step.fast.N) return;
    // server domains are far below 4 GiB each (checked at admission): 32-bit offsets
    if (m.flip) {
        if (m.k17) bulk_cell<S, true, 1, uint32_t>(m.step, cell);
        else bulk_cell<S, true, 2, uint32_t>(m.step, cell);
    } else {
        if (m.k17) bulk_cell<S, false, 1, uint32_t>(m.step, cell);
        else bulk_cell<S, false, 2, uint32_t>(m.step, cell);
    }
}

template <class S>
__global__ __launch_bounds__(256) void k_multi_wall(const MultiDesc<S> *__restrict__ D, const int *__restrict__ start, int n)
{
    const int d = domain_of(start, n, blockIdx.x);
    const MultiDesc<S> &m = D[d];
    const int wm = (blockIdx.x - start[d]) * 256 + threadIdx.x;
    if (wm >= m.step.W) return;
    if (m.flip) wall_cell<S, true>(m.step, wm);
    else wall_cell<S, false>(m.step, wm);
}

// after streaming: the open boundary and the periodic copy at the new parity
template <class S>
__global__ __launch_bounds__(256) void k_multi_open(const MultiDesc<S> *__restrict__ D, const int *__restrict__ start, int n)
{
    const int d = domain_of(start, n, blockIdx.x);
    const MultiDesc<S> &m = D[d];
    const int x = (blockIdx.x - start[d]) * 256 + threadIdx.x;
    if (x >= m.open.count) return;
    if (m.flip) open_entry<S, false>(m.open, x);
    else open_entry<S, true>(m.open, x);
}

template <class S, bool FLIP>
__device__ __forceinline__ void periodic_entry(S *f, const Topology &topo, const int *index, const int *source, int x)
{
    int ns[8], nd[8];
    decode(topo, source[x], ns, nullptr);
    decode(topo, index[x], nd, nullptr);
    S v[27];
#pragma unroll
    for (int q = 0; q < 27; ++q) v[q] = f[pop_index(pre_plane(q, FLIP), ns[pre_nb(q)])];
#pragma unroll
    for (int q = 0; q < 27; ++q) f[pop_index(pre_plane(q, FLIP), nd[pre_nb(q)])] = v[q];
}

template <class S>
__global__ __launch_bounds__(256) void k_multi_periodic(const MultiDesc<S> *__restrict__ D, const int *__restrict__ start, int n)
{
    const int d = domain_of(start, n, blockIdx.x);
    const MultiDesc<S> &m = D[d];
    const int x = (blockIdx.x - start[d]) * 256 + threadIdx.x;
    if (x >= m.nPeriodic) return;
    if (m.flip) periodic_entry<S, false>(m.step.f, m.step.topo, m.perIndex, m.perSource, x);
    else periodic_entry<S, true>(m.step.f, m.step.topo, m.perIndex, m.perSource, x);
}

// ---------------------------------------------------------------------------
// host launchers
// ---------------------------------------------------------------------------
static inline unsigned blocks_for(long long n) { return (unsigned)((n + 255) / 256); }

template <class S>
void launch_collide_wall(const StepArgs<S> &a, bool flip, hipStream_t s)
{
    if (a.W <= 0) return;
    if (flip) hipLaunchKernelGGL((k_collide_wall<S, true>), dim3(blocks_for(a.W)), dim3(256), 0, s, a);
    else hipLaunchKernelGGL((k_collide_wall<S, false>), dim3(blocks_for(a.W)), dim3(256), 0, s, a);
}


template <class S>
void launch_collide_stream(const StepArgs<S> &a, bool flip, hipStream_t s)
{
    if (a.fast.N <= 0) return;
    const dim3 g(blocks_for(a.fast.N)), b(256);
    const unsigned lds = (unsigned)a.bulkLds;
    const bool k17 = relax_of<typename StoreTraits<S>::compute>(a.bulk).k17 != 0;
    const bool wide = wide_level<S>(a.fast.N);
    if (a.wallBlocks > 0) {
        // the caller combines only levels without rotor cells in the bulk
        // kernel, rotated lanes or 64-bit offsets
        const dim3 gw(blocks_for(a.fast.N) + (unsigned)a.wallBlocks);
#define LBM3_BULK_WALLS(F, K, O) hipLaunchKernelGGL((k_collide_stream_walls<S, F, K, O>), gw, b, lds, s, a)
        const bool o = a.fast.openBits != nullptr;
        if (flip) { if (k17) { if (o) LBM3_BULK_WALLS(true, 1, true); else LBM3_BULK_WALLS(true, 1, false); }
                    else { if (o) LBM3_BULK_WALLS(true, 2, true); else LBM3_BULK_WALLS(true, 2, false); } }
        else { if (k17) { if (o) LBM3_BULK_WALLS(false, 1, true); else LBM3_BULK_WALLS(false, 1, false); }
               else { if (o) LBM3_BULK_WALLS(false, 2, true); else LBM3_BULK_WALLS(false, 2, false); } }
#undef LBM3_BULK_WALLS
        return;
    }
#define LBM3_BULK_LAUNCH(F, K, W) hipLaunchKernelGGL((k_collide_stream<S, F, K, W>), g, b, lds, s, a)
#define LBM3_BULK_OPEN(F, K) hipLaunchKernelGGL((k_collide_stream<S, F, K, false, true>), g, b, lds, s, a)
#define LBM3_BULK_OPEN_ROT(F, K) hipLaunchKernelGGL((k_collide_stream<S, F, K, false, true, false, true>), g, b, lds, s, a)
#define LBM3_BULK_ROTOR(F, K) hipLaunchKernelGGL((k_collide_stream<S, F, K, false, false, true>), g, b, lds, s, a)
    if (a.fast.rotorBits && !wide) {
        if (flip) { if (k17) LBM3_BULK_ROTOR(true, 1); else LBM3_BULK_ROTOR(true, 2); }
        else { if (k17) LBM3_BULK_ROTOR(false, 1); else LBM3_BULK_ROTOR(false, 2); }
    } else if (a.fast.openBits && !wide && a.laneRot) {
        if (flip) { if (k17) LBM3_BULK_OPEN_ROT(true, 1); else LBM3_BULK_OPEN_ROT(true, 2); }
        else { if (k17) LBM3_BULK_OPEN_ROT(false, 1); else LBM3_BULK_OPEN_ROT(false, 2); }
    } else if (a.fast.openBits && !wide) {
        if (flip) { if (k17) LBM3_BULK_OPEN(true, 1); else LBM3_BULK_OPEN(true, 2); }
        else { if (k17) LBM3_BULK_OPEN(false, 1); else LBM3_BULK_OPEN(false, 2); }
    } else if (wide) {
        if (flip) { if (k17) LBM3_BULK_LAUNCH(true, 1, true); else LBM3_BULK_LAUNCH(true, 2, true); }
        else { if (k17) LBM3_BULK_LAUNCH(false, 1, true); else LBM3_BULK_LAUNCH(false, 2, true); }
    } else {
        if (flip) { if (k17) LBM3_BULK_LAUNCH(true, 1, false); else LBM3_BULK_LAUNCH(true, 2, false); }
        else { if (k17) LBM3_BULK_LAUNCH(false, 1, false); else LBM3_BULK_LAUNCH(false, 2, false); }
    }
#undef LBM3_BULK_LAUNCH
#undef LBM3_BULK_OPEN
#undef LBM3_BULK_OPEN_ROT
#undef LBM3_BULK_ROTOR
}

template <class S>
void launch_rotor_cells(const StepArgs<S> &a, bool flip, hipStream_t s)
{
    if (a.rotorCellCount <= 0) return;
    if (flip) hipLaunchKernelGGL((k_rotor_cells<S, true>), dim3(blocks_for(a.rotorCellCount)), dim3(256), 0, s, a);
    else hipLaunchKernelGGL((k_rotor_cells<S, false>), dim3(blocks_for(a.rotorCellCount)), dim3(256), 0, s, a);
}

template <class S, int G>
void launch_open_groups(const StepArgs<S> &a, bool flip, hipStream_t s)
{
    const dim3 g((unsigned)((a.open.count + 256 / G - 1) / (256 / G))), b(256);
    const bool k17 = relax_of<typename StoreTraits<S>::compute>(a.bulk).k17 != 0;
    if (flip) { if (k17) hipLaunchKernelGGL((k_open_groups<S, true, 1, G>), g, b, 0, s, a); else hipLaunchKernelGGL((k_open_groups<S, true, 2, G>), g, b, 0, s, a); }
    else { if (k17) hipLaunchKernelGGL((k_open_groups<S, false, 1, G>), g, b, 0, s, a); else hipLaunchKernelGGL((k_open_groups<S, false, 2, G>), g, b, 0, s, a); }
}

// the group widths the open-cell groups may be built with (Sim::Tune::openGroup)
template <class S>
void launch_open_cells(const StepArgs<S> &a, bool flip, hipStream_t s)
{
    if (a.open.count <= 0) return;
    if (a.open.groupCells == 16) launch_open_groups<S, 16>(a, flip, s);
    else launch_open_groups<S, 8>(a, flip, s);
}

// velocity profile factor of each open-list entry: its rule's radial profile
// at the cell centre (getOpenBC of the pipe cases, poiseuilleFlowMain.cu:58-75)
__global__ __launch_bounds__(256) void k_open_scale(Topology t, LevelGeom g, int count, const int *cell,
                                                     const uint32_t *meta, OpenProfiles P, float *scale)
{
    const int x = blockIdx.x * 256 + threadIdx.x;
    if (x >= count) return;
    int nb[8], ijk[3];
    decode(t, cell[x], nb, ijk);
    float px, py, pz;
    cell_xyz(g, ijk, px, py, pz);
    scale[x] = profile_factor(P.p[open_meta_record(meta[x])], px, py, pz);
}

void launch_open_scale(const Topology &t, const LevelGeom &g, int count, const int *cell, const uint32_t *meta,
                       const OpenProfiles &P, float *scale, hipStream_t s)
{
    if (count <= 0) return;
    k_open_scale<<<(count + 255) / 256, 256, 0, s>>>(t, g, count, cell, meta, P, scale);
}

void launch_open_nb(const Topology &t, int count, const int *cell, int4 *out, hipStream_t s)
{
    if (count <= 0) return;
    hipLaunchKernelGGL(k_open_nb, dim3(blocks_for(count)), dim3(256), 0, s, t, count, cell, out);
}

void launch_fast_topo(const Topology &t, const int *wallMap, const int *openOf, const uint8_t *rotorCand, int2 *cellRow,
                      int4 *rows, hipStream_t s)
{
    if (t.N <= 0) return;
    hipLaunchKernelGGL(k_fast_topo, dim3(blocks_for(t.N)), dim3(256), 0, s, t, wallMap, openOf, rotorCand, cellRow, rows);
}

void launch_open_conflicts(const int *cells, int count, int mode, const int *openOf, const int *wallMap,
                           const uint8_t *cand, const FastTopo &t, int *flag, hipStream_t s)
{
    if (count <= 0) return;
    hipLaunchKernelGGL(k_open_conflicts, dim3(blocks_for(count)), dim3(256), 0, s, cells, count, mode, openOf, wallMap, cand,
                       t, flag);
}

void launch_scatter_index(const int *cell, int count, int *openOf, hipStream_t s)
{
    if (count <= 0) return;
    hipLaunchKernelGGL(k_scatter_index, dim3(blocks_for(count)), dim3(256), 0, s, cell, count, openOf);
}

// openOf[cell of entry] = entry, for the listed open-list entries only
__global__ __launch_bounds__(256) void k_scatter_entries(const int *cell, const int *list, int count, int *openOf)
{
    const int t = blockIdx.x * 256 + threadIdx.x;
    if (t < count) openOf[cell[list[t]]] = list[t];
}

void launch_scatter_entries(const int *cell, const int *list, int count, int *openOf, hipStream_t s)
{
    if (count <= 0) return;
    hipLaunchKernelGGL(k_scatter_entries, dim3(blocks_for(count)), dim3(256), 0, s, cell, list, count, openOf);
}

template <class S>
void launch_open(const OpenArgs<S> &a, bool flip, hipStream_t s)
{
    if (a.count <= 0) return;
    if (flip) hipLaunchKernelGGL((k_open<S, true>), dim3(blocks_for(a.count)), dim3(256), 0, s, a);
    else hipLaunchKernelGGL((k_open<S, false>), dim3(blocks_for(a.count)), dim3(256), 0, s, a);
}

template <class S>
void launch_periodic(S *f, const Topology &t, int count, const int *index, const int *source, bool flip, hipStream_t s)
{
    if (count <= 0) return;
    if (flip) hipLaunchKernelGGL((k_periodic<S, true>), dim3(blocks_for(count)), dim3(256), 0, s, f, t, count, index, source);
    else hipLaunchKernelGGL((k_periodic<S, false>), dim3(blocks_for(count)), dim3(256), 0, s, f, t, count, index, source);
}

template <class S>
void launch_init(S *f, const Topology &t, const InitDesc &d, const LevelGeom &g, const Codec &cd, bool flip, hipStream_t s)
{
    if (t.N <= 0) return;
    if (flip) hipLaunchKernelGGL((k_init<S, true>), dim3(blocks_for(t.N)), dim3(256), 0, s, f, t, d, g, cd);
    else hipLaunchKernelGGL((k_init<S, false>), dim3(blocks_for(t.N)), dim3(256), 0, s, f, t, d, g, cd);
}

template <class S>
void launch_open_ref_init(const OpenArgs<S> &a, int X, int Y, int Z, bool flip, hipStream_t s)
{
    if (a.count <= 0) return;
    if (flip) hipLaunchKernelGGL((k_open_ref_init<S, true>), dim3(blocks_for(a.count)), dim3(256), 0, s, a, X, Y, Z);
    else hipLaunchKernelGGL((k_open_ref_init<S, false>), dim3(blocks_for(a.count)), dim3(256), 0, s, a, X, Y, Z);
}

template <class S>
void launch_macros(const S *f, const Topology &t, typename StoreTraits<S>::compute *out, const Codec &cd, bool flip,
                   hipStream_t s)
{
    if (t.N <= 0) return;
    if (flip) hipLaunchKernelGGL((k_macros<S, true>), dim3(blocks_for(t.N)), dim3(256), 0, s, f, t, out, cd);
    else hipLaunchKernelGGL((k_macros<S, false>), dim3(blocks_for(t.N)), dim3(256), 0, s, f, t, out, cd);
}

template <class S>
int launch_kinetic_energy(const S *f, const Topology &t, const int *wallMap, const uint32_t *wallData, double *partial,
                          double *out, const Codec &cd, bool flip, hipStream_t s, double divisor, bool accumulate)
{
    const unsigned nb = blocks_for(t.N);
    if (flip) hipLaunchKernelGGL((k_cell_sums<S, true, false>), dim3(nb), dim3(256), 0, s, f, t, wallMap, wallData, partial, cd);
    else hipLaunchKernelGGL((k_cell_sums<S, false, false>), dim3(nb), dim3(256), 0, s, f, t, wallMap, wallData, partial, cd);
    hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(256), 0, s, partial, (int)nb, out, divisor, accumulate ? 1 : 0);
    return (int)nb;
}

template <class S>
int launch_mass(const S *f, const Topology &t, const int *wallMap, const uint32_t *wallData, double *partial, double *out2,
                const Codec &cd, bool flip, hipStream_t s, double divisor, bool accumulate)
{
    const unsigned nb = blocks_for(t.N);
    if (flip) hipLaunchKernelGGL((k_cell_sums<S, true, true>), dim3(nb), dim3(256), 0, s, f, t, wallMap, wallData, partial, cd);
    else hipLaunchKernelGGL((k_cell_sums<S, false, true>), dim3(nb), dim3(256), 0, s, f, t, wallMap, wallData, partial, cd);
    hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(256), 0, s, partial, (int)nb, out2, divisor, accumulate ? 1 : 0);
    hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(256), 0, s, partial + nb, (int)nb, out2 + 1, divisor, accumulate ? 1 : 0);
    return (int)nb;
}

void launch_wall_force(const double *gWall, int W, const uint32_t *wallData, const int *wallIndex, const Topology &t,
                       const LevelGeom &g, const double origin[3], int nBodies, double *partial, hipStream_t s)
{
    if (W <= 0) return;
    hipLaunchKernelGGL(k_wall_force, dim3(blocks_for(W)), dim3(256), 0, s, gWall, W, wallData, wallIndex, t, g, origin[0],
                       origin[1], origin[2], nBodies, partial);
}

template <class T>
void launch_collide_batch(T *f, int n, const Relax<T> &R, hipStream_t s)
{
    if (n <= 0) return;
    hipLaunchKernelGGL((k_collide_batch<T>), dim3(blocks_for(n)), dim3(256), 0, s, f, n, R);
}

template <class S>
void launch_multi_step(const MultiDesc<S> *D, const int *starts, int n, const int (&blocks)[4], hipStream_t s)
{
    // starts: four prefix arrays of n + 1 entries (bulk, wall, open, periodic blocks)
    if (n <= 0) return;
    const unsigned lds = LBM3_BULK_LDS;
    if (blocks[0] > 0) hipLaunchKernelGGL((k_multi_bulk<S>), dim3(blocks[0]), dim3(256), lds, s, D, starts, n);
    if (blocks[1] > 0) hipLaunchKernelGGL((k_multi_wall<S>), dim3(blocks[1]), dim3(256), 0, s, D, starts + (n + 1), n);
    if (blocks[2] > 0) hipLaunchKernelGGL((k_multi_open<S>), dim3(blocks[2]), dim3(256), 0, s, D, starts + 2 * (n + 1), n);
    if (blocks[3] > 0) hipLaunchKernelGGL((k_multi_periodic<S>), dim3(blocks[3]), dim3(256), 0, s, D, starts + 3 * (n + 1), n);
}

// ---- section cut (exportSectionCutPlotGeneral, include/plotter/exportSectionCutPlot.h:11-330)

// image of the section: per pixel {dRho, ux, uy, uz, marker, gridID} before
// conversion; marker 1 (solid) and zeros where no level writes
__global__ __launch_bounds__(256) void k_cut_clear(float *img, long long pixels)
{
    const long long p = (long long)blockIdx.x * 256 + threadIdx.x;
    if (p >= pixels) return;
    float *v = img + 6 * p;
    v[0] = v[1] = v[2] = v[3] = 0.f;
    v[4] = 1.f;
    v[5] = 0.f;
}

// the cells of one level on the section (:101-287): normal index `lower`,
// blended with the next plane by alpha where both are fluid; every
// down-th cell, each painting up x up pixels
template <class S, bool FLIP>
__global__ __launch_bounds__(256) void k_section_cut(const S *f, Topology topo, const int *wallMap, Codec cd, CutLevel d,
                                                     float *img)
{
    using T = typename StoreTraits<S>::compute;
    const int cell = blockIdx.x * 256 + threadIdx.x;
    if (cell >= topo.N) return;
    int nb[8], ijk[3];
    decode(topo, cell, nb, ijk);
    const int normal = d.plane == 0 ? ijk[2] : (d.plane == 1 ? ijk[0] : ijk[1]);
    if (normal != d.lower) return;
    const int hc = d.plane == 0 ? ijk[0] : ijk[2];
    const int vc = d.plane == 2 ? ijk[0] : ijk[1];
    if (hc % d.down != 0 || vc % d.down != 0) return;
    const int ih = (hc / d.down) * d.up, iv = (vc / d.down) * d.up;
    if (ih + d.up <= d.startH || ih >= d.startH + d.pixH || iv + d.up <= d.startV || iv >= d.startV + d.pixV) return;
    int upper = cell;
    int nbu[8];
    for (int k = 0; k < 8; ++k) nbu[k] = nb[k];
    if (d.interpolate) {
        upper = d.plane == 0 ? nb[3] : (d.plane == 1 ? nb[1] : nb[2]);
        if (upper < 0 || upper >= topo.N) return;
        int ijku[3];
        decode(topo, upper, nbu, ijku);
        // only an adjacent centre counts (sparse rows may skip or wrap)
        if (ijku[0] != ijk[0] + (d.plane == 1) || ijku[1] != ijk[1] + (d.plane == 2) || ijku[2] != ijk[2] + (d.plane == 0))
            return;
    }
    const bool lowerSolid = wallMap[cell] == -3, upperSolid = wallMap[upper] == -3;
    const bool selectedSolid = d.alpha <= 0.5f ? lowerSolid : upperSolid;
    float val[4] = {0.f, 0.f, 0.f, 0.f};
    float marker = selectedSolid ? 1.f : 0.f;
    if (!selectedSolid) {
        const bool blend = d.interpolate && !lowerSolid && !upperSolid;
        T v[27], dd, u0, u1, u2;
        load_pre<S, FLIP>(f, topo.N, lowerSolid ? nbu : nb, v, cd);
        moments(v, dd, u0, u1, u2);
        val[0] = float(dd); val[1] = float(u0); val[2] = float(u1); val[3] = float(u2);
        if (blend) {
            load_pre<S, FLIP>(f, topo.N, nbu, v, cd);
            moments(v, dd, u0, u1, u2);
            const float up4[4] = {float(dd), float(u0), float(u1), float(u2)};
            for (int k = 0; k < 4; ++k) val[k] += d.alpha * (up4[k] - val[k]);
        }
        marker = marker < 0.f ? 0.f : (marker > 1.f ? 1.f : marker);
    }
    for (int sv = 0; sv < d.up; ++sv) {
        const int y = iv + sv - d.startV;
        if (y < 0 || y >= d.pixV) continue;
        for (int sh = 0; sh < d.up; ++sh) {
            const int x = ih + sh - d.startH;
            if (x < 0 || x >= d.pixH) continue;
            float *p = img + 6 * ((long long)y * d.pixH + x);
            p[0] = val[0]; p[1] = val[1]; p[2] = val[2]; p[3] = val[3]; p[4] = marker; p[5] = float(d.gridID);
        }
    }
}

// physical units and the plane's velocity order, per pixel (:296-329):
// {p [Pa], u horizontal, u vertical, u normal [m/s], marker, res [mm]}
__global__ __launch_bounds__(256) void k_cut_finish(float *img, long long pixels, CutUnits u)
{
    const long long p = (long long)blockIdx.x * 256 + threadIdx.x;
    if (p >= pixels) return;
    float *v = img + 6 * p;
    const int g = int(v[5]);
    const float res = u.res[g], dt = u.dt[g];
    const float ss = 0.57735026918962576f * (res / 1000.f) / dt;
    const float pr = v[0] * u.rhoPhys * ss * ss;
    const float ux = v[1] * (res / 1000.f) / dt, uy = v[2] * (res / 1000.f) / dt, uz = v[3] * (res / 1000.f) / dt;
    const float h = u.plane == 0 ? ux : uz;
    const float vv = u.plane == 0 ? uy : (u.plane == 1 ? uy : ux);
    const float n = u.plane == 0 ? uz : (u.plane == 1 ? ux : uy);
    v[0] = pr; v[1] = h; v[2] = vv; v[3] = n; v[5] = res;
}

// per cell of a level, the nine fields of exportVTK (include/exportVTK.h:64-112):
// x, y, z [mm], wall, ux, uy, uz [m/s], pressure [Pa], interface overlap
template <class S, bool FLIP>
__global__ __launch_bounds__(256) void k_vtk_fields(const S *f, Topology topo, const int *wallMap, const uint32_t *wallData,
                                                    Codec cd, LevelGeom g, float rhoPhys, float *out)
{
    using T = typename StoreTraits<S>::compute;
    const int cell = blockIdx.x * 256 + threadIdx.x;
    if (cell >= topo.N) return;
    int nb[8], ijk[3];
    decode(topo, cell, nb, ijk);
    const int wm = wallMap[cell];
    const bool wall = wm == -3;
    float v[9];
    v[0] = float(ijk[0]) * g.res + g.ox;
    v[1] = float(ijk[1]) * g.res + g.oy;
    v[2] = float(ijk[2]) * g.res + g.oz;
    v[3] = wall ? 1.f : 0.f;
    v[4] = v[5] = v[6] = v[7] = v[8] = 0.f;
    if (!wall) {
        v[8] = (wm == -2 || (wm >= 0 && (wallData[wm] & 1u))) ? 1.f : 0.f;
        T q[27], d, u0, u1, u2;
        load_pre<S, FLIP>(f, topo.N, nb, q, cd);
        moments(q, d, u0, u1, u2);
        const float ss = 0.57735026918962576f * (g.res / 1000.f) / g.dtPhys;
        v[4] = float(u0) * (g.res / 1000.f) / g.dtPhys;
        v[5] = float(u1) * (g.res / 1000.f) / g.dtPhys;
        v[6] = float(u2) * (g.res / 1000.f) / g.dtPhys;
        v[7] = float(d) * rhoPhys * ss * ss;
    }
    for (int k = 0; k < 9; ++k) out[(size_t)k * topo.N + cell] = v[k];
}

template <class S> void launch_vtk_fields(const S *f, const Topology &t, const int *wallMap, const uint32_t *wallData,
                                          const Codec &cd, bool flip, const LevelGeom &g, float rhoPhys, float *out,
                                          hipStream_t s)
{
    if (t.N <= 0) return;
    const unsigned b = (unsigned)((t.N + 255) / 256);
    if (flip) k_vtk_fields<S, true><<<b, 256, 0, s>>>(f, t, wallMap, wallData, cd, g, rhoPhys, out);
    else k_vtk_fields<S, false><<<b, 256, 0, s>>>(f, t, wallMap, wallData, cd, g, rhoPhys, out);
}

template <class S> void launch_section_cut(const S *f, const Topology &t, const int *wallMap, const Codec &cd, bool flip,
                                           const CutLevel &d, float *img, hipStream_t s)
{
    if (t.N <= 0) return;
    const unsigned g = (unsigned)((t.N + 255) / 256);
    if (flip) k_section_cut<S, true><<<g, 256, 0, s>>>(f, t, wallMap, cd, d, img);
    else k_section_cut<S, false><<<g, 256, 0, s>>>(f, t, wallMap, cd, d, img);
}

void launch_cut_clear(float *img, long long pixels, hipStream_t s)
{
    if (pixels > 0) k_cut_clear<<<(unsigned)((pixels + 255) / 256), 256, 0, s>>>(img, pixels);
}

void launch_cut_finish(float *img, long long pixels, const CutUnits &u, hipStream_t s)
{
    if (pixels > 0) k_cut_finish<<<(unsigned)((pixels + 255) / 256), 256, 0, s>>>(img, pixels, u);
}

#define LBM3_INSTANTIATE(S)                                                                                          \
    template void launch_vtk_fields<S>(const S *, const Topology &, const int *, const uint32_t *, const Codec &, bool,  \
                                       const LevelGeom &, float, float *, hipStream_t);                                \
    template void launch_section_cut<S>(const S *, const Topology &, const int *, const Codec &, bool, const CutLevel &, \
                                        float *, hipStream_t);                                                       \
    template void launch_collide_stream<S>(const StepArgs<S> &, bool, hipStream_t);                                 \
    template void launch_collide_wall<S>(const StepArgs<S> &, bool, hipStream_t);                                   \
    template void launch_open_cells<S>(const StepArgs<S> &, bool, hipStream_t);                                     \
    template void launch_rotor_cells<S>(const StepArgs<S> &, bool, hipStream_t);                                    \
    template void launch_multi_step<S>(const MultiDesc<S> *, const int *, int, const int (&)[4], hipStream_t);       \
    template void launch_open<S>(const OpenArgs<S> &, bool, hipStream_t);                                           \
    template void launch_periodic<S>(S *, const Topology &, int, const int *, const int *, bool, hipStream_t);        \
    template void launch_init<S>(S *, const Topology &, const InitDesc &, const LevelGeom &, const Codec &, bool,     \
                                 hipStream_t);                                                                    \
    template void launch_open_ref_init<S>(const OpenArgs<S> &, int, int, int, bool, hipStream_t);                   \
    template void launch_macros<S>(const S *, const Topology &, typename StoreTraits<S>::compute *, const Codec &,    \
                                   bool, hipStream_t);                                                            \
    template int launch_kinetic_energy<S>(const S *, const Topology &, const int *, const uint32_t *, double *,       \
                                          double *, const Codec &, bool, hipStream_t, double, bool);              \
    template int launch_mass<S>(const S *, const Topology &, const int *, const uint32_t *, double *, double *,       \
                                const Codec &, bool, hipStream_t, double, bool);
LBM3_INSTANTIATE(float)
LBM3_INSTANTIATE(double)
LBM3_INSTANTIATE(__half)
LBM3_INSTANTIATE(__hip_bfloat16)
LBM3_INSTANTIATE(f16c)
template void launch_collide_batch<float>(float *, int, const Relax<float> &, hipStream_t);
template void launch_collide_batch<double>(double *, int, const Relax<double> &, hipStream_t);

} // namespace lbm3

// ---------------------------------------------------------------------------
// plain streaming copy: the bandwidth probe used to calibrate HBM counters
// ---------------------------------------------------------------------------
namespace lbm3 {
typedef float f4v __attribute__((ext_vector_type(4)));
__global__ __launch_bounds__(256) void k_copy(const float4 *__restrict__ a4, float4 *__restrict__ b4, long long n4,
                                              const float *__restrict__ at, float *__restrict__ bt, int tail)
{
    // four independent 16-byte loads in flight per thread before the stores,
    // non-temporal: a plain stream with no reuse
    const f4v *__restrict__ a = reinterpret_cast<const f4v *>(a4);
    f4v *__restrict__ b = reinterpret_cast<f4v *>(b4);
    const long long stride = (long long)gridDim.x * 256;
    long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    for (; i + 3 * stride < n4; i += 4 * stride) {
        const f4v v0 = __builtin_nontemporal_load(a + i), v1 = __builtin_nontemporal_load(a + i + stride);
        const f4v v2 = __builtin_nontemporal_load(a + i + 2 * stride), v3 = __builtin_nontemporal_load(a + i + 3 * stride);
        __builtin_nontemporal_store(v0, b + i);
        __builtin_nontemporal_store(v1, b + i + stride);
        __builtin_nontemporal_store(v2, b + i + 2 * stride);
        __builtin_nontemporal_store(v3, b + i + 3 * stride);
    }
    for (; i < n4; i += stride) b[i] = a[i];
    if (blockIdx.x == 0 && (int)threadIdx.x < tail) bt[threadIdx.x] = at[threadIdx.x];
}

// 16-byte vector copy of `bytes` (the buffers come from hipMalloc, 256-byte aligned)
// tiled level array <-> plain planes q * N + cell (the reference's layout), elements of E bytes
template <class E> __global__ __launch_bounds__(256) void k_relayout(const E *src, E *dst, int N, int toPlanes)
{
    const long long x = blockIdx.x * 256LL + threadIdx.x;
    if (x >= 27LL * N) return;
    const int q = int(x / N), c = int(x % N);
    if (toPlanes) dst[x] = src[pop_index(q, c)];
    else dst[pop_index(q, c)] = src[x];
}

void launch_relayout(const void *src, void *dst, int N, int elemBytes, bool toPlanes, hipStream_t s)
{
    if (N <= 0) return;
    const unsigned g = blocks_for(27LL * N);
    if (elemBytes == 2) hipLaunchKernelGGL(k_relayout<uint16_t>, dim3(g), dim3(256), 0, s, (const uint16_t *)src, (uint16_t *)dst, N, int(toPlanes));
    else if (elemBytes == 4) hipLaunchKernelGGL(k_relayout<uint32_t>, dim3(g), dim3(256), 0, s, (const uint32_t *)src, (uint32_t *)dst, N, int(toPlanes));
    else hipLaunchKernelGGL(k_relayout<uint64_t>, dim3(g), dim3(256), 0, s, (const uint64_t *)src, (uint64_t *)dst, N, int(toPlanes));
}

void launch_copy(const void *a, void *b, long long bytes, hipStream_t s)
{
    const long long n4 = bytes / 16;
    const int tail = (int)((bytes - 16 * n4) / 4);
    hipLaunchKernelGGL(k_copy, dim3(16384), dim3(256), 0, s, (const float4 *)a, (float4 *)b, n4,
                       (const float *)a + 4 * n4, (float *)b + 4 * n4, tail);
}
} // namespace lbm3
