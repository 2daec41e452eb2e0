// Multilevel coarse <-> fine interface transfer for gfx950 (reference
// include/updateInterface.h:155-890, or with LINEAR the affine variant of
// include/INTERFACE_VERSIONS_BACKUP/updateInterfaceLINEAR.h), run after every
// pair of fine steps.
//
//  k_fine_to_coarse  one coarse cell per group of 8 lanes: lane n loads fine
//                    stencil corner n, the group shares the corner values in
//                    LDS, every lane builds the interpolant and writes 1/8 of
//                    the coarse cell's 27 populations
//  k_coarse_to_fine  one 2x2x2 coarse block per group of 8 lanes: lane n loads
//                    coarse corner n, shares it, and writes fine target n
//  k_leftovers       fine cells outside every full block: centre coarse cell
//                    plus three signed face neighbours (one lane each)
//
// Populations are read and written in the pre-collision slots of the current
// esotwist parity of their level.
#include <hip/hip_runtime.h>
#include "interface.hpp"
#include "sim.hpp"

namespace lbm3 {

namespace {

template <class S, bool FLIP>
__device__ __forceinline__ void load_cell(const S *f, const FastTopo &t, int cell, typename StoreTraits<S>::compute (&v)[27],
                                          int (&nb)[8], const Codec &cd)
{
    decode_fast(t, cell, t.cellRow[cell].y, nb);
#pragma unroll
    for (int q = 0; q < 27; ++q) v[q] = StoreTraits<S>::load(f + pop_index(pre_plane(q, FLIP), nb[pre_nb(q)]), q, cd);
}

// the stencil whose lowest corner is `c0`: self, i+, j+, ij+, k+, ik+, jk+, ijk+
__device__ __forceinline__ int stencil_corner(const FastTopo &t, int c0, int n)
{
    int nb[8];
    decode_fast(t, c0, t.cellRow[c0].y, nb);
    // decode order: self, i+, j+, k+, ij+, ik+, jk+, ijk+
    const int map[8] = {0, 1, 2, 4, 3, 5, 6, 7};
    return nb[map[n]];
}

// the eight neighbour cells of table entry e (two int4)
__device__ __forceinline__ void table_nb(const int4 *t, size_t e, int (&nb)[8])
{
    const int4 a = t[2 * e], b = t[2 * e + 1];
    nb[0] = a.x; nb[1] = a.y; nb[2] = a.z; nb[3] = a.w;
    nb[4] = b.x; nb[5] = b.y; nb[6] = b.z; nb[7] = b.w;
}

template <class S, bool FLIP>
__device__ __forceinline__ void load_nb(const S *f, const int (&nb)[8], typename StoreTraits<S>::compute (&v)[27],
                                        const Codec &cd)
{
#pragma unroll
    for (int q = 0; q < 27; ++q) v[q] = StoreTraits<S>::load(f + pop_index(pre_plane(q, FLIP), nb[pre_nb(q)]), q, cd);
}

__device__ __forceinline__ void put_nb(int4 *t, size_t e, const int (&nb)[8])
{
    t[2 * e] = make_int4(nb[0], nb[1], nb[2], nb[3]);
    t[2 * e + 1] = make_int4(nb[4], nb[5], nb[6], nb[7]);
}

// neighbour tables of one transfer list: per item, the stencil corners it reads
// and the cell(s) it writes (F->C: the coarse cell; C->F: the 8 fine targets)
__global__ __launch_bounds__(256) void k_if_tables(FastTopo src, FastTopo dst, int count, const int *stencilBase,
                                                   const int *target, int c2f, int4 *srcNb, int4 *dstNb)
{
    const int gid = blockIdx.x * 256 + threadIdx.x;
    const int item = gid >> 3, n = gid & 7;
    if (item >= count) return;
    int nb[8];
    const int cs = stencil_corner(src, stencilBase[item], n);
    decode_fast(src, cs, src.cellRow[cs].y, nb);
    put_nb(srcNb, 8 * (size_t)item + n, nb);
    if (c2f) {
        const int ct = stencil_corner(dst, target[item], n);
        decode_fast(dst, ct, dst.cellRow[ct].y, nb);
        put_nb(dstNb, 8 * (size_t)item + n, nb);
    } else if (n == 0) {
        const int ct = target[item];
        decode_fast(dst, ct, dst.cellRow[ct].y, nb);
        put_nb(dstNb, item, nb);
    }
}

template <class T> struct IfArgs {
    FastTopo coarse, fine;
    const int4 *srcNb, *dstNb;     // neighbour tables of the transfer (InterfaceLists)
    void *fc, *ff;                 // populations of the two levels (storage type)
    T w1c, w1f;                    // omega1 of the two levels
    int count;
    int loCount;                   // leftovers of a combined C->F and leftover launch
    const int *index, *child;      // interface lists of the coarse level
    const int *loIndex, *loParent, *loI, *loJ, *loK;
    const int *wallMapC;
    const uint32_t *wallDataC;
    const LocalBC *wallBCC;        // [32] local BC of the coarse level by wall id
    const float *wallVelC;         // [3][WC] per-cell wall velocity of the coarse level (velMode 3) or null
    int WC;
    LevelGeom geomC;
    Topology topoC;
    Codec cd;
};

// F->C: the coarse cell at the centre of a 2x2x2 fine stencil
// (updateInterface.h:155-360)
template <class S, bool FLIPC, bool FLIPF, bool LINEAR>
__global__ __launch_bounds__(256) void k_fine_to_coarse(IfArgs<typename StoreTraits<S>::compute> a)
{
    using T = typename StoreTraits<S>::compute;
    __shared__ CellValues<T> lds[256];
    const int gid = blockIdx.x * 256 + threadIdx.x;
    const int item = gid >> 3, n = gid & 7;
    const bool active = item < a.count;
    if (active) {
        int nb[8];
        table_nb(a.srcNb, 8 * (size_t)item + n, nb);
        T v[27];
        load_nb<S, FLIPF>(reinterpret_cast<const S *>(a.ff), nb, v, a.cd);
        lds[threadIdx.x] = cell_values(v, a.w1f);
    }
    __syncthreads();
    if (!active) return;
    const int base = threadIdx.x & ~7;
    CellValues<T> cv[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) cv[k] = lds[base + k];
    T f[27];
    if constexpr (LINEAR) {
        linear_fine_to_coarse(f, cv, a.w1c);
    } else {
        const Interpolant<T> P = interpolant(cv);
        T d, u[3], g[3][3];
        evaluate(P, T(0), T(0), T(0), d, u, g);
        // density: the plain mean of the eight fine cells (updateInterface.h:235-237)
        T mean = 0;
#pragma unroll
        for (int k = 0; k < 8; ++k) mean += cv[k].d;
        d = mean * T(0.125);
        reconstruct(f, d, u, g[0][1] + g[1][0], g[1][2] + g[2][1], g[0][2] + g[2][0], g[0][0] - g[1][1], g[0][0] - g[2][2],
                    T(2) / a.w1c);
    }
    int nb[8];
    table_nb(a.dstNb, item, nb);
    S *fc = reinterpret_cast<S *>(a.fc);
    // lane n writes populations n, n+8, n+16, n+24
#pragma unroll
    for (int q = 0; q < 27; ++q)
        if ((q & 7) == n) StoreTraits<S>::store(fc + pop_index(pre_plane(q, FLIPC), nb[pre_nb(q)]), f[q], q, a.cd);
}

// C->F: the eight fine cells inside a 2x2x2 coarse block, at +-1/4
// (updateInterface.h:362-687); `blk` the block's index among the C->F blocks
template <class S, bool FLIPC, bool FLIPF, bool LINEAR>
__device__ __forceinline__ void c2f_block(const IfArgs<typename StoreTraits<S>::compute> &a, int blk)
{
    using T = typename StoreTraits<S>::compute;
    __shared__ CellValues<T> lds[256];
    const int gid = blk * 256 + threadIdx.x;
    const int item = gid >> 3, n = gid & 7;
    const bool active = item < a.count;
    if (active) {
        int nb[8];
        table_nb(a.srcNb, 8 * (size_t)item + n, nb);
        T v[27];
        load_nb<S, FLIPC>(reinterpret_cast<const S *>(a.fc), nb, v, a.cd);
        lds[threadIdx.x] = cell_values(v, a.w1c);
    }
    __syncthreads();
    if (!active) return;
    const int base = threadIdx.x & ~7;
    CellValues<T> cv[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) cv[k] = lds[base + k];
    const T x = (n & 1) ? T(0.25) : T(-0.25), y = (n & 2) ? T(0.25) : T(-0.25), z = (n & 4) ? T(0.25) : T(-0.25);
    T f[27];
    if constexpr (LINEAR) {
        linear_coarse_to_fine(f, cv, a.w1f, x, y, z);
    } else {
        const Interpolant<T> P = interpolant(cv);
        T d, u[3], g[3][3];
        evaluate(P, x, y, z, d, u, g);
        reconstruct(f, d, u, g[0][1] + g[1][0], g[1][2] + g[2][1], g[0][2] + g[2][0], g[0][0] - g[1][1], g[0][0] - g[2][2],
                    T(0.5) / a.w1f);
    }
    int nb[8];
    table_nb(a.dstNb, 8 * (size_t)item + n, nb);
    S *ff = reinterpret_cast<S *>(a.ff);
#pragma unroll
    for (int q = 0; q < 27; ++q) StoreTraits<S>::store(ff + pop_index(pre_plane(q, FLIPF), nb[pre_nb(q)]), f[q], q, a.cd);
}

template <class S, bool FLIPC, bool FLIPF, bool LINEAR>
__global__ __launch_bounds__(256) void k_coarse_to_fine(IfArgs<typename StoreTraits<S>::compute> a)
{
    c2f_block<S, FLIPC, FLIPF, LINEAR>(a, blockIdx.x);
}

// fine cells without a full coarse block (updateInterface.h:689-884), `count`
// of them; `blk` the block's index among the leftover blocks
template <class S, bool FLIPC, bool FLIPF>
__device__ __forceinline__ void leftover_block(const IfArgs<typename StoreTraits<S>::compute> &a, int blk, int count)
{
    using T = typename StoreTraits<S>::compute;
    const int x = blk * 256 + threadIdx.x;
    if (x >= count) return;
    const S *fc = reinterpret_cast<const S *>(a.fc);
    const int centre = a.loParent[x];
    T v[27];
    int nb[8];
    load_cell<S, FLIPC>(fc, a.coarse, centre, v, nb, a.cd);
    const CellValues<T> c0 = cell_values(v, a.w1c);
    // wall velocity of the centre cell: the local BC of its wall id (the cell's own velocity if none is set)
    const int wm = a.wallMapC[centre];
    const uint32_t wd = wm >= 0 ? a.wallDataC[wm] : 0u;
    T uw[3] = {c0.u[0], c0.u[1], c0.u[2]};
    if (wm >= 0) {
        const LocalBC &b = a.wallBCC[(wd >> 27) & 31u];
        if (b.velMode == 1) { uw[0] = T(b.ux); uw[1] = T(b.uy); uw[2] = T(b.uz); }
        else if (b.velMode == 2) {
            int nbx[8], ijk[3];
            decode(a.topoC, centre, nbx, ijk);
            const float px = float(ijk[0]) * a.geomC.res + a.geomC.ox - b.origin[0];
            const float py = float(ijk[1]) * a.geomC.res + a.geomC.oy - b.origin[1];
            const float pz = float(ijk[2]) * a.geomC.res + a.geomC.oz - b.origin[2];
            const float s = a.geomC.dtPhys / a.geomC.res;
            uw[0] = T((b.omega[1] * pz - b.omega[2] * py) * s);
            uw[1] = T((b.omega[2] * px - b.omega[0] * pz) * s);
            uw[2] = T((b.omega[0] * py - b.omega[1] * px) * s);
        } else if (b.velMode == 3) {
            uw[0] = T(a.wallVelC[wm]);
            uw[1] = T(a.wallVelC[(size_t)a.WC + wm]);
            uw[2] = T(a.wallVelC[2 * (size_t)a.WC + wm]);
        }
    }
    CellValues<T> r = c0;
    const int nbr[3] = {a.loI[x], a.loJ[x], a.loK[x]};
    const T q = T(0.25);
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const bool opposite = nbr[k] < 0;
        const int cell = opposite ? -nbr[k] - 1 : nbr[k];
        load_cell<S, FLIPC>(fc, a.coarse, cell, v, nb, a.cd);
        const CellValues<T> c = cell_values(v, a.w1c);
        // a neighbour on the + side adds the linear slope, one on the - side removes it
        const T sgn = opposite ? T(-1) : T(1);
        r.d += sgn * q * (c.d - c0.d);
        r.sxy += sgn * q * (c.sxy - c0.sxy); r.syz += sgn * q * (c.syz - c0.syz); r.sxz += sgn * q * (c.sxz - c0.sxz);
        r.nxy += sgn * q * (c.nxy - c0.nxy); r.nxz += sgn * q * (c.nxz - c0.nxz);
        const int dir = k == 0 ? POO : (k == 1 ? OPO : OOP);
        const bool wall = opposite && (((wd >> dir) & 1u) || ((wd >> inv(dir)) & 1u));
        if (!wall) {
#pragma unroll
            for (int m = 0; m < 3; ++m) r.u[m] += sgn * q * (c.u[m] - c0.u[m]);
        } else {
            // quadratic through the wall (1/2 away), the centre and the opposite cell
#pragma unroll
            for (int m = 0; m < 3; ++m) r.u[m] += (T(1) / T(24)) * (T(10) * uw[m] - T(9) * c0.u[m] - c.u[m]);
        }
    }
    T f[27];
    reconstruct(f, r.d, r.u, r.sxy, r.syz, r.sxz, r.nxy, r.nxz, T(0.5) / a.w1f);
    const int cell = a.loIndex[x];
    decode_fast(a.fine, cell, a.fine.cellRow[cell].y, nb);
    S *ff = reinterpret_cast<S *>(a.ff);
#pragma unroll
    for (int m = 0; m < 27; ++m) StoreTraits<S>::store(ff + pop_index(pre_plane(m, FLIPF), nb[pre_nb(m)]), f[m], m, a.cd);
}

template <class S, bool FLIPC, bool FLIPF>
__global__ __launch_bounds__(256) void k_leftovers(IfArgs<typename StoreTraits<S>::compute> a)
{
    leftover_block<S, FLIPC, FLIPF>(a, blockIdx.x, a.count);
}

// C->F and the leftovers in one launch: they read the coarse level and write
// disjoint fine cells, so nothing orders them; the first `c2fBlocks` blocks
// take the C->F items, the others the leftovers
template <class S, bool FLIPC, bool FLIPF, bool LINEAR>
__global__ __launch_bounds__(256) void k_c2f_leftovers(IfArgs<typename StoreTraits<S>::compute> a, int c2fBlocks)
{
    if ((int)blockIdx.x < c2fBlocks) c2f_block<S, FLIPC, FLIPF, LINEAR>(a, blockIdx.x);
    else leftover_block<S, FLIPC, FLIPF>(a, blockIdx.x - c2fBlocks, a.loCount);
}

inline unsigned groups_of_8(long long items) { return (unsigned)((8 * items + 255) / 256); }

template <class S, bool FC, bool FF> void launch_all(const IfArgs<typename StoreTraits<S>::compute> &base, const Level &c,
                                                      bool linear, hipStream_t s)
{
    IfArgs<typename StoreTraits<S>::compute> a = base;
    if (c.F2C.interfaceCount > 0) {
        a.count = c.F2C.interfaceCount; a.index = c.F2C.index.p; a.child = c.F2C.childMap.p;
        a.srcNb = c.F2C.srcNb.p; a.dstNb = c.F2C.dstNb.p;
        if (linear) hipLaunchKernelGGL((k_fine_to_coarse<S, FC, FF, true>), dim3(groups_of_8(a.count)), dim3(256), 0, s, a);
        else hipLaunchKernelGGL((k_fine_to_coarse<S, FC, FF, false>), dim3(groups_of_8(a.count)), dim3(256), 0, s, a);
    }
    const bool c2f = c.C2F.interfaceCount > 0, lo = c.C2F.leftoverCount > 0;
    if (c2f) {
        a.count = c.C2F.interfaceCount; a.index = c.C2F.index.p; a.child = c.C2F.childMap.p;
        a.srcNb = c.C2F.srcNb.p; a.dstNb = c.C2F.dstNb.p;
    }
    if (c2f && lo) {
        a.loCount = c.C2F.leftoverCount;
        const int nb = (int)groups_of_8(a.count);
        const dim3 g((unsigned)nb + (unsigned)((a.loCount + 255) / 256));
        if (linear) hipLaunchKernelGGL((k_c2f_leftovers<S, FC, FF, true>), g, dim3(256), 0, s, a, nb);
        else hipLaunchKernelGGL((k_c2f_leftovers<S, FC, FF, false>), g, dim3(256), 0, s, a, nb);
    } else if (c2f) {
        if (linear) hipLaunchKernelGGL((k_coarse_to_fine<S, FC, FF, true>), dim3(groups_of_8(a.count)), dim3(256), 0, s, a);
        else hipLaunchKernelGGL((k_coarse_to_fine<S, FC, FF, false>), dim3(groups_of_8(a.count)), dim3(256), 0, s, a);
    } else if (lo) {
        a.count = c.C2F.leftoverCount;
        hipLaunchKernelGGL((k_leftovers<S, FC, FF>), dim3((a.count + 255) / 256), dim3(256), 0, s, a);
    }
}

// derive the neighbour tables of a transfer list (its first use after a build)
int build_if_tables(InterfaceLists &L, const FastTopo &src, const FastTopo &dst, const int *stencilBase, const int *target,
                    bool c2f, hipStream_t s)
{
    if (L.nbReady || L.interfaceCount <= 0) return 0;
    const long long n = L.interfaceCount;
    if (L.srcNb.alloc(16 * n) != hipSuccess || L.dstNb.alloc((c2f ? 16 : 2) * n) != hipSuccess) return LBM3_E_HIP;
    hipLaunchKernelGGL(k_if_tables, dim3(groups_of_8(n)), dim3(256), 0, s, src, dst, L.interfaceCount, stencilBase, target,
                       c2f ? 1 : 0, L.srcNb.p, L.dstNb.p);
    if (hipGetLastError() != hipSuccess) return LBM3_E_HIP;
    L.nbReady = true;
    return 0;
}

template <class S> void interface_typed(Sim &s, int coarse)
{
    using T = typename StoreTraits<S>::compute;
    Level &c = s.L[coarse], &f = s.L[coarse + 1];
    IfArgs<T> a{};
    a.coarse = FastTopo{c.N, c.cellRow.p, c.rowTab.p};
    a.fine = FastTopo{f.N, f.cellRow.p, f.rowTab.p};

    a.fc = c.f.p;
    a.ff = f.f.p;
    // omega1 = 1 / (3 nu + 1/2) of each level, in the compute precision
    a.w1c = T(1) / (T(3) * T(c.info.nu) + T(0.5));
    a.w1f = T(1) / (T(3) * T(f.info.nu) + T(0.5));
    a.loIndex = c.C2F.leftoverIndex.p; a.loParent = c.C2F.leftoverParent.p;
    a.loI = c.C2F.leftoverNbrI.p; a.loJ = c.C2F.leftoverNbrJ.p; a.loK = c.C2F.leftoverNbrK.p;
    a.wallMapC = c.wallMap.p;
    a.wallDataC = c.wallData.p;
    a.wallBCC = c.wallBC.p;
    a.wallVelC = c.wallVel.p;
    a.WC = c.W;
    a.geomC = c.geom;
    a.topoC = Topology{c.N, c.shifter.p, c.ri.p, c.rj.p, c.rk.p, c.rjp.p, c.rkp.p, c.rjkp.p};
    a.cd = s.codec;
    const bool fc = c.info.esotwistFlipper != 0, ff = f.info.esotwistFlipper != 0;
    const bool lin = s.C.interfaceScheme == LBM3_INTERFACE_LINEAR;
    if (fc) { if (ff) launch_all<S, true, true>(a, c, lin, s.stream); else launch_all<S, true, false>(a, c, lin, s.stream); }
    else { if (ff) launch_all<S, false, true>(a, c, lin, s.stream); else launch_all<S, false, false>(a, c, lin, s.stream); }
}

} // namespace

// updateInterface (updateInterface.h:886-890): F->C, then C->F on the coarse
// values it has just written, then the leftovers
void update_interface(Sim &s, int coarse)
{
    switch (s.storage) {
    case LBM3_STORE_F64: interface_typed<double>(s, coarse); break;
    case LBM3_STORE_F16: interface_typed<__half>(s, coarse); break;
    case LBM3_STORE_BF16: interface_typed<__hip_bfloat16>(s, coarse); break;
    case LBM3_STORE_F16C: interface_typed<f16c>(s, coarse); break;
    default: interface_typed<float>(s, coarse); break;
    }
}

// the neighbour tables of the transfers between `coarse` and the next level:
// F->C reads fine stencils (from childMap) and writes the coarse cell (index),
// C->F reads coarse stencils (from index) and writes the fine ones (childMap)
int prepare_interface(Sim &s, int coarse)
{
    Level &c = s.L[coarse], &f = s.L[coarse + 1];
    const FastTopo ct{c.N, c.cellRow.p, c.rowTab.p}, ft{f.N, f.cellRow.p, f.rowTab.p};
    if (const int rc = build_if_tables(c.F2C, ft, ct, c.F2C.childMap.p, c.F2C.index.p, false, s.stream)) return rc;
    return build_if_tables(c.C2F, ct, ft, c.C2F.index.p, c.C2F.childMap.p, true, s.stream);
}

} // namespace lbm3
