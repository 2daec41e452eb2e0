// GPU grid builder for gfx950: STL binning, z-ray voxelizer, markers, cell
// lists, neighbour rows, IBB links and the compression into the per-step arrays.
//
// Reference pipeline (same stages, same results): gridBuilderFunctions.h
// initializeGridInfo :9, buildIJKFull :299, deleteExcessCells :402,
// buildWallMarkers :590, gridBuilderToGrid :1032, buildGrids :1577;
// voxelizerFunctions.h voxelizeSTL :67, sumRayMaps :537; markerFunctions.h;
// NBRFunctions.h; interpolatedBouncebackFunctions.h :181-702; STLFunctions.h :5-128.
//
// Data layout: every level keeps a dense index map over its full box
// (cell index or -1).  Compacting a dense occupancy in linear order yields the
// cells already sorted by (k, j, i), so the reference's sort and all of its
// binary searches (parent map, periodic and open-boundary sources) become
// O(1) lookups.  Floating-point arithmetic in this file is never contracted,
// and fp32 division/sqrt are correctly rounded, so the voxelizer snap, the hit
// heights and the link lengths are bit-identical to the sequential restatement.
#pragma clang fp contract(off)

#include <hipcub/hipcub.hpp>
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>
#include "sim.hpp"

namespace lbm3 {
namespace {

#define BCHK(expr)                                    \
    do {                                              \
        const int rc_ = check_hip((expr), #expr);     \
        if (rc_) return rc_;                          \
    } while (0)

template <class F> __global__ __launch_bounds__(256) void k_for(long long n, F f)
{
    const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    if (i < n) f(i);
}

template <class F> void for_n(long long n, hipStream_t s, F f)
{
    if (n <= 0) return;
    hipLaunchKernelGGL((k_for<F>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n, f);
}

// ---------------------------------------------------------------------------
// scans (hipcub)
// ---------------------------------------------------------------------------
struct Scratch {
    DBuf<unsigned char> tmp;
    int ensure(size_t bytes)
    {
        if ((size_t)tmp.n < bytes) return check_hip(tmp.alloc((long long)bytes), "scan scratch");
        return 0;
    }
};

template <class T> int exclusive_sum(Scratch &sc, const T *in, T *out, long long n, hipStream_t s)
{
    size_t bytes = 0;
    BCHK(hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, in, out, (int)n, s));
    if (sc.ensure(bytes)) return LBM3_E_HIP;
    BCHK(hipcub::DeviceScan::ExclusiveSum(sc.tmp.p, bytes, in, out, (int)n, s));
    return 0;
}

struct MaxOp {
    __host__ __device__ int operator()(int a, int b) const { return a > b ? a : b; }
};

int inclusive_max(Scratch &sc, const int *in, int *out, long long n, hipStream_t s)
{
    size_t bytes = 0;
    BCHK(hipcub::DeviceScan::InclusiveScan(nullptr, bytes, in, out, MaxOp(), (int)n, s));
    if (sc.ensure(bytes)) return LBM3_E_HIP;
    BCHK(hipcub::DeviceScan::InclusiveScan(sc.tmp.p, bytes, in, out, MaxOp(), (int)n, s));
    return 0;
}

template <class T> T read_one(const T *p, hipStream_t s)
{
    T v{};
    (void)hipMemcpyAsync(&v, p, sizeof(T), hipMemcpyDeviceToHost, s);
    (void)hipStreamSynchronize(s);
    return v;
}

// positions of the set flags, in order; returns the count
int compact(Scratch &sc, const uint8_t *flags, long long n, DBuf<int> &out, hipStream_t s)
{
    DBuf<int> pos;
    BCHK(pos.alloc(n + 1));
    const uint8_t *fl = flags;
    int *p = pos.p;
    for_n(n + 1, s, [=] __device__(long long i) { p[i] = (i < n && fl[i]) ? 1 : 0; });
    DBuf<int> sc2;
    BCHK(sc2.alloc(n + 1));
    if (exclusive_sum(sc, pos.p, sc2.p, n + 1, s)) return -1;
    const int count = read_one(sc2.p + n, s);
    BCHK(out.alloc(count));
    int *o = out.p;
    const int *x = sc2.p;
    for_n(n, s, [=] __device__(long long i) {
        if (fl[i]) o[x[i]] = (int)i;
    });
    return count;
}

// ---------------------------------------------------------------------------
// builder state
// ---------------------------------------------------------------------------
struct DevSTL {
    int T = 0;
    DBuf<float> v[9];
    lbm3_bounds B{};
    float binSize = 0, oxBin = 0, oyBin = 0, ozBin = 0;
    int bx = 0, by = 0, bz = 0;
    DBuf<int> firstInBin, binArray;
    std::vector<int> edges;    // shared-edge counts, 3 per kept triangle (AB, BC, CA)
};

struct RayMap {
    int nRays = 0;
    long long total = 0;
    DBuf<long long> scan;
    DBuf<int> map;
};

struct BLevel {
    lbm3_info I{};
    lbm3_info Sk{};
    int N = 0;
    DBuf<int> i, j, k, jPlus, kPlus, jMinus, kMinus, dense, parentMap;
    DBuf<uint8_t> isGeo, keep, wall, refine, deep, f2c, c2f, parentIf, ifOverlap, parentF2C, needCoarse, skKeep;
    DBuf<int> wallIDFull, wallIDc, wallList;
    int W = 0;
    DBuf<uint8_t> linkExist, linkPierce;
    DBuf<float> linkLen;
    RayMap total;
    std::vector<RayMap> maps;
};

struct RayView {
    const long long *scan;
    const int *map;
};

// ---------------------------------------------------------------------------
// initializeGridInfo (host, float arithmetic as the reference)
// ---------------------------------------------------------------------------
void init_info(const lbm3_case &C, std::vector<BLevel> &G, lbm3_bounds B)
{
    for (size_t L = 0; L < G.size(); ++L) {
        lbm3_info &I = G[L].I;
        memset(&I, 0, sizeof(I));
        I.useRotors = 1;
        if (L == 0) {
            I.res = C.resGlobal;
            lbm3_info &S = G[0].Sk;
            memset(&S, 0, sizeof(S));
            S.gridID = -1.f;
            S.res = I.res * 2.f;
            S.cellCountX = (int)((B.xMax - B.xMin) / S.res);
            S.cellCountY = (int)((B.yMax - B.yMin) / S.res);
            S.cellCountZ = (int)((B.zMax - B.zMin) / S.res);
            S.cellCount = S.cellCountX * S.cellCountY * S.cellCountZ;
            S.ox = B.xMin + ((B.xMax - B.xMin) - ((float)S.cellCountX * S.res) + S.res) / 2.0f;
            S.oy = B.yMin + ((B.yMax - B.yMin) - ((float)S.cellCountY * S.res) + S.res) / 2.0f;
            S.oz = B.zMin + ((B.zMax - B.zMin) - ((float)S.cellCountZ * S.res) + S.res) / 2.0f;
            I.cellCountX = S.cellCountX * 2; I.cellCountY = S.cellCountY * 2; I.cellCountZ = S.cellCountZ * 2;
            I.ox = S.ox - I.res * 0.5f; I.oy = S.oy - I.res * 0.5f; I.oz = S.oz - I.res * 0.5f;
            I.dtPhys = C.dtPhysGlobal;
        } else {
            const lbm3_info &P = G[L - 1].I;
            I.gridID = (float)L;
            I.res = P.res * 0.5f;
            I.cellCountX = P.cellCountX * 2; I.cellCountY = P.cellCountY * 2; I.cellCountZ = P.cellCountZ * 2;
            I.ox = P.ox - I.res * 0.5f; I.oy = P.oy - I.res * 0.5f; I.oz = P.oz - I.res * 0.5f;
            I.dtPhys = P.dtPhys * 0.5f;
        }
        I.nu = (I.dtPhys * C.nuPhys) / ((I.res / 1000.f) * (I.res / 1000.f));
    }
}

// ---------------------------------------------------------------------------
// closed-mesh check (checkSTLEdges, STLFunctions.h:130-315): every edge of a
// closed STL is shared by exactly two triangles.  One lane per triangle counts,
// for its edges AB and BC, the triangles of the bin holding its vertex B that
// have the same edge in either orientation (itself included), and for CA those
// of the bin holding A; the binning puts every triangle into each bin its
// widened bounding box overlaps, so a neighbour sharing an edge is in both
// bins.  The reference prints "Check OK" whatever the counts (its error flag
// array is never set, STLFunctions.h:161-162,289) although its message (:314)
// says the build stops; here an open STL is refused with LBM3_E_BUILD.
// ---------------------------------------------------------------------------
__device__ __forceinline__ bool same_point(const float *a, const float *b)
{
    return a[0] == b[0] && a[1] == b[1] && a[2] == b[2];
}

int stl_edges(DevSTL &S, hipStream_t s)
{
    DBuf<int> ec;
    BCHK(ec.alloc(3LL * S.T));
    const float *v0 = S.v[0].p, *v1 = S.v[1].p, *v2 = S.v[2].p, *v3 = S.v[3].p, *v4 = S.v[4].p, *v5 = S.v[5].p;
    const float *v6 = S.v[6].p, *v7 = S.v[7].p, *v8 = S.v[8].p;
    const int *first = S.firstInBin.p, *bins = S.binArray.p;
    const float ox = S.oxBin, oy = S.oyBin, oz = S.ozBin, bs = S.binSize;
    const int BX = S.bx, BY = S.by;
    int *out = ec.p;
    for_n(S.T, s, [=] __device__(long long t) {
        const float t1[9] = {v0[t], v1[t], v2[t], v3[t], v4[t], v5[t], v6[t], v7[t], v8[t]};
        int cnt[3] = {0, 0, 0};
        for (int side = 0; side < 2; ++side) {
            const float *p = side == 0 ? t1 + 3 : t1;      // vertex B, then vertex A
            const int ib = (int)((p[0] - ox) / bs), jb = (int)((p[1] - oy) / bs), kb = (int)((p[2] - oz) / bs);
            const long long bin = ((long long)kb * BY + jb) * BX + ib;
            const int e0 = side == 0 ? 0 : 2, e1 = side == 0 ? 2 : 3;
            for (int r = first[bin]; r < first[bin + 1]; ++r) {
                const int u = bins[r];
                const float t2[9] = {v0[u], v1[u], v2[u], v3[u], v4[u], v5[u], v6[u], v7[u], v8[u]};
                for (int e = e0; e < e1; ++e) {
                    const float *a1 = t1 + 3 * e, *b1 = t1 + 3 * ((e + 1) % 3);
                    for (int f = 0; f < 3; ++f) {
                        const float *a2 = t2 + 3 * f, *b2 = t2 + 3 * ((f + 1) % 3);
                        cnt[e] += (same_point(a1, a2) && same_point(b1, b2)) ? 1 : 0;
                        cnt[e] += (same_point(a1, b2) && same_point(b1, a2)) ? 1 : 0;
                    }
                }
            }
        }
        out[3 * t] = cnt[0];
        out[3 * t + 1] = cnt[1];
        out[3 * t + 2] = cnt[2];
    });
    BCHK(hipGetLastError());
    S.edges.assign(3 * (size_t)S.T, 0);
    BCHK(hipMemcpyAsync(S.edges.data(), ec.p, sizeof(int) * S.edges.size(), hipMemcpyDeviceToHost, s));
    BCHK(hipStreamSynchronize(s));
    long long faulty = 0;
    for (const int c : S.edges) faulty += c != 2;
    if (faulty)
        return set_error(LBM3_E_BUILD, "checkSTLEdges: " + std::to_string(faulty) +
                                           " edges are not shared by exactly two triangles; the STL is not closed");
    return 0;
}

// ---------------------------------------------------------------------------
// STL upload and binning (sortSTLToBins, STLFunctions.h:5-128)
// ---------------------------------------------------------------------------
int stl_upload(DevSTL &S, const float *tri, int n, float RES, hipStream_t s)
{
    // drop triangles with a repeated vertex, bounds of the rest (readSTL, STLFunctions.h:362-410)
    std::vector<float> soa[9];
    lbm3_bounds B{3.4e38f, 3.4e38f, 3.4e38f, -3.4e38f, -3.4e38f, -3.4e38f, -3.4e38f, -3.4e38f, -3.4e38f};
    for (int t = 0; t < n; ++t) {
        const float *v = tri + 9 * (size_t)t;
        const bool degenerate = (v[0] == v[3] && v[1] == v[4] && v[2] == v[5]) || (v[6] == v[3] && v[7] == v[4] && v[8] == v[5]) ||
                                (v[6] == v[0] && v[7] == v[1] && v[8] == v[2]);
        if (degenerate) continue;
        for (int c = 0; c < 9; ++c) soa[c].push_back(v[c]);
        B.xMin = std::min(B.xMin, std::min(v[0], std::min(v[3], v[6])));
        B.yMin = std::min(B.yMin, std::min(v[1], std::min(v[4], v[7])));
        B.zMin = std::min(B.zMin, std::min(v[2], std::min(v[5], v[8])));
        B.xMax = std::max(B.xMax, std::max(v[0], std::max(v[3], v[6])));
        B.yMax = std::max(B.yMax, std::max(v[1], std::max(v[4], v[7])));
        B.zMax = std::max(B.zMax, std::max(v[2], std::max(v[5], v[8])));
    }
    S.T = (int)soa[0].size();
    if (S.T == 0) return set_error(LBM3_E_BUILD, "voxelizeSTL failed: the passed STL has 0 triangles.");
    for (int c = 0; c < 9; ++c) {
        BCHK(S.v[c].alloc(S.T));
        BCHK(hipMemcpyAsync(S.v[c].p, soa[c].data(), sizeof(float) * S.T, hipMemcpyHostToDevice, s));
    }
    S.B = B;
    const float totalVolume = (B.xMax - B.xMin) * (B.yMax - B.yMin) * (B.zMax - B.zMin);
    const float targetBinVolume = totalVolume / ((float)S.T / (float)10);
    S.binSize = powf(targetBinVolume, 1.f / 3.f);
    const float bs = S.binSize;
    S.oxBin = B.xMin - bs - 2.f * RES;
    S.oyBin = B.yMin - bs - 2.f * RES;
    S.ozBin = B.zMin - bs - 2.f * RES;
    S.bx = (int)((B.xMax - B.xMin + 2.f * bs + 4.f * RES) / bs) + 1;
    S.by = (int)((B.yMax - B.yMin + 2.f * bs + 4.f * RES) / bs) + 1;
    S.bz = (int)((B.zMax - B.zMin + 2.f * bs + 4.f * RES) / bs) + 1;
    const long long nb = (long long)S.bx * S.by * S.bz;
    DBuf<int> cnt, cursor;
    BCHK(cnt.alloc(nb + 1));
    BCHK(cursor.alloc(nb));
    BCHK(hipMemsetAsync(cnt.p, 0, sizeof(int) * (size_t)(nb + 1), s));
    BCHK(hipMemsetAsync(cursor.p, 0, sizeof(int) * (size_t)nb, s));
    BCHK(S.firstInBin.alloc(nb + 1));
    const float *ax = S.v[0].p, *ay = S.v[1].p, *az = S.v[2].p, *bx = S.v[3].p, *by = S.v[4].p, *bz = S.v[5].p;
    const float *cx = S.v[6].p, *cy = S.v[7].p, *cz = S.v[8].p;
    const float ox = S.oxBin, oy = S.oyBin, oz = S.ozBin, tol = bs * 0.2f + RES * 3.f;
    const int BX = S.bx, BY = S.by, BZ = S.bz;
    // total bin entries in 64 bits: a tolerance much larger than the body (tiny STL on a
    // coarse grid) would overflow the int scan, so it is checked before anything is placed
    DBuf<unsigned long long> entries;
    BCHK(entries.alloc(1));
    BCHK(hipMemsetAsync(entries.p, 0, sizeof(unsigned long long), s));
    for (int pass = 0; pass < 2; ++pass) {
        int *c = cnt.p, *cur = cursor.p, *first = S.firstInBin.p, *arr = S.binArray.p;
        unsigned long long *ent = entries.p;
        for_n(S.T, s, [=] __device__(long long t) {
            const float x0 = fminf(ax[t] - ox, fminf(bx[t] - ox, cx[t] - ox)) - tol;
            const float y0 = fminf(ay[t] - oy, fminf(by[t] - oy, cy[t] - oy)) - tol;
            const float z0 = fminf(az[t] - oz, fminf(bz[t] - oz, cz[t] - oz)) - tol;
            const float x1 = fmaxf(ax[t] - ox, fmaxf(bx[t] - ox, cx[t] - ox)) + tol;
            const float y1 = fmaxf(ay[t] - oy, fmaxf(by[t] - oy, cy[t] - oy)) + tol;
            const float z1 = fmaxf(az[t] - oz, fmaxf(bz[t] - oz, cz[t] - oz)) + tol;
            const int i0 = max(0, (int)(x0 / bs)), j0 = max(0, (int)(y0 / bs)), k0 = max(0, (int)(z0 / bs));
            const int i1 = min(BX, (int)(x1 / bs) + 1), j1 = min(BY, (int)(y1 / bs) + 1), k1 = min(BZ, (int)(z1 / bs) + 1);
            if (pass == 0 && i1 > i0 && j1 > j0 && k1 > k0)
                atomicAdd(ent, (unsigned long long)(i1 - i0) * (unsigned long long)(j1 - j0) * (unsigned long long)(k1 - k0));
            for (int kb = k0; kb < k1; ++kb)
                for (int jb = j0; jb < j1; ++jb)
                    for (int ib = i0; ib < i1; ++ib) {
                        const long long bin = ((long long)kb * BY + jb) * BX + ib;
                        if (pass == 0) atomicAdd(&c[bin], 1);
                        else arr[first[bin] + atomicAdd(&cur[bin], 1)] = (int)t;
                    }
        });
        if (pass == 0) {
            unsigned long long total64 = 0;
            BCHK(hipMemcpyAsync(&total64, entries.p, sizeof(total64), hipMemcpyDeviceToHost, s));
            BCHK(hipStreamSynchronize(s));
            if (total64 > 0x7fffffffULL)
                return set_error(LBM3_E_BUILD, "sortSTLToBins: the STL is too small for the grid spacing (more than 2^31 "
                                               "bin entries); use fewer, larger triangles or a finer grid");
            Scratch sc;
            if (exclusive_sum(sc, cnt.p, S.firstInBin.p, nb + 1, s)) return LBM3_E_HIP;
            const int total = read_one(S.firstInBin.p + nb, s);
            BCHK(S.binArray.alloc(total));
        }
    }
    BCHK(hipGetLastError());
    return stl_edges(S, s);
}

// ---------------------------------------------------------------------------
// voxelizer (voxelizeSTL, voxelizerFunctions.h:5-535)
// ---------------------------------------------------------------------------
struct TriSnap {
    int axI, ayI, bxI, byI, cxI, cyI;
    int i0, i1, j0, j1;
};

__device__ __forceinline__ TriSnap snap(float ax, float ay, float bx, float by, float cx, float cy, float res, int X, int Y)
{
    TriSnap t;
    const float scale = 500.0f / res;
    t.axI = (int)roundf(ax * scale) * 2 + 1; t.ayI = (int)roundf(ay * scale) * 2 + 1;
    t.bxI = (int)roundf(bx * scale) * 2 + 1; t.byI = (int)roundf(by * scale) * 2 + 1;
    t.cxI = (int)roundf(cx * scale) * 2 + 1; t.cyI = (int)roundf(cy * scale) * 2 + 1;
    const int xMin = max(0, min(min(t.axI, t.bxI), min(t.cxI, (X - 1) * 1000)));
    const int xMax = min((X - 1) * 1000, max(max(t.axI, t.bxI), max(t.cxI, 0)));
    const int yMin = max(0, min(min(t.ayI, t.byI), min(t.cyI, (Y - 1) * 1000)));
    const int yMax = min((Y - 1) * 1000, max(max(t.ayI, t.byI), max(t.cyI, 0)));
    t.i0 = (xMin + 999) / 1000; t.i1 = xMax / 1000 + 1;
    t.j0 = (yMin + 999) / 1000; t.j1 = yMax / 1000 + 1;
    return t;
}

// exact edge-function hit test with the tie rule for rays on an edge
__device__ __forceinline__ bool ray_hit(int i, int j, const TriSnap &t, long long wab, long long wbc, long long wca,
                                        long long abx, long long bcx, long long cax)
{
    if (wab > 0 && wbc > 0 && wca > 0) return true;
    if (wab < 0 || wbc < 0 || wca < 0) return false;
    const long long rx = (long long)i * 1000, ry = (long long)j * 1000;
    const long long ax0 = t.axI - rx, ay0 = t.ayI - ry, bx0 = t.bxI - rx, by0 = t.byI - ry, cx0 = t.cxI - rx, cy0 = t.cyI - ry;
    if (wab == 0) {
        if (abx == 0) return cx0 < 0;
        if (bx0 > 0 && cy0 * bx0 > by0 * cx0) return true;
        if (bx0 < 0 && cy0 * bx0 < by0 * cx0) return true;
    }
    if (wbc == 0) {
        if (bcx == 0) return ax0 < 0;
        if (cx0 > 0 && ay0 * cx0 > cy0 * ax0) return true;
        if (cx0 < 0 && ay0 * cx0 < cy0 * ax0) return true;
    }
    if (wca == 0) {
        if (cax == 0) return bx0 < 0;
        if (ax0 > 0 && by0 * ax0 > ay0 * bx0) return true;
        if (ax0 < 0 && by0 * ax0 < ay0 * bx0) return true;
    }
    return false;
}

constexpr int RAY_CHUNK = 64;   // rays per thread on large triangles

int voxelize(RayMap &R, const DevSTL &S, const lbm3_info &I, Scratch &sc, hipStream_t s)
{
    const int X = I.cellCountX, Y = I.cellCountY;
    R.nRays = X * Y;
    const float res = I.res, ox = I.ox, oy = I.oy, oz = I.oz;
    const float *ax = S.v[0].p, *ay = S.v[1].p, *az = S.v[2].p, *bx = S.v[3].p, *by = S.v[4].p, *bz = S.v[5].p;
    const float *cx = S.v[6].p, *cy = S.v[7].p, *cz = S.v[8].p;
    // threads per triangle, balanced by ray count
    DBuf<int> tpt, tscan;
    BCHK(tpt.alloc(S.T + 1));
    BCHK(tscan.alloc(S.T + 1));
    int *tp = tpt.p;
    const int T = S.T;
    for_n(T + 1, s, [=] __device__(long long t) {
        if (t == T) { tp[t] = 0; return; }
        const TriSnap q = snap(ax[t] - ox, ay[t] - oy, bx[t] - ox, by[t] - oy, cx[t] - ox, cy[t] - oy, res, X, Y);
        const long long rays = (long long)max(0, q.i1 - q.i0) * (long long)max(0, q.j1 - q.j0);
        tp[t] = (int)((rays + RAY_CHUNK - 1) / RAY_CHUNK);
    });
    if (exclusive_sum(sc, tpt.p, tscan.p, T + 1, s)) return LBM3_E_HIP;
    const int threads = read_one(tscan.p + T, s);
    DBuf<int> cnt;
    BCHK(cnt.alloc(R.nRays + 1));
    BCHK(hipMemsetAsync(cnt.p, 0, sizeof(int) * (size_t)(R.nRays + 1), s));
    BCHK(R.scan.alloc(R.nRays + 1));
    DBuf<long long> cnt64;
    BCHK(cnt64.alloc(R.nRays + 1));
    for (int pass = 0; pass < 2; ++pass) {
        int *c = cnt.p;
        const long long *scan = R.scan.p;
        int *out = R.map.p;
        const int *ts = tscan.p;
        for_n(threads, s, [=] __device__(long long th) {
            // triangle of this thread: last t with ts[t] <= th
            int lo = 0, hi = T - 1;
            while (lo < hi) {
                const int mid = (lo + hi + 1) / 2;
                if (ts[mid] <= th) lo = mid; else hi = mid - 1;
            }
            const int t = lo;
            const float Ax = ax[t] - ox, Ay = ay[t] - oy, Az = az[t] - oz;
            const float Bx = bx[t] - ox, By = by[t] - oy, Bz = bz[t] - oz;
            const float Cx = cx[t] - ox, Cy = cy[t] - oy, Cz = cz[t] - oz;
            const TriSnap q = snap(Ax, Ay, Bx, By, Cx, Cy, res, X, Y);
            const int iSpan = q.i1 - q.i0, jSpan = q.j1 - q.j0;
            if (iSpan <= 0 || jSpan <= 0) return;
            const long long xL0 = 1000LL * q.i0, yL0 = 1000LL * q.j0;
            const long long axL = q.axI - xL0, ayL = q.ayI - yL0, bxL = q.bxI - xL0, byL = q.byI - yL0, cxL = q.cxI - xL0, cyL = q.cyI - yL0;
            const long long abx = bxL - axL, aby = byL - ayL, bcx = cxL - bxL, bcy = cyL - byL, cax = axL - cxL, cay = ayL - cyL;
            const long long area = abx * bcy - aby * bcx;
            if (area == 0) return;
            const long long qz = area < 0 ? -1 : 1;
            const long long wab0 = qz * (abx * (-byL) - aby * (-bxL));
            const long long wbc0 = qz * (bcx * (-cyL) - bcy * (-cxL));
            const long long wca0 = qz * (cax * (-ayL) - cay * (-axL));
            const long long dab_i = -qz * aby * 1000, dab_j = qz * abx * 1000;
            const long long dbc_i = -qz * bcy * 1000, dbc_j = qz * bcx * 1000;
            const long long dca_i = -qz * cay * 1000, dca_j = qz * cax * 1000;
            const float v1x = Bx - Ax, v1y = By - Ay, v1z = Bz - Az, v2x = Cx - Ax, v2y = Cy - Ay, v2z = Cz - Az;
            const float nx = v1y * v2z - v1z * v2y, ny = v1z * v2x - v1x * v2z, nz = v1x * v2y - v1y * v2x;
            const float maxZ = fmaxf(fmaxf(Az, Bz), Cz), minZ = fminf(fminf(Az, Bz), Cz);
            const float midZ = 0.5f * (maxZ + minZ);
            const long long task0 = (long long)(th - ts[t]) * RAY_CHUNK;
            const long long tasks = (long long)iSpan * jSpan;
            for (long long task = task0; task < task0 + RAY_CHUNK && task < tasks; ++task) {
                const int j = q.j0 + (int)(task / iSpan), i = q.i0 + (int)(task % iSpan);
                const long long wab = wab0 + (long long)(i - q.i0) * dab_i + (long long)(j - q.j0) * dab_j;
                const long long wbc = wbc0 + (long long)(i - q.i0) * dbc_i + (long long)(j - q.j0) * dbc_j;
                const long long wca = wca0 + (long long)(i - q.i0) * dca_i + (long long)(j - q.j0) * dca_j;
                if (!ray_hit(i, j, q, wab, wbc, wca, abx, bcx, cax)) continue;
                const int ray = j * X + i;
                if (pass == 0) { atomicAdd(&c[ray], 1); continue; }
                const float rayX = (float)i * res, rayY = (float)j * res;
                float rayZ;
                if (nz != 0.0f) {
                    rayZ = Az - (nx * (rayX - Ax) + ny * (rayY - Ay)) / nz;
                    if (rayZ > maxZ) rayZ = maxZ;
                    else if (rayZ < minZ) rayZ = minZ;
                } else rayZ = midZ;
                out[scan[ray] + atomicAdd(&c[ray], 1)] = (int)ceilf(rayZ / res);
            }
        });
        if (pass == 0) {
            const int *ci = cnt.p;
            long long *c64 = cnt64.p;
            const int nr = R.nRays;
            for_n(nr + 1, s, [=] __device__(long long r) { c64[r] = r < nr ? (long long)ci[r] : 0LL; });
            if (exclusive_sum(sc, cnt64.p, R.scan.p, (long long)R.nRays + 1, s)) return LBM3_E_HIP;
            R.total = read_one(R.scan.p + R.nRays, s);
            BCHK(R.map.alloc(R.total));
            BCHK(hipMemsetAsync(cnt.p, 0, sizeof(int) * (size_t)(R.nRays + 1), s));
        }
    }
    // sort every ray's hit list (short lists: insertion sort)
    const long long *scan = R.scan.p;
    int *m = R.map.p;
    for_n(R.nRays, s, [=] __device__(long long r) {
        const long long s0 = scan[r], n = scan[r + 1] - s0;
        for (long long a = 1; a < n; ++a) {
            const int key = m[s0 + a];
            long long b = a - 1;
            while (b >= 0 && m[s0 + b] > key) { m[s0 + b + 1] = m[s0 + b]; --b; }
            m[s0 + b + 1] = key;
        }
    });
    return check_hip(hipGetLastError(), "voxelizer");
}

// union of solid intervals (sumRayMaps, voxelizerFunctions.h:537-690)
int raymap_union(RayMap &sum, const RayMap &bonus, Scratch &sc, hipStream_t s)
{
    const int n = sum.nRays;
    RayMap out;
    out.nRays = n;
    BCHK(out.scan.alloc(n + 1));
    DBuf<long long> cnt;
    BCHK(cnt.alloc(n + 1));
    for (int pass = 0; pass < 2; ++pass) {
        const long long *ss = sum.scan.p, *bs = bonus.scan.p, *os = out.scan.p;
        const int *sm = sum.map.p, *bm = bonus.map.p;
        int *om = out.map.p;
        long long *c = cnt.p;
        for_n(n + 1, s, [=] __device__(long long r) {
            if (r == n) { if (pass == 0) c[r] = 0; return; }
            long long si = ss[r], se = ss[r + 1], bi = bs[r], be = bs[r + 1];
            long long w = pass ? os[r] : 0;
            bool active = false;
            long long lastStart = 0, lastEnd = 0;
            while (si < se || bi < be) {
                long long cs, ce;
                if (bi >= be || (si < se && sm[si] <= bm[bi])) { cs = sm[si]; ce = sm[si + 1]; si += 2; }
                else { cs = bm[bi]; ce = bm[bi + 1]; bi += 2; }
                if (!active) { lastStart = cs; lastEnd = ce; active = true; }
                else if (cs <= lastEnd) { if (ce > lastEnd) lastEnd = ce; }
                else {
                    if (pass) { om[w] = (int)lastStart; om[w + 1] = (int)lastEnd; }
                    w += 2; lastStart = cs; lastEnd = ce;
                }
            }
            if (active) { if (pass) { om[w] = (int)lastStart; om[w + 1] = (int)lastEnd; } w += 2; }
            if (!pass) c[r] = w;
        });
        if (!pass) {
            if (exclusive_sum(sc, cnt.p, out.scan.p, (long long)n + 1, s)) return LBM3_E_HIP;
            out.total = read_one(out.scan.p + n, s);
            BCHK(out.map.alloc(out.total));
        }
    }
    std::swap(sum.scan.p, out.scan.p); std::swap(sum.scan.n, out.scan.n);
    std::swap(sum.map.p, out.map.p); std::swap(sum.map.n, out.map.n);
    sum.total = out.total;
    return 0;
}

int raymap_copy(RayMap &d, const RayMap &src, hipStream_t s)
{
    d.nRays = src.nRays;
    d.total = src.total;
    BCHK(d.scan.alloc(src.nRays + 1));
    BCHK(hipMemcpyAsync(d.scan.p, src.scan.p, sizeof(long long) * (size_t)(src.nRays + 1), hipMemcpyDeviceToDevice, s));
    BCHK(d.map.alloc(src.total));
    if (src.total) BCHK(hipMemcpyAsync(d.map.p, src.map.p, sizeof(int) * (size_t)src.total, hipMemcpyDeviceToDevice, s));
    return 0;
}

// ---------------------------------------------------------------------------
// per-cell marker primitives (markerFunctions.h)
// ---------------------------------------------------------------------------
__device__ __forceinline__ bool in_wall(const RayView R, int ray, int k)
{
    for (long long a = R.scan[ray]; a < R.scan[ray + 1]; a += 2) {
        if (R.map[a] > k) break;
        if (R.map[a + 1] > k) return true;
    }
    return false;
}

__device__ __forceinline__ bool finer_fluid(const RayView R, int Xc, int ic, int jc, int kc, int ds)
{
    const int i0 = ic * ds, j0 = jc * ds, k0 = kc * ds, kLast = k0 + ds - 1;
    for (int ja = 0; ja < ds; ++ja)
        for (int ia = 0; ia < ds; ++ia) {
            const int ray = (Xc * ds) * (j0 + ja) + (i0 + ia);
            const long long s0 = R.scan[ray], e0 = R.scan[ray + 1];
            if (s0 == e0) return true;
            for (long long a = s0; a < e0; a += 2) {
                const int kEnd = R.map[a + 1];
                if (kEnd < k0) { if (a + 2 < e0) continue; else return true; }
                else if (kEnd <= kLast) return true;
                if (R.map[a] <= k0) break;
                else return true;
            }
        }
    return false;
}

__device__ __forceinline__ bool finer_wall(const RayView R, int Xc, int ic, int jc, int kc, int ds)
{
    const int i0 = ic * ds, j0 = jc * ds, k0 = kc * ds, kLast = k0 + ds - 1;
    for (int ja = 0; ja < ds; ++ja)
        for (int ia = 0; ia < ds; ++ia) {
            const int ray = (Xc * ds) * (j0 + ja) + (i0 + ia);
            for (long long a = R.scan[ray]; a < R.scan[ray + 1]; a += 2) {
                const int kStart = R.map[a];
                if (kStart > kLast) break;
                if (kStart >= k0) return true;
                if (R.map[a + 1] <= k0) continue;
                return true;
            }
        }
    return false;
}

struct Views {
    int N;
    const int *i, *j, *k, *jPlus, *kPlus;
    const uint8_t *isGeo;
};
Views views(const BLevel &B) { return Views{B.N, B.i.p, B.j.p, B.k.p, B.jPlus.p, B.kPlus.p, B.isGeo.p}; }

// spreadMarkers via geometric plus-links, both ways (markerFunctions.h:332-399)
int spread(const BLevel &B, uint8_t *t, const uint8_t *src, hipStream_t s)
{
    BCHK(hipMemcpyAsync(t, src, (size_t)B.N, hipMemcpyDeviceToDevice, s));
    const Views v = views(B);
    for_n(B.N, s, [=] __device__(long long cl) {
        const int c = (int)cl, N = v.N;
        int nb[7];
        nb[0] = c + 1 < N ? c + 1 : 0;
        nb[1] = v.jPlus[c];
        nb[2] = nb[1] + 1 < N ? nb[1] + 1 : 0;
        nb[3] = v.kPlus[c];
        nb[4] = nb[3] + 1 < N ? nb[3] + 1 : 0;
        nb[5] = v.jPlus[nb[3]];
        nb[6] = nb[5] + 1 < N ? nb[5] + 1 : 0;
        const uint8_t geo = v.isGeo[c];
        bool marker = src[c] != 0;
        if (!marker)
            for (int q = 0; q < 7; ++q)
                if (((geo >> q) & 1) && src[nb[q]]) { marker = true; break; }
        if (marker) {
            t[c] = 1;
            for (int q = 0; q < 7; ++q)
                if ((geo >> q) & 1) t[nb[q]] = 1;   // every writer stores 1
        }
    });
    return 0;
}

// ---------------------------------------------------------------------------
// cell lists from dense occupancy, neighbour rows, geometric bits
// ---------------------------------------------------------------------------
// cells = set positions of `exists` over the level box, sorted (k,j,i) by construction
int cells_from_dense(BLevel &B, const uint8_t *exists, Scratch &sc, hipStream_t s)
{
    const long long box = (long long)B.I.cellCountX * B.I.cellCountY * B.I.cellCountZ;
    DBuf<int> lin;
    const int n = compact(sc, exists, box, lin, s);
    if (n < 0) return LBM3_E_HIP;
    B.N = n;
    B.I.cellCount = n;
    BCHK(B.i.alloc(n)); BCHK(B.j.alloc(n)); BCHK(B.k.alloc(n));
    BCHK(B.dense.alloc(box));
    BCHK(hipMemsetAsync(B.dense.p, 0xff, sizeof(int) * (size_t)box, s));
    int *I = B.i.p, *J = B.j.p, *K = B.k.p, *D = B.dense.p;
    const int *l = lin.p;
    const long long X = B.I.cellCountX, XY = X * B.I.cellCountY;
    for_n(n, s, [=] __device__(long long c) {
        const long long p = l[c];
        K[c] = (int)(p / XY);
        J[c] = (int)((p % XY) / X);
        I[c] = (int)(p % X);
        D[p] = (int)c;
    });
    return 0;
}

// dense map of the current cell list (after a compaction)
int rebuild_dense(BLevel &B, hipStream_t s)
{
    const long long box = (long long)B.I.cellCountX * B.I.cellCountY * B.I.cellCountZ;
    BCHK(B.dense.alloc(box));
    BCHK(hipMemsetAsync(B.dense.p, 0xff, sizeof(int) * (size_t)box, s));
    int *D = B.dense.p;
    const int *I = B.i.p, *J = B.j.p, *K = B.k.p;
    const long long X = B.I.cellCountX, XY = X * B.I.cellCountY;
    for_n(B.N, s, [=] __device__(long long c) { D[(long long)K[c] * XY + (long long)J[c] * X + I[c]] = (int)c; });
    return 0;
}

// connectNBRInRows + buildNBRPlus (NBRFunctions.h:44-152) + markGeometricNBRPlus (:154-194)
int build_nbr(BLevel &B, hipStream_t s)
{
    BCHK(B.jPlus.alloc(B.N)); BCHK(B.kPlus.alloc(B.N)); BCHK(B.isGeo.alloc(B.N));
    const int X = B.I.cellCountX, Y = B.I.cellCountY, Z = B.I.cellCountZ;
    const int *D = B.dense.p;
    int *jp = B.jPlus.p, *kp = B.kPlus.p;
    // j rows: columns of equal (i,k); each cell links to the next present j, the last wraps to the first
    for_n((long long)X * Z, s, [=] __device__(long long r) {
        const int i = (int)(r % X), k = (int)(r / X);
        int first = -1, prev = -1;
        for (int j = 0; j < Y; ++j) {
            const int c = D[((long long)k * Y + j) * X + i];
            if (c < 0) continue;
            if (prev >= 0) jp[prev] = c; else first = c;
            prev = c;
        }
        if (prev >= 0) jp[prev] = first;
    });
    for_n((long long)X * Y, s, [=] __device__(long long r) {
        const int i = (int)(r % X), j = (int)(r / X);
        int first = -1, prev = -1;
        for (int k = 0; k < Z; ++k) {
            const int c = D[((long long)k * Y + j) * X + i];
            if (c < 0) continue;
            if (prev >= 0) kp[prev] = c; else first = c;
            prev = c;
        }
        if (prev >= 0) kp[prev] = first;
    });
    const int N = B.N;
    const int *I = B.i.p, *J = B.j.p, *K = B.k.p;
    uint8_t *geo = B.isGeo.p;
    for_n(N, s, [=] __device__(long long cl) {
        const int c = (int)cl;
        const int ic = I[c], jc = J[c], kc = K[c];
        const int ip = c + 1 < N ? c + 1 : 0;
        const int jq = jp[c], ijq = jq + 1 < N ? jq + 1 : 0;
        const int kq = kp[c], ikq = kq + 1 < N ? kq + 1 : 0;
        const int jkq = jp[kq], ijkq = jkq + 1 < N ? jkq + 1 : 0;
        uint8_t b = 0;
        auto at = [&](int n, int di, int dj, int dk) { return I[n] == ic + di && J[n] == jc + dj && K[n] == kc + dk; };
        if (at(ip, 1, 0, 0)) b |= 1;
        if (at(jq, 0, 1, 0)) b |= 2;
        if (at(ijq, 1, 1, 0)) b |= 4;
        if (at(kq, 0, 0, 1)) b |= 8;
        if (at(ikq, 1, 0, 1)) b |= 16;
        if (at(jkq, 0, 1, 1)) b |= 32;
        if (at(ijkq, 1, 1, 1)) b |= 64;
        geo[c] = b;
    });
    return check_hip(hipGetLastError(), "neighbour rows");
}

// parent cell through the coarse dense map (binarySearchIJK of (i/2, j/2, k/2))
int parent_map(BLevel &B, const BLevel &P, hipStream_t s)
{
    BCHK(B.parentMap.alloc(B.N));
    int *pm = B.parentMap.p;
    const int *I = B.i.p, *J = B.j.p, *K = B.k.p, *D = P.dense.p;
    const long long X = P.I.cellCountX, XY = X * P.I.cellCountY;
    for_n(B.N, s, [=] __device__(long long c) { pm[c] = D[(long long)(K[c] / 2) * XY + (long long)(J[c] / 2) * X + I[c] / 2]; });
    return 0;
}

struct Views2 {
    RayView r[LBM3_MAX_LEVELS];
};

int build_interface_lists(Sim &sim, std::vector<BLevel> &G, int L, Scratch &sc);

// ---------------------------------------------------------------------------
// buildWallMarkers (:590-739) and gridBuilderToGrid (:1032-1569)
// ---------------------------------------------------------------------------
__device__ __forceinline__ int interp_ray_triangle(float ax, float ay, float az, float bx, float by, float bz, float cx, float cy,
                                                   float cz, float ex, float ey, float ez, float edgeTol, float &t)
{
    // Moller-Trumbore with physical edge tolerance (interpolatedBouncebackFunctions.h:397-480)
    const float e1x = bx - ax, e1y = by - ay, e1z = bz - az, e2x = cx - ax, e2y = cy - ay, e2z = cz - az;
    const float nx = e1y * e2z - e1z * e2y, ny = e1z * e2x - e1x * e2z, nz = e1x * e2y - e1y * e2x;
    const float area2 = sqrtf(nx * nx + ny * ny + nz * nz);
    if (area2 == 0.f) return 0;
    const float px = ey * e2z - ez * e2y, py = ez * e2x - ex * e2z, pz = ex * e2y - ey * e2x;
    const float det = e1x * px + e1y * py + e1z * pz;
    if (det > -1e-8f && det < 1e-8f) return 0;
    const float invDet = 1.f / det;
    const float lenAB = sqrtf(e1x * e1x + e1y * e1y + e1z * e1z);
    const float lenAC = sqrtf(e2x * e2x + e2y * e2y + e2z * e2z);
    const float bcx = cx - bx, bcy = cy - by, bcz = cz - bz;
    const float lenBC = sqrtf(bcx * bcx + bcy * bcy + bcz * bcz);
    const float uTol = edgeTol * lenAC / area2, vTol = edgeTol * lenAB / area2, wTol = edgeTol * lenBC / area2;
    const float tx = -ax, ty = -ay, tz = -az;
    const float u = (tx * px + ty * py + tz * pz) * invDet;
    if (u < -uTol) return 0;
    const float qx = ty * e1z - tz * e1y, qy = tz * e1x - tx * e1z, qz = tx * e1y - ty * e1x;
    const float v = (ex * qx + ey * qy + ez * qz) * invDet;
    if (v < -vTol || u + v > 1.f + wTol) return 0;
    t = (e2x * qx + e2y * qy + e2z * qz) * invDet;
    return 1;
}

struct StlView {
    const float *v[9];
    const int *first, *bins;
    float binSize, ox, oy, oz;
    int bx, by, bz;
};

// getOpenBC of every boundary cell of a level (rm[0]) and of its projections
// onto the x, y and z boundary planes where the cell lies within `layers` of
// them (rm[1..3], the periodic-overlap test, gridBuilderFunctions.h:1232-1274),
// as indices into `recs`, the records those calls return (-1: no call).  From
// the case's rule table (records: the rules, then the default), or from the
// case's own hook through the open_bc callback (records: the distinct BCStructs
// it returned, in order of first appearance).
static int open_rule_maps(Sim &sim, BLevel &B, int L, const lbm3_case *dc, Scratch &sc, DBuf<int> (&rm)[4],
                          std::vector<lbm3_bc> &recs)
{
    hipStream_t s = sim.stream;
    const long long N = B.N;
    const Views v = views(B);
    const lbm3_info Il = B.I;
    const int layers = 1 << L, lev = L;
    const uint8_t *wall = B.wall.p;
    for (DBuf<int> &a : rm) BCHK(a.alloc(N > 0 ? N : 1));
    int *r0 = rm[0].p, *r1 = rm[1].p, *r2 = rm[2].p, *r3 = rm[3].p;
    if (!sim.hooks.open_bc) {
        for_n(N, s, [=] __device__(long long c) {
            r0[c] = r1[c] = r2[c] = r3[c] = -1;
            if (wall[c]) return;
            const int i = v.i[c], j = v.j[c], k = v.k[c];
            const int X = Il.cellCountX, Y = Il.cellCountY, Z = Il.cellCountZ;
            if (i == 0 || i == X - 1 || j == 0 || j == Y - 1 || k == 0 || k == Z - 1) r0[c] = open_rule(*dc, lev, X, Y, Z, i, j, k);
            if (i < layers || i >= X - layers) r1[c] = open_rule(*dc, lev, X, Y, Z, i < layers ? 0 : X - 1, j, k);
            if (j < layers || j >= Y - layers) r2[c] = open_rule(*dc, lev, X, Y, Z, i, j < layers ? 0 : Y - 1, k);
            if (k < layers || k >= Z - layers) r3[c] = open_rule(*dc, lev, X, Y, Z, i, j, k < layers ? 0 : Z - 1);
        });
        const lbm3_case &C = sim.C;
        recs.clear();
        for (int r = 0; r < C.nOpenRules; ++r) recs.push_back(C.openRules[r].bc);
        recs.push_back(C.openDefault);
        return 0;
    }
    // queries: for each of the four calls, the cells that make it, compacted
    DBuf<uint8_t> fl;
    BCHK(fl.alloc(N > 0 ? N : 1));
    DBuf<int> lists[4];
    long long count[4], total = 0;
    for (int t = 0; t < 4; ++t) {
        uint8_t *f = fl.p;
        for_n(N, s, [=] __device__(long long c) {
            bool q = false;
            if (!wall[c]) {
                const int i = v.i[c], j = v.j[c], k = v.k[c];
                const int X = Il.cellCountX, Y = Il.cellCountY, Z = Il.cellCountZ;
                if (t == 0) q = i == 0 || i == X - 1 || j == 0 || j == Y - 1 || k == 0 || k == Z - 1;
                else if (t == 1) q = i < layers || i >= X - layers;
                else if (t == 2) q = j < layers || j >= Y - layers;
                else q = k < layers || k >= Z - layers;
            }
            f[c] = q;
        });
        count[t] = compact(sc, fl.p, N, lists[t], s);
        if (count[t] < 0) return LBM3_E_HIP;
        total += count[t];
    }
    recs.clear();
    for (int t = 0; t < 4; ++t) for_n(N, s, [=] __device__(long long c) { (t == 0 ? r0 : t == 1 ? r1 : t == 2 ? r2 : r3)[c] = -1; });
    if (total == 0) return 0;
    if (total > (1LL << 31) - 1) return set_error(LBM3_E_MEMORY, "too many open-boundary hook calls on one level");
    DBuf<int> qi, qj, qk;
    BCHK(qi.alloc(total));
    BCHK(qj.alloc(total));
    BCHK(qk.alloc(total));
    long long off = 0;
    for (int t = 0; t < 4; ++t) {
        const int *lc = lists[t].p;
        int *oi = qi.p + off, *oj = qj.p + off, *ok = qk.p + off;
        for_n(count[t], s, [=] __device__(long long x) {
            const int c = lc[x];
            int i = v.i[c], j = v.j[c], k = v.k[c];
            const int X = Il.cellCountX, Y = Il.cellCountY, Z = Il.cellCountZ;
            if (t == 1) i = i < layers ? 0 : X - 1;
            else if (t == 2) j = j < layers ? 0 : Y - 1;
            else if (t == 3) k = k < layers ? 0 : Z - 1;
            oi[x] = i; oj[x] = j; ok[x] = k;
        });
        off += count[t];
    }
    DBuf<lbm3_bc> out;
    BCHK(out.alloc(total));
    lbm3_info info = Il;
    if (sim.hooks.open_bc(sim.hooks.user, &info, qi.p, qj.p, qk.p, (int)total, out.p, (void *)s) != 0)
        return set_error(LBM3_E_ARG, "the getOpenBC hook callback failed");
    BCHK(hipGetLastError());
    std::vector<lbm3_bc> host((size_t)total);
    BCHK(hipMemcpyAsync(host.data(), out.p, sizeof(lbm3_bc) * host.size(), hipMemcpyDeviceToHost, s));
    BCHK(hipStreamSynchronize(s));
    std::vector<int> idx((size_t)total);
    dedupe_bc(host, idx, recs);
    DBuf<int> di;
    BCHK(di.alloc(total));
    BCHK(hipMemcpyAsync(di.p, idx.data(), sizeof(int) * idx.size(), hipMemcpyHostToDevice, s));
    off = 0;
    for (int t = 0; t < 4; ++t) {
        const int *lc = lists[t].p, *ix = di.p + off;
        int *dst = t == 0 ? r0 : t == 1 ? r1 : t == 2 ? r2 : r3;
        for_n(count[t], s, [=] __device__(long long x) { dst[lc[x]] = ix[x]; });
        off += count[t];
    }
    BCHK(hipStreamSynchronize(s));   // the host vectors above outlive the copies
    return 0;
}

int walls_and_grid(Sim &sim, std::vector<BLevel> &G, std::vector<DevSTL> &stl, int nBodies, const lbm3_case *dc, Scratch &sc,
                   long long *fluidUpdates)
{
    const lbm3_case &C = sim.C;
    const int levels = (int)G.size();
    hipStream_t s = sim.stream;
    // ---- buildWallMarkers ----
    for (int L = 0; L < levels; ++L) {
        BLevel &B = G[L];
        const bool finest = L == levels - 1, coarsest = L == 0;
        const int N = B.N;
        const Views v = views(B);
        const int X = B.I.cellCountX;
        uint8_t *wall = B.wall.p;
        {
            const RayView rt{B.total.scan.p, B.total.map.p};
            const int *pm = B.parentMap.p;
            const uint8_t *pw = coarsest ? nullptr : G[L - 1].wall.p;
            for_n(N, s, [=] __device__(long long c) {
                bool w = in_wall(rt, X * v.j[c] + v.i[c], v.k[c]);
                if (!coarsest && pm[c] >= 0) w = pw[pm[c]] != 0;
                wall[c] = w;
            });
        }
        BCHK(B.wallIDFull.alloc(N));
        BCHK(hipMemsetAsync(B.wallIDFull.p, 0xff, sizeof(int) * (size_t)N, s));
        DBuf<uint8_t> body, src;
        BCHK(body.alloc(N));
        BCHK(src.alloc(N));
        for (int b = 0; b < nBodies; ++b) {
            const RayView rb{B.maps[b].scan.p, B.maps[b].map.p};
            uint8_t *bp = body.p;
            for_n(N, s, [=] __device__(long long c) { bp[c] = in_wall(rb, X * v.j[c] + v.i[c], v.k[c]); });
            if (!coarsest) {
                const BLevel &P = G[L - 1];
                const RayView rp{P.maps[b].scan.p, P.maps[b].map.p};
                const int *pm = B.parentMap.p, *PI = P.i.p, *PJ = P.j.p, *PK = P.k.p;
                const int PX = P.I.cellCountX;
                for_n(N, s, [=] __device__(long long c) {
                    const int p = pm[c];
                    if (p >= 0) bp[c] = in_wall(rp, PX * PJ[p] + PI[p], PK[p]);
                });
            }
            BCHK(hipMemcpyAsync(src.p, body.p, (size_t)N, hipMemcpyDeviceToDevice, s));
            if (spread(B, body.p, src.p, s)) return LBM3_E_HIP;
            int *wid = B.wallIDFull.p;
            const int *pm = B.parentMap.p;
            const int *pwid = coarsest ? nullptr : G[L - 1].wallIDFull.p;
            const int bb = b;
            for_n(N, s, [=] __device__(long long c) {
                if (!bp[c] || wall[c]) return;
                int id = bb;
                if (!coarsest && pm[c] >= 0 && pwid[pm[c]] >= 0) id = pwid[pm[c]];
                wid[c] = id;
            });
        }
        {
            DBuf<uint8_t> adj;
            BCHK(adj.alloc(N));
            uint8_t *a = adj.p;
            const int *wid = B.wallIDFull.p;
            for_n(N, s, [=] __device__(long long c) { a[c] = wid[c] >= 0; });
            B.W = compact(sc, adj.p, N, B.wallList, s);
            if (B.W < 0) return LBM3_E_HIP;
            BCHK(B.wallIDc.alloc(B.W));
            int *wc = B.wallIDc.p;
            const int *wl = B.wallList.p;
            for_n(B.W, s, [=] __device__(long long w) { wc[w] = wid[wl[w]]; });
        }
        const int W = B.W;
        BCHK(B.linkExist.alloc(27LL * (W > 0 ? W : 1)));
        BCHK(B.linkPierce.alloc(27LL * (W > 0 ? W : 1)));
        BCHK(B.linkLen.alloc(27LL * (W > 0 ? W : 1)));
        BCHK(hipMemsetAsync(B.linkExist.p, 0, (size_t)B.linkExist.n, s));
        BCHK(hipMemsetAsync(B.linkPierce.p, 0, (size_t)B.linkPierce.n, s));
        // link existence (interpolatedBouncebackFunctions.h:181-395)
        {
            const lbm3_info Il = B.I;
            const int *wl = B.wallList.p, *JM = B.jMinus.p, *KM = B.kMinus.p;
            const int *pm = coarsest ? nullptr : B.parentMap.p;
            const uint8_t *f2c = finest ? nullptr : B.f2c.p, *c2f = finest ? nullptr : B.c2f.p;
            uint8_t *le = B.linkExist.p, *lp = B.linkPierce.p;
            for_n(W, s, [=] __device__(long long wq) {
                const int w = (int)wq, c = wl[w], Nn = v.N;
                const int ic = v.i[c], jc = v.j[c], kc = v.k[c];
                auto isIf = [&](int x) {
                    bool r = false;
                    if (!coarsest && pm[x] >= 0) r = true;
                    if (!finest && (f2c[x] || c2f[x])) r = true;
                    return r;
                };
                const bool meIf = isIf(c);
                const int ip = c + 1 >= Nn ? 0 : c + 1, im = c - 1 < 0 ? Nn - 1 : c - 1;
                const int *JP = v.jPlus, *KP = v.kPlus;
                const int jp = JP[c], kp = KP[c], jm = JM[c], km = KM[c];
                int nb[27];
                nb[OOO] = c; nb[POO] = ip; nb[MOO] = im; nb[OOM] = km; nb[OOP] = kp; nb[OMO] = jm; nb[OPO] = jp;
                nb[POM] = KM[ip]; nb[MOP] = KP[im]; nb[POP] = KP[ip]; nb[MOM] = KM[im];
                nb[MMO] = JM[im]; nb[PPO] = JP[ip]; nb[OPM] = KM[jp]; nb[OMP] = KP[jm];
                nb[MPO] = JP[im]; nb[PMO] = JM[ip]; nb[OPP] = JP[kp]; nb[OMM] = KM[jm];
                nb[MPM] = KM[JP[im]]; nb[PMP] = KP[JM[ip]]; nb[MMP] = KP[JM[im]]; nb[PPM] = KM[JP[ip]];
                nb[PMM] = KM[JM[ip]]; nb[MPP] = JP[KP[im]]; nb[MMM] = KM[JM[im]]; nb[PPP] = JP[KP[ip]];
                uint32_t geo = 0;
                bool all = true, ipc = true, jpc = true, kpc = true, imc = true, jmc = true, kmc = true;
                for (int d = 1; d < 27; ++d) {
                    const int ie = ic + cx(d), je = jc + cy(d), ke = kc + cz(d), n = nb[d];
                    if (v.i[n] == ie && v.j[n] == je && v.k[n] == ke) { geo |= 1u << d; continue; }
                    if (ie < 0 || ie >= Il.cellCountX || je < 0 || je >= Il.cellCountY || ke < 0 || ke >= Il.cellCountZ) continue;
                    all = false;
                    if (cx(d) > 0) ipc = false;
                    if (cy(d) > 0) jpc = false;
                    if (cz(d) > 0) kpc = false;
                    if (cx(d) < 0) imc = false;
                    if (cy(d) < 0) jmc = false;
                    if (cz(d) < 0) kmc = false;
                }
                const bool interior = ic != 0 && ic != Il.cellCountX - 1 && jc != 0 && jc != Il.cellCountY - 1 && kc != 0 &&
                                      kc != Il.cellCountZ - 1;
                if (interior && !all) return;
                for (int d = 1; d < 27; ++d) {
                    const int n = nb[d];
                    if (!interior && !((geo >> d) & 1u)) continue;
                    if (!wall[n]) continue;
                    if (!interior) {
                        if (cx(d) > 0 && !ipc) continue;
                        if (cy(d) > 0 && !jpc) continue;
                        if (cz(d) > 0 && !kpc) continue;
                        if (cx(d) < 0 && !imc) continue;
                        if (cy(d) < 0 && !jmc) continue;
                        if (cz(d) < 0 && !kmc) continue;
                    }
                    le[(size_t)d * W + w] = 1;
                    if (meIf || isIf(n)) lp[(size_t)d * W + w] = 1;
                }
            });
        }
        // link lengths (interpolatedBouncebackFunctions.h:482-660)
        {
            float *ll = B.linkLen.p;
            for_n(27LL * (W > 0 ? W : 1), s, [=] __device__(long long x) { ll[x] = 2.f; });
            const lbm3_info Il = B.I;
            const int *wl = B.wallList.p;
            const uint8_t *le = B.linkExist.p, *lp = B.linkPierce.p;
            for (size_t b = 0; b < stl.size(); ++b) {
                const DevSTL &S = stl[b];
                StlView sv;
                for (int q = 0; q < 9; ++q) sv.v[q] = S.v[q].p;
                sv.first = S.firstInBin.p; sv.bins = S.binArray.p;
                sv.binSize = S.binSize; sv.ox = S.oxBin; sv.oy = S.oyBin; sv.oz = S.ozBin;
                sv.bx = S.bx; sv.by = S.by; sv.bz = S.bz;
                for_n(W, s, [=] __device__(long long w) {
                    const int c = wl[w];
                    const float xc = (float)v.i[c] * Il.res + Il.ox, yc = (float)v.j[c] * Il.res + Il.oy, zc = (float)v.k[c] * Il.res + Il.oz;
                    const int ib = (int)((xc - sv.ox) / sv.binSize), jb = (int)((yc - sv.oy) / sv.binSize), kb = (int)((zc - sv.oz) / sv.binSize);
                    if (ib < 0 || ib >= sv.bx || jb < 0 || jb >= sv.by || kb < 0 || kb >= sv.bz) return;
                    const long long bin = ((long long)kb * sv.by + jb) * sv.bx + ib;
                    for (int r = sv.first[bin]; r < sv.first[bin + 1]; ++r) {
                        const int t = sv.bins[r];
                        const float ax = sv.v[0][t] - xc, ay = sv.v[1][t] - yc, az = sv.v[2][t] - zc;
                        const float bx = sv.v[3][t] - xc, by = sv.v[4][t] - yc, bz = sv.v[5][t] - zc;
                        const float cxx = sv.v[6][t] - xc, cyy = sv.v[7][t] - yc, czz = sv.v[8][t] - zc;
                        for (int d = 1; d < 27; ++d) {
                            const size_t slot = (size_t)d * W + w;
                            if (!le[slot]) continue;
                            if (lp[slot]) { ll[slot] = 0.5f; continue; }
                            float tt;
                            if (!interp_ray_triangle(ax, ay, az, bx, by, bz, cxx, cyy, czz, (float)cx(d), (float)cy(d), (float)cz(d),
                                                     0.1f * Il.res, tt)) continue;
                            float q = tt / Il.res;
                            if (q > 0 && q <= 1.f) {
                                q = q < 0.00001f ? 0.00001f : (q > 1.f ? 1.f : q);
                                const float qp = ll[slot];
                                if (qp < 0.f || q < qp) ll[slot] = q;
                            } else if (q > -0.2f && q <= 1.f + 0.2f) {
                                q = q < 0.00001f ? 0.00001f : (q > 1.f ? 1.f : q);
                                const float qp = ll[slot];
                                if (qp < 0.f) { if (q < -qp) ll[slot] = -q; }
                                else if (qp > 1.f) ll[slot] = -q;
                            }
                        }
                    }
                });
            }
            for_n(W, s, [=] __device__(long long w) {
                for (int d = 1; d < 27; ++d) {
                    const size_t slot = (size_t)d * W + w;
                    if (!le[slot]) continue;
                    if (lp[slot]) { ll[slot] = 0.5f; continue; }
                    const float qp = ll[slot];
                    if (qp > 1.f) ll[slot] = 0.5f;
                    else if (qp < 0.f) ll[slot] = -qp;
                }
            });
        }
        // interface overlap (:715-734)
        {
            uint8_t *ov = B.ifOverlap.p;
            const uint8_t *f2c = finest ? nullptr : B.f2c.p;
            const int *pm = coarsest ? nullptr : B.parentMap.p;
            const uint8_t *pc2f = coarsest ? nullptr : G[L - 1].c2f.p;
            for_n(N, s, [=] __device__(long long c) {
                bool o = ov[c] != 0;
                if (!finest && f2c[c]) o = true;
                if (!coarsest && pm[c] >= 0 && pc2f[pm[c]]) o = true;
                ov[c] = o;
            });
        }
    }

    // ---- gridBuilderToGrid ----
    long long fluid = 0;
    for (int L = 0; L < levels; ++L) {
        BLevel &B = G[L];
        Level &l = sim.L[L];
        const int N = B.N, W = B.W;
        const Views v = views(B);
        l.info = B.I;
        l.N = N;
        l.W = W;
        // compressed IJK/NBR (:1044-1126)
        DBuf<int> head, rowId, firstIn;
        BCHK(head.alloc(N + 1));
        BCHK(rowId.alloc(N + 1));
        BCHK(firstIn.alloc(N));
        {
            int *h = head.p, *fi = firstIn.p;
            for_n(N + 1, s, [=] __device__(long long cl) {
                const int c = (int)cl;
                if (c == v.N) { h[c] = 0; return; }
                bool hd = true;
                if (c > 0) {
                    const int jk = v.jPlus[v.kPlus[c]], jkp = v.jPlus[v.kPlus[c - 1]];
                    hd = v.i[c] != v.i[c - 1] + 1 || v.j[c] != v.j[c - 1] || v.k[c] != v.k[c - 1] ||
                         v.jPlus[c] != v.jPlus[c - 1] + 1 || v.kPlus[c] != v.kPlus[c - 1] + 1 || jk != jkp + 1;
                }
                h[c] = hd ? 1 : 0;
                fi[c] = hd ? c : 0;
            });
            if (exclusive_sum(sc, head.p, rowId.p, (long long)N + 1, s)) return LBM3_E_HIP;
            DBuf<int> fmax;
            BCHK(fmax.alloc(N));
            if (inclusive_max(sc, firstIn.p, fmax.p, N, s)) return LBM3_E_HIP;
            std::swap(firstIn.p, fmax.p);
        }
        l.nRows = read_one(rowId.p + N, s);
        BCHK(l.shifter.alloc(N));
        BCHK(l.ri.alloc(l.nRows)); BCHK(l.rj.alloc(l.nRows)); BCHK(l.rk.alloc(l.nRows));
        BCHK(l.rjp.alloc(l.nRows)); BCHK(l.rkp.alloc(l.nRows)); BCHK(l.rjkp.alloc(l.nRows));
        {
            const int *h = head.p, *rid = rowId.p, *fi = firstIn.p;
            int *sh = l.shifter.p, *ri = l.ri.p, *rj = l.rj.p, *rk = l.rk.p, *rjp = l.rjp.p, *rkp = l.rkp.p, *rjkp = l.rjkp.p;
            for_n(N, s, [=] __device__(long long cl) {
                const int c = (int)cl;
                if (h[c]) {
                    const int r = rid[c];
                    sh[c] = r;
                    ri[r] = v.i[c]; rj[r] = v.j[c]; rk[r] = v.k[c];
                    rjp[r] = v.jPlus[c]; rkp[r] = v.kPlus[c]; rjkp[r] = v.jPlus[v.kPlus[c]];
                } else sh[c] = fi[c] - c;
            });
        }
        // wallMap, wallData, packed link codes (:1127-1180, packWallData :662-695)
        BCHK(l.wallMap.alloc(N));
        BCHK(l.wallIndex.alloc(W));
        BCHK(l.wallData.alloc(W));
        BCHK(l.linkLength.alloc(7LL * W));
        {
            int *wm = l.wallMap.p;
            const uint8_t *wall = B.wall.p, *ov = B.ifOverlap.p;
            for_n(N, s, [=] __device__(long long c) { wm[c] = wall[c] ? -3 : (ov[c] ? -2 : -1); });
            const int *wl = B.wallList.p, *wid = B.wallIDc.p;
            int *wi = l.wallIndex.p;
            uint32_t *wd = l.wallData.p, *lk = l.linkLength.p;
            const uint8_t *le = B.linkExist.p;
            const float *ll = B.linkLen.p;
            for_n(W, s, [=] __device__(long long wq) {
                const int w = (int)wq, c = wl[w];
                wi[w] = c;
                wm[c] = w;
                uint32_t data = (ov[c] ? 1u : 0u) | ((uint32_t)wid[w] << 27);
                uint32_t packed[7] = {0u, 0u, 0u, 0u, 0u, 0u, 0u};
                int li = 0;
                for (int d = 1; d < 27; ++d) {
                    if (!le[(size_t)d * W + w]) continue;
                    data |= 1u << d;
                    float q = ll[(size_t)d * W + w];
                    q = q < 0.f ? 0.f : (q > 1.f ? 1.f : q);
                    packed[li >> 2] |= (uint32_t)(q * 254.f + 0.5f) << (8 * (li & 3));
                    ++li;
                }
                wd[w] = data;
                for (int p = 0; p < 7; ++p) lk[(size_t)p * W + w] = packed[p];
            });
        }
        // getOpenBC at the boundary cells and their boundary projections
        DBuf<int> rm[4];
        DBuf<lbm3_bc> drec;
        if (const int rc = open_rule_maps(sim, B, L, dc, sc, rm, l.openBC)) return rc;
        BCHK(drec.alloc(l.openBC.empty() ? 1 : (long long)l.openBC.size()));
        if (!l.openBC.empty())
            BCHK(hipMemcpyAsync(drec.p, l.openBC.data(), sizeof(lbm3_bc) * l.openBC.size(), hipMemcpyHostToDevice, s));
        const lbm3_bc *rec = drec.p;
        const int *rSelf = rm[0].p, *rPx = rm[1].p, *rPy = rm[2].p, *rPz = rm[3].p;
        // periodic overlap: no force tracking near periodic boundaries (:1232-1274)
        {
            int *wm = l.wallMap.p;
            uint32_t *wd = l.wallData.p;
            const uint8_t *wall = B.wall.p;
            for_n(N, s, [=] __device__(long long c) {
                if (wall[c]) return;
                bool per = false;
                if (rPx[c] >= 0) per |= rec[rPx[c]].periodic != 0;
                if (rPy[c] >= 0) per |= rec[rPy[c]].periodic != 0;
                if (rPz[c] >= 0) per |= rec[rPz[c]].periodic != 0;
                if (!per) return;
                const int m = wm[c];
                if (m >= 0) wd[m] |= 1u;
                else if (m == -1) wm[c] = -2;
            });
        }
        // open boundaries and their sources (:1276-1484)
        {
            DBuf<int> bcid, rule;
            BCHK(bcid.alloc(N));
            BCHK(rule.alloc(N));
            DBuf<int> mx;
            BCHK(mx.alloc(1));
            const int minus1 = -1;
            BCHK(hipMemcpyAsync(mx.p, &minus1, sizeof(int), hipMemcpyHostToDevice, s));
            int *bi = bcid.p, *ru = rule.p, *m = mx.p;
            const uint8_t *wall = B.wall.p;
            const lbm3_info Il = B.I;
            for_n(N, s, [=] __device__(long long c) {
                bi[c] = -1;
                ru[c] = -1;
                if (wall[c]) return;
                const int r = rSelf[c];
                if (r < 0) return;
                const lbm3_bc &b = rec[r];
                ru[c] = r;
                bi[c] = b.periodic ? (b.periodic ? -2 : -1) : b.openBCID;
                if (!b.periodic) atomicMax(m, b.openBCID);
            });
            l.nOpenIDs = read_one(mx.p, s) + 1;
            l.openCount.assign(l.nOpenIDs, 0);
            std::vector<DBuf<int>> lists(l.nOpenIDs);
            DBuf<uint8_t> fl;
            BCHK(fl.alloc(N));
            l.nOpen = 0;
            for (int id = 0; id < l.nOpenIDs; ++id) {
                uint8_t *f = fl.p;
                for_n(N, s, [=] __device__(long long c) { f[c] = bi[c] == id; });
                l.openCount[id] = compact(sc, fl.p, N, lists[id], s);
                if (l.openCount[id] < 0) return LBM3_E_HIP;
                l.nOpen += l.openCount[id];
            }
            BCHK(l.openCell.alloc(l.nOpen));
            BCHK(l.openSource.alloc(l.nOpen));
            BCHK(l.openMeta.alloc(l.nOpen));
            long long off = 0;
            for (int id = 0; id < l.nOpenIDs; ++id) {
                if (l.openCount[id])
                    BCHK(hipMemcpyAsync(l.openCell.p + off, lists[id].p, sizeof(int) * (size_t)l.openCount[id], hipMemcpyDeviceToDevice, s));
                off += l.openCount[id];
            }
            DBuf<int> bad;
            BCHK(bad.alloc(1));
            BCHK(hipMemsetAsync(bad.p, 0, sizeof(int), s));
            const int *oc = l.openCell.p, *D = B.dense.p, *wmap = l.wallMap.p;
            const uint32_t *wdat = l.wallData.p;
            int *osrc = l.openSource.p, *bd = bad.p;
            uint32_t *meta = l.openMeta.p;
            for_n(l.nOpen, s, [=] __device__(long long x) {
                const int c = oc[x];
                const int i = v.i[c], j = v.j[c], k = v.k[c];
                const int X = Il.cellCountX, Y = Il.cellCountY, Z = Il.cellCountZ;
                const int nx = i == 0 ? -1 : (i == X - 1 ? 1 : 0), ny = j == 0 ? -1 : (j == Y - 1 ? 1 : 0), nz = k == 0 ? -1 : (k == Z - 1 ? 1 : 0);
                const int faces = (nx != 0) + (ny != 0) + (nz != 0);
                const int r = ru[c];
                const lbm3_bc &b = rec[r];
                const bool edge = b.dirichletU && faces == 2;
                int pref = 0;
                if (edge) {
                    if (nx == 0) pref = i <= X - 1 - i ? 1 : -1;
                    else if (ny == 0) pref = j <= Y - 1 - j ? 1 : -1;
                    else pref = k <= Z - 1 - k ? 1 : -1;
                }
                int src = -1;
                for (int cand = 0; cand < (edge ? 3 : 1); ++cand) {
                    int is = i - nx, js = j - ny, ks = k - nz;
                    if (cand > 0) {
                        const int sh = cand == 1 ? pref : -pref;
                        if (nx == 0) is += sh; else if (ny == 0) js += sh; else ks += sh;
                    }
                    if (is <= 0 || is >= X - 1 || js <= 0 || js >= Y - 1 || ks <= 0 || ks >= Z - 1) continue;
                    const int f = D[((long long)ks * Y + js) * X + is];
                    if (f < 0) continue;
                    if (b.dirichletU && faces > 1 && wall[f]) continue;
                    src = f;
                    break;
                }
                osrc[x] = src;
                if (src < 0) atomicAdd(bd, 1);
                const int wm = wmap[c];
                bool track = wm != -2;
                if (wm >= 0 && (wdat[wm] & 1u)) track = false;
                meta[x] = open_meta(normal_index(nx, ny, nz), r, track);
            });
            if (read_one(bad.p, s) > 0) return set_error(LBM3_E_BUILD, "Building open BC failed: no valid source for some cells.");
            // periodic cells and sources (:1486-1564)
            uint8_t *f = fl.p;
            for_n(N, s, [=] __device__(long long c) { f[c] = bi[c] == -2; });
            l.nPeriodic = compact(sc, fl.p, N, l.perIndex, s);
            if (l.nPeriodic < 0) return LBM3_E_HIP;
            BCHK(l.perSource.alloc(l.nPeriodic));
            const int *pi = l.perIndex.p;
            int *ps = l.perSource.p;
            const int cpc = 1 << L;
            BCHK(hipMemsetAsync(bad.p, 0, sizeof(int), s));
            for_n(l.nPeriodic, s, [=] __device__(long long x) {
                const int c = pi[x];
                const int X = Il.cellCountX, Y = Il.cellCountY, Z = Il.cellCountZ;
                int is = v.i[c], js = v.j[c], ks = v.k[c];
                if (is == 0) is = X - 2 * cpc; else if (is == X - 1) is = 2 * cpc - 1;
                if (js == 0) js = Y - 2 * cpc; else if (js == Y - 1) js = 2 * cpc - 1;
                if (ks == 0) ks = Z - 2 * cpc; else if (ks == Z - 1) ks = 2 * cpc - 1;
                int src = (is >= 0 && is < X && js >= 0 && js < Y && ks >= 0 && ks < Z) ? D[((long long)ks * Y + js) * X + is] : -1;
                if (src >= 0 && wall[src]) src = -1;
                ps[x] = src;
                if (src < 0) atomicAdd(bd, 1);
            });
            if (read_one(bad.p, s) > 0)
                return set_error(LBM3_E_BUILD, "Building periodic BC failed, some source cells do not exist or are marked as wall.");
        }
        // wall count and the level's voxelizer result
        {
            DBuf<int> cnt;
            BCHK(cnt.alloc(1));
            BCHK(hipMemsetAsync(cnt.p, 0, sizeof(int), s));
            int *cp = cnt.p;
            const uint8_t *wall = B.wall.p;
            for_n(N, s, [=] __device__(long long c) { if (wall[c]) atomicAdd(cp, 1); });
            l.wallCount = read_one(cnt.p, s);
            fluid += (long long)(N - l.wallCount) << L;
        }
        std::swap(l.rayHitScan.p, B.total.scan.p); std::swap(l.rayHitScan.n, B.total.scan.n);
        std::swap(l.rayMap.p, B.total.map.p); std::swap(l.rayMap.n, B.total.map.n);
        if (const int rc = finish_level_setup(sim, L)) return rc;
    }
    // interfaces need both levels of a pair compressed
    for (int L = 0; L + 1 < levels; ++L)
        if (const int rc = build_interface_lists(sim, G, L, sc)) return rc;
    BCHK(hipStreamSynchronize(s));
    sim.fluidUpdates = fluid;
    if (fluidUpdates) *fluidUpdates = fluid;
    (void)C;
    return check_hip(hipGetLastError(), "builder");
}


// fillFineToCoarseInterface / fillCoarseToFineInterface (gridBuilderFunctions.h:741-1030)
int build_interface_lists(Sim &sim, std::vector<BLevel> &G, int L, Scratch &sc)
{
    hipStream_t s = sim.stream;
    BLevel &B = G[L], &F = G[L + 1];
    Level &l = sim.L[L];
    const int N = B.N, NF = F.N;
    const Views v = views(B), vf = views(F);
    // first child (all-even fine indices) of every coarse cell
    DBuf<int> child;
    BCHK(child.alloc(N));
    BCHK(hipMemsetAsync(child.p, 0xff, sizeof(int) * (size_t)N, s));
    {
        int *ch = child.p;
        const int *pm = F.parentMap.p;
        for_n(NF, s, [=] __device__(long long cf) {
            const int cc = pm[cf];
            if (cc < 0) return;
            if (vf.i[cf] % 2 == 0 && vf.j[cf] % 2 == 0 && vf.k[cf] % 2 == 0) ch[cc] = (int)cf;
        });
    }
    DBuf<uint8_t> mk;
    BCHK(mk.alloc(N));
    uint8_t *m = mk.p;
    const uint8_t *wall = B.wall.p, *f2c = B.f2c.p, *c2f = B.c2f.p;
    // fine-to-coarse: every non-wall F->C cell with its first child
    for_n(N, s, [=] __device__(long long c) { m[c] = f2c[c] && !wall[c]; });
    l.F2C.interfaceCount = compact(sc, mk.p, N, l.F2C.index, s);
    if (l.F2C.interfaceCount < 0) return LBM3_E_HIP;
    BCHK(l.F2C.childMap.alloc(l.F2C.interfaceCount));
    {
        const int *ix = l.F2C.index.p, *ch = child.p;
        int *cm = l.F2C.childMap.p;
        for_n(l.F2C.interfaceCount, s, [=] __device__(long long x) { cm[x] = ch[ix[x]]; });
    }
    // coarse-to-fine: full 2x2x2 coarse blocks whose 8 fine cells at the block centre exist
    for_n(N, s, [=] __device__(long long c) { m[c] = c2f[c] && !wall[c]; });
    DBuf<uint8_t> full;
    BCHK(full.alloc(N));
    {
        uint8_t *fu = full.p;
        const int *ch = child.p, *FJM = F.jMinus.p, *FKM = F.kMinus.p;
        for_n(N, s, [=] __device__(long long cl) {
            const int c = (int)cl;
            fu[c] = 0;
            if (!m[c]) return;
            int nb[8];
            nb[0] = c; nb[2] = v.jPlus[c]; nb[4] = v.kPlus[c]; nb[6] = v.jPlus[nb[4]];
            nb[1] = c + 1 >= v.N ? 0 : c + 1; nb[3] = nb[2] + 1 >= v.N ? 0 : nb[2] + 1;
            nb[5] = nb[4] + 1 >= v.N ? 0 : nb[4] + 1; nb[7] = nb[6] + 1 >= v.N ? 0 : nb[6] + 1;
            for (int q = 1; q < 8; ++q) if (!m[nb[q]]) return;
            const int ic = v.i[c], jc = v.j[c], kc = v.k[c];
            for (int q = 1; q < 8; ++q)
                if (v.i[nb[q]] != ic + (q & 1) || v.j[nb[q]] != jc + ((q >> 1) & 1) || v.k[nb[q]] != kc + ((q >> 2) & 1)) return;
            const int top = ch[nb[7]];
            if (top < 0) return;
            int fn[8];
            fn[7] = top;
            fn[6] = top - 1 < 0 ? vf.N - 1 : top - 1;
            fn[4] = FJM[fn[6]];
            fn[0] = FKM[fn[4]];
            fn[2] = vf.jPlus[fn[0]];
            fn[1] = fn[0] + 1 >= vf.N ? 0 : fn[0] + 1;
            fn[3] = fn[2] + 1 >= vf.N ? 0 : fn[2] + 1;
            fn[5] = fn[4] + 1 >= vf.N ? 0 : fn[4] + 1;
            for (int q = 0; q < 8; ++q)
                if (vf.i[fn[q]] != 2 * ic + 1 + (q & 1) || vf.j[fn[q]] != 2 * jc + 1 + ((q >> 1) & 1) ||
                    vf.k[fn[q]] != 2 * kc + 1 + ((q >> 2) & 1)) return;
            fu[c] = 1;
        });
    }
    l.C2F.interfaceCount = compact(sc, full.p, N, l.C2F.index, s);
    if (l.C2F.interfaceCount < 0) return LBM3_E_HIP;
    BCHK(l.C2F.childMap.alloc(l.C2F.interfaceCount));
    DBuf<uint8_t> recv;
    BCHK(recv.alloc(NF));
    BCHK(hipMemsetAsync(recv.p, 0, (size_t)NF, s));
    {
        const int *ix = l.C2F.index.p, *ch = child.p, *FJM = F.jMinus.p, *FKM = F.kMinus.p;
        int *cm = l.C2F.childMap.p;
        uint8_t *rc = recv.p;
        for_n(l.C2F.interfaceCount, s, [=] __device__(long long x) {
            const int c = ix[x];
            const int kp = v.kPlus[c], jkp = v.jPlus[kp], ijkp = jkp + 1 >= v.N ? 0 : jkp + 1;
            const int top = ch[ijkp];
            const int jk = top - 1 < 0 ? vf.N - 1 : top - 1;
            const int self = FKM[FJM[jk]];
            cm[x] = self;
            int fnb[8];
            fnb[0] = self; fnb[2] = vf.jPlus[self]; fnb[4] = vf.kPlus[self]; fnb[6] = vf.jPlus[fnb[4]];
            fnb[1] = self + 1 >= vf.N ? 0 : self + 1; fnb[3] = fnb[2] + 1 >= vf.N ? 0 : fnb[2] + 1;
            fnb[5] = fnb[4] + 1 >= vf.N ? 0 : fnb[4] + 1; fnb[7] = fnb[6] + 1 >= vf.N ? 0 : fnb[6] + 1;
            for (int q = 0; q < 8; ++q) rc[fnb[q]] = 1;
        });
    }
    // leftovers: fine cells that need coarse values but sit in no full block
    DBuf<uint8_t> lo;
    BCHK(lo.alloc(NF));
    {
        uint8_t *lp = lo.p;
        const uint8_t *nc = F.needCoarse.p, *fw = F.wall.p, *rc = recv.p;
        for_n(NF, s, [=] __device__(long long c) { lp[c] = nc[c] && !fw[c] && !rc[c]; });
    }
    l.C2F.leftoverCount = compact(sc, lo.p, NF, l.C2F.leftoverIndex, s);
    if (l.C2F.leftoverCount < 0) return LBM3_E_HIP;
    const int nl = l.C2F.leftoverCount;
    BCHK(l.C2F.leftoverParent.alloc(nl));
    BCHK(l.C2F.leftoverNbrI.alloc(nl));
    BCHK(l.C2F.leftoverNbrJ.alloc(nl));
    BCHK(l.C2F.leftoverNbrK.alloc(nl));
    {
        const int *li = l.C2F.leftoverIndex.p, *pm = F.parentMap.p, *JM = B.jMinus.p, *KM = B.kMinus.p;
        int *lpar = l.C2F.leftoverParent.p, *nI = l.C2F.leftoverNbrI.p, *nJ = l.C2F.leftoverNbrJ.p, *nK = l.C2F.leftoverNbrK.p;
        for_n(nl, s, [=] __device__(long long x) {
            const int cf = li[x], cc = pm[cf];
            lpar[x] = cc;
            const int ic = v.i[cc], jc = v.j[cc], kc = v.k[cc];
            auto match = [&](int n, int di, int dj, int dk) {
                return m[n] && v.i[n] == ic + di && v.j[n] == jc + dj && v.k[n] == kc + dk;
            };
            auto pick = [&](bool plusFirst, int plus, int minus, int dx, int dy, int dz) {
                int r = cc;
                if (plusFirst) {
                    if (match(plus, dx, dy, dz)) r = plus;
                    else if (match(minus, -dx, -dy, -dz)) r = -minus - 1;
                } else {
                    if (match(minus, -dx, -dy, -dz)) r = minus;
                    else if (match(plus, dx, dy, dz)) r = -plus - 1;
                }
                return r;
            };
            const int ip = cc + 1 >= v.N ? 0 : cc + 1, im = cc - 1 < 0 ? v.N - 1 : cc - 1;
            nI[x] = pick(vf.i[cf] % 2 == 1, ip, im, 1, 0, 0);
            nJ[x] = pick(vf.j[cf] % 2 == 1, v.jPlus[cc], JM[cc], 0, 1, 0);
            nK[x] = pick(vf.k[cf] % 2 == 1, v.kPlus[cc], KM[cc], 0, 0, 1);
        });
    }
    return check_hip(hipGetLastError(), "interface lists");
}

} // namespace

// ---------------------------------------------------------------------------
// the build
// ---------------------------------------------------------------------------
int gpu_build(Sim &sim, const float *tri, const int *triCounts, int nBodies, lbm3_bounds domain, long long *fluidUpdates)
{
    const lbm3_case &C = sim.C;
    const int levels = C.gridLevelCount;
    hipStream_t s = sim.stream;
    Scratch sc;
    std::vector<BLevel> G(levels);
    init_info(C, G, domain);
    // STL bodies
    std::vector<DevSTL> stl(nBodies);
    {
        long long off = 0;
        sim.stlEdges.assign((size_t)nBodies, {});
        for (int b = 0; b < nBodies; ++b) {
            const int rc = stl_upload(stl[b], tri + 9 * off, triCounts[b], C.resGlobal, s);
            sim.stlEdges[(size_t)b] = stl[b].edges;       // kept also when the check refuses the body
            if (rc) return rc;
            off += triCounts[b];
        }
    }
    if (nBodies > LBM3_MAX_BODIES) return set_error(LBM3_E_ARG, "too many STL bodies");
    // voxelizers per level
    for (int L = 0; L < levels; ++L) {
        G[L].maps.resize(nBodies);
        for (int b = 0; b < nBodies; ++b)
            if (const int rc = voxelize(G[L].maps[b], stl[b], G[L].I, sc, s)) return rc;
        RayMap &T = G[L].total;
        if (nBodies > 0) {
            if (raymap_copy(T, G[L].maps[0], s)) return LBM3_E_HIP;
        } else {
            T.nRays = G[L].I.cellCountX * G[L].I.cellCountY;
            T.total = 0;
            BCHK(T.scan.alloc(T.nRays + 1));
            BCHK(hipMemsetAsync(T.scan.p, 0, sizeof(long long) * (size_t)(T.nRays + 1), s));
        }
        for (int b = 1; b < nBodies; ++b)
            if (raymap_union(T, G[L].maps[b], sc, s)) return LBM3_E_HIP;
    }
    Views2 rv{};
    for (int L = 0; L < levels; ++L) rv.r[L] = RayView{G[L].total.scan.p, G[L].total.map.p};
    DBuf<lbm3_case> dcase;
    BCHK(dcase.alloc(1));
    BCHK(hipMemcpyAsync(dcase.p, &C, sizeof(lbm3_case), hipMemcpyHostToDevice, s));
    const lbm3_case *dc = dcase.p;

    // ---- buildIJKFull (:299-400) ----
    for (int L = 0; L < levels; ++L) {
        BLevel &B = G[L];
        const bool finest = L == levels - 1, coarsest = L == 0;
        const long long box = (long long)B.I.cellCountX * B.I.cellCountY * B.I.cellCountZ;
        DBuf<uint8_t> exists;
        BCHK(exists.alloc(box));
        if (coarsest) {
            // skeleton keep cells: any finer fluid on any level, then one dense spread
            const lbm3_info Sk = B.Sk;
            DBuf<uint8_t> m;
            BCHK(m.alloc(Sk.cellCount));
            BCHK(B.skKeep.alloc(Sk.cellCount));
            uint8_t *mp = m.p, *kp = B.skKeep.p;
            const int lev = levels;
            for_n(Sk.cellCount, s, [=] __device__(long long c) {
                const int XY = Sk.cellCountX * Sk.cellCountY;
                const int k = (int)(c / XY), rem = (int)(c % XY), j = rem / Sk.cellCountX, i = rem % Sk.cellCountX;
                bool any = false;
                for (int l = 0; l < lev && !any; ++l) any = finer_fluid(rv.r[l], Sk.cellCountX, i, j, k, 1 << (l + 1));
                mp[c] = any;
            });
            for_n(Sk.cellCount, s, [=] __device__(long long c) {
                const int X = Sk.cellCountX, Y = Sk.cellCountY, Z = Sk.cellCountZ;
                if (mp[c]) { kp[c] = 1; return; }
                const int k = (int)(c / (X * Y)), rem = (int)(c % (X * Y)), j = rem / X, i = rem % X;
                uint8_t r = 0;
                for (int dk = -1; dk <= 1 && !r; ++dk)
                    for (int dj = -1; dj <= 1 && !r; ++dj)
                        for (int di = -1; di <= 1 && !r; ++di) {
                            if (!dk && !dj && !di) continue;
                            const int a = i + di, b = j + dj, d = k + dk;
                            if (a < 0 || a >= X || b < 0 || b >= Y || d < 0 || d >= Z) continue;
                            r = mp[((long long)d * Y + b) * X + a];
                        }
                kp[c] = r;
            });
            uint8_t *ex = exists.p;
            const int X = B.I.cellCountX, XY = B.I.cellCountX * B.I.cellCountY, SX = Sk.cellCountX, SXY = Sk.cellCountX * Sk.cellCountY;
            for_n(box, s, [=] __device__(long long p) {
                const int k = (int)(p / XY), j = (int)((p % XY) / X), i = (int)(p % X);
                ex[p] = kp[(long long)(k / 2) * SXY + (long long)(j / 2) * SX + i / 2];
            });
        } else {
            const BLevel &P = G[L - 1];
            uint8_t *ex = exists.p;
            const int *D = P.dense.p;
            const uint8_t *ref = P.refine.p;
            const int X = B.I.cellCountX, XY = B.I.cellCountX * B.I.cellCountY, PX = P.I.cellCountX, PXY = P.I.cellCountX * P.I.cellCountY;
            for_n(box, s, [=] __device__(long long p) {
                const int k = (int)(p / XY), j = (int)((p % XY) / X), i = (int)(p % X);
                const int pc = D[(long long)(k / 2) * PXY + (long long)(j / 2) * PX + i / 2];
                ex[p] = pc >= 0 && ref[pc];
            });
        }
        if (const int rc = cells_from_dense(B, exists.p, sc, s)) return rc;
        exists.release();
        BCHK(B.keep.alloc(B.N));
        if (!finest) { BCHK(B.refine.alloc(B.N)); BCHK(B.deep.alloc(B.N)); BCHK(B.f2c.alloc(B.N)); BCHK(B.c2f.alloc(B.N)); }
        if (!coarsest) { BCHK(B.parentIf.alloc(B.N)); BCHK(B.parentF2C.alloc(B.N)); BCHK(B.needCoarse.alloc(B.N)); }
        if (const int rc = build_nbr(B, s)) return rc;
        if (!coarsest) {
            const BLevel &P = G[L - 1];
            if (parent_map(B, P, s)) return LBM3_E_HIP;
            const int *pm = B.parentMap.p;
            const uint8_t *pc2f = P.c2f.p, *pf2c = P.f2c.p;
            uint8_t *pif = B.parentIf.p, *pf = B.parentF2C.p;
            for_n(B.N, s, [=] __device__(long long c) {
                const int p = pm[c];
                pif[c] = (pc2f[p] + pf2c[p]) != 0;
                pf[c] = pf2c[p];
            });
        }
        if (!finest) {
            // markRefinementCells on this level (markerFunctions.h:401-450)
            const Views v = views(B);
            const int L0 = L, lev = levels, Xl = B.I.cellCountX;
            DBuf<uint8_t> m, buf;
            BCHK(m.alloc(B.N));
            BCHK(buf.alloc(B.N));
            uint8_t *mp = m.p, *bp = buf.p, *keep = B.keep.p, *deep = B.deep.p, *f2c = B.f2c.p, *c2f = B.c2f.p, *ref = B.refine.p;
            const uint8_t *pif = B.parentIf.p;
            for_n(B.N, s, [=] __device__(long long c) {
                bool any = false;
                for (int l = L0; l < lev && !any; ++l) any = finer_fluid(rv.r[l], Xl, v.i[c], v.j[c], v.k[c], 1 << (l - L0));
                mp[c] = any;
            });
            if (spread(B, keep, mp, s)) return LBM3_E_HIP;
            if (!coarsest) for_n(B.N, s, [=] __device__(long long c) { keep[c] = keep[c] | pif[c]; });
            for_n(B.N, s, [=] __device__(long long c) {
                bool any = false;
                for (int l = L0; l < lev && !any; ++l) any = finer_wall(rv.r[l], Xl, v.i[c], v.j[c], v.k[c], 1 << (l - L0));
                deep[c] = any;
            });
            for (int sp = 0; sp < C.wallRefinementCount; ++sp) {
                BCHK(hipMemcpyAsync(bp, deep, (size_t)B.N, hipMemcpyDeviceToDevice, s));
                if (spread(B, deep, bp, s)) return LBM3_E_HIP;
            }
            const lbm3_info Il = B.I;
            // getRefinementModifier: the case's own hook (both answers per cell,
            // for an incoming marker false and true) or its rule table
            DBuf<uint8_t> hookOut;
            if (sim.hooks.refinement) {
                BCHK(hookOut.alloc(B.N > 0 ? B.N : 1));
                lbm3_info info = Il;
                if (sim.hooks.refinement(sim.hooks.user, &info, v.i, v.j, v.k, B.N, hookOut.p, (void *)s) != 0)
                    return set_error(LBM3_E_ARG, "the getRefinementModifier hook callback failed");
                BCHK(hipGetLastError());
            }
            const uint8_t *ho = hookOut.p;
            for_n(B.N, s, [=] __device__(long long c) {
                const float x = (float)v.i[c] * Il.res + Il.ox, y = (float)v.j[c] * Il.res + Il.oy, z = (float)v.k[c] * Il.res + Il.oz;
                bool d = ho ? ((ho[c] >> (deep[c] ? 1 : 0)) & 1u) != 0 : refine_hook(*dc, L0, x, y, z, deep[c] != 0);
                if (L0 > 0) d = d && !pif[c];
                deep[c] = d && keep[c];
            });
            if (spread(B, f2c, deep, s)) return LBM3_E_HIP;
            for_n(B.N, s, [=] __device__(long long c) { f2c[c] = f2c[c] && keep[c] && !deep[c]; });
            if (spread(B, bp, f2c, s)) return LBM3_E_HIP;
            if (spread(B, c2f, bp, s)) return LBM3_E_HIP;
            for_n(B.N, s, [=] __device__(long long c) {
                c2f[c] = c2f[c] && keep[c] && !deep[c] && !f2c[c];
                ref[c] = deep[c] || f2c[c] || c2f[c];
            });
        }
    }

    // ---- deleteExcessCells (:402-588) ----
    for (int L = 0; L < levels; ++L) {
        BLevel &B = G[L];
        const bool finest = L == levels - 1, coarsest = L == 0;
        const int N = B.N;
        if (!coarsest) {
            if (parent_map(B, G[L - 1], s)) return LBM3_E_HIP;
            const int *pm = B.parentMap.p;
            uint8_t *pif = B.parentIf.p;
            for_n(N, s, [=] __device__(long long c) { pif[c] = pm[c] >= 0; });
        }
        const Views v = views(B);
        const lbm3_info Il = B.I;
        const RayView rt = rv.r[L];
        uint8_t *keep = B.keep.p;
        const uint8_t *pif = B.parentIf.p;
        DBuf<uint8_t> src;
        BCHK(src.alloc(N));
        uint8_t *sp = src.p;
        for_n(N, s, [=] __device__(long long c) {
            bool fluid = !in_wall(rt, Il.cellCountX * v.j[c] + v.i[c], v.k[c]);
            if (!coarsest) fluid = fluid || pif[c];
            sp[c] = fluid;
        });
        if (spread(B, keep, sp, s)) return LBM3_E_HIP;
        if (!finest) {
            const uint8_t *deep = B.deep.p;
            for_n(N, s, [=] __device__(long long c) { keep[c] = keep[c] && !deep[c]; });
        }
        if (!coarsest) {
            DBuf<uint8_t> l3;
            BCHK(l3.alloc(N));
            if (spread(B, l3.p, B.parentF2C.p, s)) return LBM3_E_HIP;
            BCHK(hipMemcpyAsync(sp, l3.p, (size_t)N, hipMemcpyDeviceToDevice, s));
            if (spread(B, l3.p, sp, s)) return LBM3_E_HIP;
            BCHK(hipMemcpyAsync(sp, l3.p, (size_t)N, hipMemcpyDeviceToDevice, s));
            if (spread(B, l3.p, sp, s)) return LBM3_E_HIP;
            uint8_t *nc = B.needCoarse.p;
            const uint8_t *l3p = l3.p;
            for_n(N, s, [=] __device__(long long c) {
                if (pif[c] && !l3p[c]) keep[c] = 0;
                nc[c] = pif[c] && keep[c];
            });
            if (spread(B, sp, B.parentF2C.p, s)) return LBM3_E_HIP;
            for_n(N, s, [=] __device__(long long c) { nc[c] = nc[c] && !sp[c]; });
        }
        // compaction of kept cells (cell order preserved)
        DBuf<int> kept;
        const int n = compact(sc, keep, N, kept, s);
        if (n < 0) return LBM3_E_HIP;
        auto gather_i = [&](DBuf<int> &a) -> int {
            DBuf<int> t;
            BCHK(t.alloc(n));
            int *tp = t.p;
            const int *ap = a.p, *kp = kept.p;
            for_n(n, s, [=] __device__(long long c) { tp[c] = ap[kp[c]]; });
            std::swap(a.p, t.p); std::swap(a.n, t.n);
            return 0;
        };
        auto gather_b = [&](DBuf<uint8_t> &a) -> int {
            DBuf<uint8_t> t;
            BCHK(t.alloc(n));
            uint8_t *tp = t.p;
            const uint8_t *ap = a.p;
            const int *kp = kept.p;
            for_n(n, s, [=] __device__(long long c) { tp[c] = ap[kp[c]]; });
            std::swap(a.p, t.p); std::swap(a.n, t.n);
            return 0;
        };
        if (gather_i(B.i) || gather_i(B.j) || gather_i(B.k)) return LBM3_E_HIP;
        if (!coarsest && (gather_i(B.parentMap) || gather_b(B.needCoarse))) return LBM3_E_HIP;
        if (!finest && (gather_b(B.f2c) || gather_b(B.c2f))) return LBM3_E_HIP;
        B.N = n;
        B.I.cellCount = n;
        B.skKeep.release();
        B.deep.release();
        B.refine.release();
        BCHK(B.wall.alloc(n));
        BCHK(B.ifOverlap.alloc(n));
        BCHK(hipMemsetAsync(B.ifOverlap.p, 0, (size_t)n, s));
        if (rebuild_dense(B, s) || build_nbr(B, s)) return LBM3_E_HIP;
        BCHK(B.jMinus.alloc(n)); BCHK(B.kMinus.alloc(n));
        {
            int *jm = B.jMinus.p, *km = B.kMinus.p;
            const int *jp = B.jPlus.p, *kp = B.kPlus.p;
            for_n(n, s, [=] __device__(long long c) { jm[jp[c]] = (int)c; km[kp[c]] = (int)c; });
        }
        // bounds of the kept cells
        {
            DBuf<int> mm;
            BCHK(mm.alloc(6));
            const int init[6] = {0x7fffffff, -0x7fffffff, 0x7fffffff, -0x7fffffff, 0x7fffffff, -0x7fffffff};
            BCHK(hipMemcpyAsync(mm.p, init, sizeof(init), hipMemcpyHostToDevice, s));
            int *m = mm.p;
            const int *I = B.i.p, *J = B.j.p, *K = B.k.p;
            for_n(n, s, [=] __device__(long long c) {
                atomicMin(&m[0], I[c]); atomicMax(&m[1], I[c]);
                atomicMin(&m[2], J[c]); atomicMax(&m[3], J[c]);
                atomicMin(&m[4], K[c]); atomicMax(&m[5], K[c]);
            });
            int h[6];
            BCHK(hipMemcpyAsync(h, mm.p, sizeof(h), hipMemcpyDeviceToHost, s));
            BCHK(hipStreamSynchronize(s));
            lbm3_info &I2 = B.I;
            I2.Bounds.xMin = I2.ox + I2.res * (float)h[0] - 0.5f * I2.res;
            I2.Bounds.xMax = I2.ox + I2.res * (float)h[1] + 0.5f * I2.res;
            I2.Bounds.yMin = I2.oy + I2.res * (float)h[2] - 0.5f * I2.res;
            I2.Bounds.yMax = I2.oy + I2.res * (float)h[3] + 0.5f * I2.res;
            I2.Bounds.zMin = I2.oz + I2.res * (float)h[4] - 0.5f * I2.res;
            I2.Bounds.zMax = I2.oz + I2.res * (float)h[5] + 0.5f * I2.res;
        }
    }

    return walls_and_grid(sim, G, stl, nBodies, dcase.p, sc, fluidUpdates);
}

// ---------------------------------------------------------------------------
// buildRotors (rotorFunctions.h:315-585): each rotor STL voxelized at half the
// level spacing; 4x4x4 blocks of interpolation cells marked where their 6^3
// voxel neighbourhood holds a solid voxel; per interpolation cell the 8 corner
// counts (solid voxels among the 2x2x2 around each corner) packed 4 bits each
// ---------------------------------------------------------------------------
namespace {

struct RotorGrid {
    int VX, VY, bx, by;
    const long long *scan;
    const int *map;
    // voxel (i, j, k) inside: a hit pair [kStart, kEnd) of ray (i, j) holds k
    // (rotorFunctions.h:384-404; pairs read only while both entries exist)
    __device__ bool solid(int i, int j, int k) const
    {
        if (i < 0 || i >= VX || j < 0 || j >= VY) return false;
        const int ray = VX * j + i;
        const long long e = scan[ray + 1];
        for (long long h = scan[ray]; h + 1 < e; h += 2) {
            if (map[h] > k) break;
            if (map[h + 1] > k) return true;
        }
        return false;
    }
};

int build_rotor(Sim &sim, Level &l, const float *tri, int n, Level::Rotor &Rt, Scratch &sc, hipStream_t s)
{
    DevSTL S;
    if (const int rc = stl_upload(S, tri, n, sim.C.resGlobal, s)) return rc;
    lbm3_info V;
    memset(&V, 0, sizeof V);
    V.ox = S.B.xMin; V.oy = S.B.yMin; V.oz = S.B.zMin;
    V.res = l.info.res * 0.5f;
    V.cellCountX = (int)((S.B.xMax - S.B.xMin) / V.res) + 2;
    V.cellCountY = (int)((S.B.yMax - S.B.yMin) / V.res) + 2;
    V.cellCountZ = (int)((S.B.zMax - S.B.zMin) / V.res) + 2;
    RayMap R;
    if (const int rc = voxelize(R, S, V, sc, s)) return rc;
    RotorDev &d = Rt.dev;
    d = RotorDev{};
    d.axis = LBM3_AXIS_NONE;
    d.res = V.res;
    d.xMin = V.ox - 1.5f * V.res;
    d.yMin = V.oy - 1.5f * V.res;
    d.zMin = V.oz - 1.5f * V.res;
    d.cx = ((V.cellCountX + 2 + 3) / 4) * 4;
    d.cy = ((V.cellCountY + 2 + 3) / 4) * 4;
    d.cz = ((V.cellCountZ + 2 + 3) / 4) * 4;
    d.xMax = d.xMin + d.res * (float)d.cx;
    d.yMax = d.yMin + d.res * (float)d.cy;
    d.zMax = d.zMin + d.res * (float)d.cz;
    const int bx = d.cx / 4, by = d.cy / 4, bz = d.cz / 4;
    const long long blocks = (long long)bx * by * bz;
    if (blocks > (1LL << 30)) return set_error(LBM3_E_MEMORY, "rotor interpolation grid too large");
    Rt.blockCount = (int)blocks;
    const RotorGrid G{V.cellCountX, V.cellCountY, bx, by, R.scan.p, R.map.p};
    DBuf<uint8_t> mark;
    BCHK(mark.alloc(blocks > 0 ? blocks : 1));
    uint8_t *mk = mark.p;
    for_n(blocks, s, [=] __device__(long long b) {
        const int kb = int(b / (G.bx * G.by)), jb = int((b % (G.bx * G.by)) / G.bx), ib = int(b % G.bx);
        uint8_t hit = 0;
        for (int j = jb * 4 - 2; j < jb * 4 + 4 && !hit; ++j)
            for (int i = ib * 4 - 2; i < ib * 4 + 4 && !hit; ++i)
                for (int k = kb * 4 - 2; k < kb * 4 + 4; ++k)
                    if (G.solid(i, j, k)) { hit = 1; break; }
        mk[b] = hit;
    });
    const int nz = compact(sc, mark.p, blocks, Rt.index, s);
    if (nz < 0) return nz;
    d.nonZero = nz;
    BCHK(Rt.rotorMap.alloc(blocks > 0 ? blocks : 1));
    int *rm = Rt.rotorMap.p;
    const int *ix = Rt.index.p;
    for_n(blocks, s, [=] __device__(long long b) { rm[b] = -1; });
    for_n(nz, s, [=] __device__(long long x) { rm[ix[x]] = (int)x; });
    BCHK(Rt.interp.alloc(64LL * (nz > 0 ? nz : 1)));
    uint32_t *ip = Rt.interp.p;
    for_n(64LL * nz, s, [=] __device__(long long t) {
        const int x = int(t / 64), q = int(t % 64);
        const int b = ix[x];
        const int kb = b / (G.bx * G.by), jb = (b % (G.bx * G.by)) / G.bx, ib = b % G.bx;
        const int ki = q / 16, ji = (q / 4) % 4, ii = q % 4;
        uint32_t packed = 0u;
        int corner = 0;
        for (int kc = ki; kc <= ki + 1; ++kc)
            for (int jc = ji; jc <= ji + 1; ++jc)
                for (int ic = ii; ic <= ii + 1; ++ic) {
                    uint32_t count = 0u;
                    for (int km = kc; km <= kc + 1; ++km)
                        for (int jm = jc; jm <= jc + 1; ++jm)
                            for (int im = ic; im <= ic + 1; ++im)
                                count += G.solid(ib * 4 - 2 + im, jb * 4 - 2 + jm, kb * 4 - 2 + km) ? 1u : 0u;
                    packed |= count << (4u * (unsigned)corner++);
                }
        ip[t] = packed;
    });
    BCHK(Rt.g.alloc(81LL * (nz > 0 ? nz : 1)));
    BCHK(hipMemsetAsync(Rt.g.p, 0, sizeof(double) * (size_t)Rt.g.n, s));
    d.rotorMap = Rt.rotorMap.p;
    d.interp = Rt.interp.p;
    d.g = Rt.g.p;
    BCHK(hipStreamSynchronize(s));
    return 0;
}

} // namespace

int upload_rotors(Sim &sim, int level);   // capi.hip

int gpu_build_rotors(Sim &sim, const float *tri, const int *triCounts, int nRotors, int levelMask)
{
    Scratch sc;
    const hipStream_t s = sim.stream;
    for (int L = 0; L < sim.levels; ++L) {
        Level &l = sim.L[L];
        l.rotors.clear();
        l.info.useRotors = (levelMask >> L) & 1;
        if (!l.info.useRotors || nRotors == 0) {
            if (const int rc = upload_rotors(sim, L)) return rc;
            continue;
        }
        l.rotors.resize(nRotors);
        long long off = 0;
        for (int r = 0; r < nRotors; ++r) {
            if (const int rc = build_rotor(sim, l, tri + 9 * off, triCounts[r], l.rotors[r], sc, s)) return rc;
            off += triCounts[r];
        }
        if (const int rc = upload_rotors(sim, L)) return rc;
    }
    return 0;
}

} // namespace lbm3

using namespace lbm3;

extern "C" int lbm3_build_rotors(lbm3_sim *h, const float *triangles, const int *triCounts, int nRotors, int levelMask)
{
    if (!h || nRotors < 0 || nRotors > LBM3_MAX_ROTORS || (nRotors > 0 && (!triangles || !triCounts)))
        return set_error(LBM3_E_ARG, "bad rotor arguments");
    for (int L = 0; L < h->s.levels; ++L)
        if (((levelMask >> L) & 1) && !h->s.L[L].built) return set_error(LBM3_E_STATE, "rotors need the level built first");
    for (int r = 0; r < nRotors; ++r)
        if (triCounts[r] <= 0) return set_error(LBM3_E_ARG, "a rotor needs triangles");
    return gpu_build_rotors(h->s, triangles, triCounts, nRotors, levelMask);
}

extern "C" int lbm3_build_grids(lbm3_sim *h, const float *triangles, const int *triCounts, int nBodies, lbm3_bounds domain,
                                long long *fluidUpdatesPerIteration)
{
    if (!h || nBodies < 0 || (nBodies > 0 && (!triangles || !triCounts))) return set_error(LBM3_E_ARG, "bad build arguments");
    return gpu_build(h->s, triangles, triCounts, nBodies, domain, fluidUpdatesPerIteration);
}
