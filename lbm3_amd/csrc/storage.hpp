// Population storage in HBM: shifted populations f - w as float, half,
// bfloat16 or double.  Arithmetic always runs in the compute type (float for
// the 32/16-bit stores, double for the 64-bit store); conversion happens at the
// load and the store only.
#pragma once
#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>
#include <hip/hip_bf16.h>

// Population loads and stores of the 32/16-bit stores are ordinary cached
// accesses by default.  Non-temporal hints (-DLBM3_LOAD_NT=1 / -DLBM3_STORE_NT=1)
// were measured 1.47x slower on the 256x128x128 sphere bulk kernel: half of
// the planes are read and written one cell off the 128-byte line grid (the i+
// neighbour), and the neighbouring wave's half of each such line must still
// be in L2 when it arrives.
#ifndef LBM3_LOAD_NT
#define LBM3_LOAD_NT 0
#endif
#ifndef LBM3_STORE_NT
#define LBM3_STORE_NT 0
#endif

namespace lbm3 {

// Tiled population layout.  Cells are grouped in tiles of LBM3_TILE
// consecutive cells; a tile holds its 27 slot planes one after another, each
// LBM3_TILE entries long: slot (plane, cell) sits at
//   (cell / TILE) * 27 * TILE + plane * TILE + cell % TILE.
// A wave of 64 consecutive cells then reads and writes 27 runs of 64 entries
// inside one few-KiB block (plus the blocks of its i+/j+/k+ neighbours), not
// 27 runs spread over the whole level as with plain planes q * N + cell, which
// keeps DRAM pages and translation entries per wave few.  The level array
// holds 27 * padded_cells(N) entries; the padding cells are never touched.
#ifndef LBM3_TILE_CELLS
#define LBM3_TILE_CELLS 256
#endif
constexpr int LBM3_TILE = LBM3_TILE_CELLS;

__host__ __device__ __forceinline__ size_t tile_base(int cell)
{
    return (size_t)(cell / LBM3_TILE) * (27 * LBM3_TILE) + (size_t)(cell % LBM3_TILE);
}
__host__ __device__ __forceinline__ size_t pop_index(int plane, int cell)
{
    return tile_base(cell) + (size_t)plane * LBM3_TILE;
}
__host__ __device__ __forceinline__ long long padded_cells(long long n)
{
    return (n + LBM3_TILE - 1) / LBM3_TILE * LBM3_TILE;
}

// a level whose population array does not fit 32-bit byte offsets
template <class S> __host__ __device__ inline bool wide_level(long long n)
{
    return 27LL * padded_cells(n) * (long long)sizeof(S) > 0xFFFFFFFFLL;
}

template <class V> __device__ __forceinline__ V ld(const V *p)
{
#if LBM3_LOAD_NT
    return __builtin_nontemporal_load(p);
#else
    return *p;
#endif
}
template <class V> __device__ __forceinline__ void st(V *p, V v)
{
#if LBM3_STORE_NT
    __builtin_nontemporal_store(v, p);
#else
    *p = v;
#endif
}

// Encoding of the 16-bit stores: the stored value of direction q is
// (f_q - w_q - off_q) / w_q, i.e. the deviation of the shifted population from
// the shifted equilibrium of a reference velocity, in units of the weight.
// The 32/64-bit stores hold f_q - w_q and ignore it.
struct Codec {
    float off[27];
    float w[27];
    float iw[27];
};

template <class S> struct StoreTraits;

template <> struct StoreTraits<float> {
    using compute = float;
    static __device__ __forceinline__ float load(const float *p, int, const Codec &) { return ld(p); }
    static __device__ __forceinline__ void store(float *p, float v, int, const Codec &) { st(p, v); }
};

template <> struct StoreTraits<double> {
    using compute = double;
    static __device__ __forceinline__ double load(const double *p, int, const Codec &) { return *p; }
    static __device__ __forceinline__ void store(double *p, double v, int, const Codec &) { *p = v; }
};

template <> struct StoreTraits<__half> {
    using compute = float;
    static __device__ __forceinline__ float load(const __half *p, int q, const Codec &c)
    {
        const unsigned short bits = ld(reinterpret_cast<const unsigned short *>(p));
        return fmaf(__half2float(__ushort_as_half(bits)), c.w[q], c.off[q]);
    }
    static __device__ __forceinline__ void store(__half *p, float v, int q, const Codec &c)
    {
        st(reinterpret_cast<unsigned short *>(p), __half_as_ushort(__float2half_rn((v - c.off[q]) * c.iw[q])));
    }
};

template <> struct StoreTraits<__hip_bfloat16> {
    using compute = float;
    static __device__ __forceinline__ float load(const __hip_bfloat16 *p, int q, const Codec &c)
    {
        const unsigned short bits = ld(reinterpret_cast<const unsigned short *>(p));
        return fmaf(__uint_as_float(static_cast<unsigned int>(bits) << 16), c.w[q], c.off[q]);
    }
    static __device__ __forceinline__ void store(__hip_bfloat16 *p, float v, int q, const Codec &c)
    {
        // round to nearest even on the upper 16 bits (finite values)
        unsigned int u = __float_as_uint((v - c.off[q]) * c.iw[q]);
        u += 0x7FFFu + ((u >> 16) & 1u);
        st(reinterpret_cast<unsigned short *>(p), static_cast<unsigned short>(u >> 16));
    }
};

} // namespace lbm3
