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
