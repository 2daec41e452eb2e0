// Population storage in HBM: shifted populations f - w as float, half,
// bfloat16, a 16-bit float with 11 mantissa bits (f16c) or double.  Arithmetic always runs in the compute type (float for
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

// cells per open-boundary group (k_open_groups).  Measured on the 256x128x128
// sphere (FP32 whole step): 4 cells 206.8 us, 8 cells 205.2 us, 16 cells
// 210.9 us, 32 cells 214.2 us, 64 cells 231.4 us, one open cell per lane
// 208.9 us -- small groups keep the x-face rows' other cells in the bulk
// kernel, large groups share fewer lines with it
#ifndef LBM3_OPEN_GROUP
#define LBM3_OPEN_GROUP 8
#endif
// the same for 16-bit population storage: 16 cells, so that a group's piece
// of a population line is 32 bytes as with 8 FP32 cells.  Measured (FP16 /
// BF16 whole step, MLUPS): 8 cells 30,626 / 31,020, 16 cells 31,509 / 32,021,
// 32 cells 30,593 / 31,416
#ifndef LBM3_OPEN_GROUP16
#define LBM3_OPEN_GROUP16 16
#endif
__host__ __device__ constexpr int open_group_cells(int storageBytes)
{
    return storageBytes == 2 ? LBM3_OPEN_GROUP16 : LBM3_OPEN_GROUP;
}

__host__ __device__ __forceinline__ size_t tile_base(int cell)
{
    return (size_t)((unsigned)cell / LBM3_TILE) * (27 * LBM3_TILE) + (size_t)((unsigned)cell % LBM3_TILE);
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

// Raw buffer access for levels below 4 GiB: a wave-uniform buffer resource for
// the level array, a per-lane 32-bit byte offset and a wave-uniform byte offset
// per slot plane, so that the 54 population accesses of a cell need no vector
// address arithmetic (buffer_load/store ... offen with an SGPR soffset).
// Word 3 of the resource is the raw-buffer format of gfx9-class GPUs.
// `base` and `bytes` must be the same in every lane of the wave (a kernel
// argument, or the descriptor of the one domain a block belongs to): they are
// read from the first lane, so that the resource sits in scalar registers even
// where the compiler cannot prove them uniform (a descriptor read through a
// pointer), instead of every access looping over the wave's resources.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t pop_rsrc(const void *base, long long bytes)
{
    const unsigned long long p = reinterpret_cast<unsigned long long>(base);
    const unsigned lo = __builtin_amdgcn_readfirstlane(unsigned(p));
    const unsigned hi = __builtin_amdgcn_readfirstlane(unsigned(p >> 32));
    const int num = __builtin_amdgcn_readfirstlane(bytes >= 0xFFFFFFFFLL ? int(0xFFFFFFFFu) : int((unsigned)bytes));
    void *b = reinterpret_cast<void *>((static_cast<unsigned long long>(hi) << 32) | lo);
    return __builtin_amdgcn_make_buffer_rsrc(b, (short)0, num, 0x00020000);
}

template <class S> struct StoreTraits;

template <> struct StoreTraits<float> {
    using compute = float;
    static __device__ __forceinline__ float load(const float *p, int, const Codec &) { return ld(p); }
    static __device__ __forceinline__ void store(float *p, float v, int, const Codec &) { st(p, v); }
    static __device__ __forceinline__ float
    bload(__amdgpu_buffer_rsrc_t r, uint32_t voff, int soff, int q, const Codec &c)
    {
        (void)q; (void)c;
        return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, (int)voff, soff, 0));
    }
    static __device__ __forceinline__ void bstore(__amdgpu_buffer_rsrc_t r, uint32_t voff, int soff, float v, int q, const Codec &c)
    {
        (void)q; (void)c;
        __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(int, v), r, (int)voff, soff, 0);
    }
};

template <> struct StoreTraits<double> {
    using compute = double;
    static __device__ __forceinline__ double load(const double *p, int, const Codec &) { return *p; }
    static __device__ __forceinline__ void store(double *p, double v, int, const Codec &) { *p = v; }
    static __device__ __forceinline__ double
    bload(__amdgpu_buffer_rsrc_t r, uint32_t voff, int soff, int q, const Codec &c)
    {
        (void)q; (void)c;
        return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, (int)voff, soff, 0));
    }
    static __device__ __forceinline__ void bstore(__amdgpu_buffer_rsrc_t r, uint32_t voff, int soff, double v, int q, const Codec &c)
    {
        (void)q; (void)c;
        typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
        __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, v), r, (int)voff, soff, 0);
    }
};

// FP16 encode: 1 materialises the scaled float (an empty asm operand) so that
// it is converted by a plain f32->f16 conversion; 0 lets the compiler fuse
// the scaling and the conversion into v_fma_mixlo_f16, which writes half a
// register.  Measured on the 256x128x128 sphere FP16 bulk kernel: 92.8 ->
// 84.0 us (open cells in groups), 115.5 -> 103.1 us (open cells inside).  The
// decode keeps the fused v_fma_mix_f32: forcing a separate conversion there
// measured 160 us.
#ifndef LBM3_F16_CVT
#define LBM3_F16_CVT 1
#endif
__device__ __forceinline__ unsigned short f16_bits(float y)
{
#if LBM3_F16_CVT
    __asm__ volatile("" : "+v"(y));
#endif
    return __half_as_ushort(__float2half_rn(y));
}
__device__ __forceinline__ float f16_value(unsigned short bits) { return __half2float(__ushort_as_half(bits)); }

template <> struct StoreTraits<__half> {
    using compute = float;
    static __device__ __forceinline__ float load(const __half *p, int q, const Codec &c)
    {
        const unsigned short bits = ld(reinterpret_cast<const unsigned short *>(p));
        return fmaf(f16_value(bits), c.w[q], c.off[q]);
    }
    static __device__ __forceinline__ void store(__half *p, float v, int q, const Codec &c)
    {
        st(reinterpret_cast<unsigned short *>(p), f16_bits((v - c.off[q]) * c.iw[q]));
    }
    static __device__ __forceinline__ float
    bload(__amdgpu_buffer_rsrc_t r, uint32_t voff, int soff, int q, const Codec &c)
    {
        const unsigned short bits = (unsigned short)__builtin_amdgcn_raw_buffer_load_b16(r, (int)voff, soff, 0);
        return fmaf(f16_value(bits), c.w[q], c.off[q]);
    }
    static __device__ __forceinline__ void bstore(__amdgpu_buffer_rsrc_t r, uint32_t voff, int soff, float v, int q, const Codec &c)
    {
        __builtin_amdgcn_raw_buffer_store_b16((short)f16_bits((v - c.off[q]) * c.iw[q]), r, (int)voff, soff, 0);
    }
};

// f16c: a 16-bit float with one sign bit, 4 exponent bits (bias 15, no
// infinities: the largest finite value is just below 2) and 11 mantissa bits,
// subnormal below 2^-14 -- one mantissa bit more than IEEE half, paid for with
// the exponent range the deviations of the shifted encoding never use (|x| <
// 2).  Its bits are those of the float x * 2^-112 from bit 12 to bit 26, whose
// exponent field is then the 4-bit exponent: decoding shifts them back into a
// float and takes the 2^112 in the codec's weight (Codec::w holds w * 2^112,
// Codec::iw holds 2^-112 / w for this storage); encoding rounds the float to
// 11 mantissa bits (nearest, ties to even) and saturates at the largest
// finite value.  Both need float subnormals (the kernels keep them).
struct f16c {
    unsigned short bits;
};

__device__ __forceinline__ float f16c_decode(unsigned int h)
{
    return __uint_as_float(((h & 0x7FFFu) << 12) | ((h & 0x8000u) << 16));
}

__device__ __forceinline__ unsigned short f16c_encode(float y)
{
    const unsigned int u = __float_as_uint(y);
    const unsigned int mag = u & 0x7FFFFFFFu;
    const unsigned int r = min((mag + 0x7FFu + ((mag >> 12) & 1u)) >> 12, 0x7FFFu);
    return (unsigned short)(r | ((u >> 16) & 0x8000u));
}

template <> struct StoreTraits<f16c> {
    using compute = float;
    static __device__ __forceinline__ float load(const f16c *p, int q, const Codec &c)
    {
        const unsigned short bits = ld(reinterpret_cast<const unsigned short *>(p));
        return fmaf(f16c_decode(bits), c.w[q], c.off[q]);
    }
    static __device__ __forceinline__ void store(f16c *p, float v, int q, const Codec &c)
    {
        st(reinterpret_cast<unsigned short *>(p), f16c_encode((v - c.off[q]) * c.iw[q]));
    }
    static __device__ __forceinline__ float
    bload(__amdgpu_buffer_rsrc_t r, uint32_t voff, int soff, int q, const Codec &c)
    {
        const unsigned short bits = (unsigned short)__builtin_amdgcn_raw_buffer_load_b16(r, (int)voff, soff, 0);
        return fmaf(f16c_decode(bits), c.w[q], c.off[q]);
    }
    static __device__ __forceinline__ void bstore(__amdgpu_buffer_rsrc_t r, uint32_t voff, int soff, float v, int q, const Codec &c)
    {
        __builtin_amdgcn_raw_buffer_store_b16((short)f16c_encode((v - c.off[q]) * c.iw[q]), r, (int)voff, soff, 0);
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
    static __device__ __forceinline__ float
    bload(__amdgpu_buffer_rsrc_t r, uint32_t voff, int soff, int q, const Codec &c)
    {
        const unsigned short bits = (unsigned short)__builtin_amdgcn_raw_buffer_load_b16(r, (int)voff, soff, 0);
        return fmaf(__uint_as_float(static_cast<unsigned int>(bits) << 16), c.w[q], c.off[q]);
    }
    static __device__ __forceinline__ void bstore(__amdgpu_buffer_rsrc_t r, uint32_t voff, int soff, float v, int q, const Codec &c)
    {
        unsigned int u = __float_as_uint((v - c.off[q]) * c.iw[q]);
        u += 0x7FFFu + ((u >> 16) & 1u);
        __builtin_amdgcn_raw_buffer_store_b16((short)(unsigned short)(u >> 16), r, (int)voff, soff, 0);
    }
};

} // namespace lbm3
