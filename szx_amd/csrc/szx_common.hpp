// szx_common.hpp -- device/host helpers shared by the gfx950 codec kernels.
//
// The codec is exact integer and bit work around a handful of IEEE float
// operations; everything here is written so that the GPU performs the same
// operations, in the same precision and rounding, as the reference CPU codec
// (no FMA contraction: the library is compiled with -ffp-contract=off; float
// denormals are preserved: -fno-gpu-flush-denormals-to-zero).
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace szxamd {

// Word traits: the codec works on the raw bits of each value.
template <typename T> struct Word;
template <> struct Word<float> {
    using U = uint32_t;
    static constexpr int W = 4;
    static constexpr int LEN_LO = 9;     // szx_float.cc:1711-1712
    static constexpr int LEN_BASE = 10;  // 9 + radExpo - reqExpo + 1, szx_float.cc:1710
};
template <> struct Word<double> {
    using U = uint64_t;
    static constexpr int W = 8;
    static constexpr int LEN_LO = 12;    // szx_double.cc:2024-2025
    static constexpr int LEN_BASE = 12;  // szx_double.cc:2023
};

__host__ __device__ inline uint32_t f32_bits(float v) { return __builtin_bit_cast(uint32_t, v); }
__host__ __device__ inline float bits_f32(uint32_t u) { return __builtin_bit_cast(float, u); }
__host__ __device__ inline uint64_t f64_bits(double v) { return __builtin_bit_cast(uint64_t, v); }
__host__ __device__ inline double bits_f64(uint64_t u) { return __builtin_bit_cast(double, u); }

template <typename T> __host__ __device__ inline typename Word<T>::U to_bits(T v);
template <> __host__ __device__ inline uint32_t to_bits<float>(float v) { return f32_bits(v); }
template <> __host__ __device__ inline uint64_t to_bits<double>(double v) { return f64_bits(v); }
template <typename T> __host__ __device__ inline T from_bits(typename Word<T>::U u);
template <> __host__ __device__ inline float from_bits<float>(uint32_t u) { return bits_f32(u); }
template <> __host__ __device__ inline double from_bits<double>(uint64_t u) { return bits_f64(u); }

__host__ __device__ inline uint32_t bswap32(uint32_t v)
{
    return (v >> 24) | ((v >> 8) & 0xFF00u) | ((v << 8) & 0xFF0000u) | (v << 24);
}

// Unbiased exponent field of the float radius (getExponent_float,
// szx_BytesToolkit.cc:335-346): a plain bit extract, so subnormals give -127.
__host__ __device__ inline int rad_expo(float r) { return (int)((f32_bits(r) >> 23) & 0xFF) - 127; }

// Unbiased exponent field of the error bound widened to double
// (getPrecisionReqLength_double, szx_BytesToolkit.cc:375-387), computed from
// the float bits with integer operations only so that no float->double
// conversion (and no denormal mode) is involved.  A float subnormal m*2^-149
// is a normal double with exponent floor(log2 m) - 149.
__host__ __device__ inline int req_expo(float eb)
{
    uint32_t u = f32_bits(eb) & 0x7FFFFFFFu;
    uint32_t e = u >> 23;
    if (e == 0xFF) return 1024;
    if (e != 0) return (int)e - 127;
    if (u == 0) return -1023;
    int lz = __builtin_clz(u);
    return (31 - lz) - 149;
}

// reqLength of a non-constant block (szx_float.cc:1708-1717,
// szx_double.cc:2020-2031).  clamped = 1 when the length hit 8*W; the stored
// median is then 0.
template <typename T>
__host__ __device__ inline int req_length(int radexp, int reqexp, int &clamped)
{
    int len = Word<T>::LEN_BASE + radexp - reqexp;
    clamped = 0;
    if (len < Word<T>::LEN_LO) len = Word<T>::LEN_LO;
    if (len > 8 * Word<T>::W) { len = 8 * Word<T>::W; clamped = 1; }
    return len;
}

// Stream layout constants (szx_float.cc:1091-1104).
constexpr int HEADER_BYTES = 13;

__host__ __device__ inline size_t ceil_div(size_t a, size_t b) { return (a + b - 1) / b; }

// ---------------------------------------------------------------------------
// Decoupled look-back state for the single-pass tile scan.  `status` packs a
// 2-bit flag with a 62-bit byte count; the matching block-count (aggregate or
// inclusive) sits in its own slot and is published before the flag with
// release ordering, so a reader that acquires the flag also sees the count.
// ---------------------------------------------------------------------------
struct TileState {
    unsigned long long status;
    unsigned long long cnt_agg;
    unsigned long long cnt_inc;
    unsigned long long pad;
};

constexpr unsigned long long FLAG_AGG = 1ull << 62;
constexpr unsigned long long FLAG_INC = 2ull << 62;
constexpr unsigned long long VAL_MASK = (1ull << 62) - 1;

__device__ inline unsigned long long load_acquire(const unsigned long long *p)
{
    return __hip_atomic_load(p, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline void store_release(unsigned long long *p, unsigned long long v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline void store_relaxed(unsigned long long *p, unsigned long long v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

}  // namespace szxamd
