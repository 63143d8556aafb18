// szx_wave.hpp -- segmented wave64 primitives for gfx950 (CDNA4).
//
// A "segment" is an aligned group of SEG lanes (SEG a power of two <= 64):
// the lanes holding one SZx block.  Reductions, scans and neighbour moves
// inside a segment use DPP (quad_perm, row_shr, row_mirror, row_bcast,
// wave_shr) and ds_swizzle, which run in the VALU / LDS crossbar without the
// address arithmetic of ds_bpermute.  Only the 32-lane crossing of a 64-lane
// segment falls back to a shuffle.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace szxamd {

// DPP control codes (GFX9 encoding).
constexpr int DPP_QUAD_XOR1 = 0xB1;     // quad_perm [1,0,3,2]
constexpr int DPP_QUAD_XOR2 = 0x4E;     // quad_perm [2,3,0,1]
constexpr int DPP_ROW_SHR(int d) { return 0x110 + d; }
constexpr int DPP_WAVE_SHR1 = 0x138;
constexpr int DPP_ROW_MIRROR = 0x140;
constexpr int DPP_ROW_HALF_MIRROR = 0x141;
constexpr int DPP_ROW_BCAST15 = 0x142;
constexpr int DPP_ROW_BCAST31 = 0x143;

template <int CTRL, int ROWMASK = 0xF, int BANKMASK = 0xF>
__device__ inline uint32_t dpp32(uint32_t old, uint32_t v)
{
    return (uint32_t)__builtin_amdgcn_update_dpp((int)old, (int)v, CTRL, ROWMASK, BANKMASK, false);
}

template <int PATTERN>
__device__ inline uint32_t swz32(uint32_t v)
{
    return (uint32_t)__builtin_amdgcn_ds_swizzle((int)v, PATTERN);
}

// ds_swizzle bit mode: lane' = ((lane & and) | or) ^ xor, within 32 lanes.
constexpr int SWZ(int and_m, int or_m, int xor_m) { return and_m | (or_m << 5) | (xor_m << 10); }

// 32- or 64-bit lanes as 32-bit pieces.
template <typename V> struct Pieces;
template <> struct Pieces<uint32_t> {
    static constexpr int N = 1;
    __device__ static uint32_t get(uint32_t v, int) { return v; }
    __device__ static uint32_t put(uint32_t, int, uint32_t p) { return p; }
};
template <> struct Pieces<uint64_t> {
    static constexpr int N = 2;
    __device__ static uint32_t get(uint64_t v, int i) { return (uint32_t)(v >> (32 * i)); }
    __device__ static uint64_t put(uint64_t v, int i, uint32_t p)
    {
        return i ? ((v & 0xFFFFFFFFull) | ((uint64_t)p << 32)) : ((v & ~0xFFFFFFFFull) | p);
    }
};

// DPP move with an undefined old value: every lane of the permutations used
// with it (quad_perm, row_mirror, row_half_mirror) reads a valid source, and
// the compiler can then fold the move into the ALU op that consumes it.
template <int CTRL>
__device__ inline uint32_t dppu32(uint32_t v)
{
    return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, CTRL, 0xF, 0xF, true);
}

template <int CTRL, typename V>
__device__ inline V dppu(V v)
{
    V r = v;
#pragma unroll
    for (int i = 0; i < Pieces<V>::N; i++) r = Pieces<V>::put(r, i, dppu32<CTRL>(Pieces<V>::get(v, i)));
    return r;
}

template <int CTRL, int ROWMASK = 0xF, typename V>
__device__ inline V dppv(V old, V v)
{
    V r = old;
#pragma unroll
    for (int i = 0; i < Pieces<V>::N; i++)
        r = Pieces<V>::put(r, i, dpp32<CTRL, ROWMASK>(Pieces<V>::get(old, i), Pieces<V>::get(v, i)));
    return r;
}

template <int PATTERN, typename V>
__device__ inline V swzv(V v)
{
    V r = v;
#pragma unroll
    for (int i = 0; i < Pieces<V>::N; i++)
        r = Pieces<V>::put(r, i, swz32<PATTERN>(Pieces<V>::get(v, i)));
    return r;
}

// Value of the partner lane for a butterfly step of distance D (1..32) in a
// segment of at least 2D lanes.  Any pairing that combines the two halves of
// every 2D-lane group works for a commutative reduction.
template <int D, typename V>
__device__ inline V partner(V v)
{
    if constexpr (D == 1) return dppu<DPP_QUAD_XOR1>(v);
    else if constexpr (D == 2) return dppu<DPP_QUAD_XOR2>(v);
    else if constexpr (D == 4) return dppu<DPP_ROW_HALF_MIRROR>(v);
    else if constexpr (D == 8) return dppu<DPP_ROW_MIRROR>(v);
    else if constexpr (D == 16) return swzv<SWZ(0x1F, 0, 0x10)>(v);
    else return __shfl_xor(v, 32, 64);
}

// All-lane segmented reduction with a commutative, associative op.
template <int SEG, typename V, typename Op>
__device__ inline V seg_reduce(V v, Op op)
{
    if constexpr (SEG >= 2) v = op(v, partner<1>(v));
    if constexpr (SEG >= 4) v = op(v, partner<2>(v));
    if constexpr (SEG >= 8) v = op(v, partner<4>(v));
    if constexpr (SEG >= 16) v = op(v, partner<8>(v));
    if constexpr (SEG >= 32) v = op(v, partner<16>(v));
    if constexpr (SEG >= 64) v = op(v, partner<32>(v));
    return v;
}

// Inclusive segmented prefix sum of a u32.
template <int SEG>
__device__ inline uint32_t seg_scan_incl(uint32_t v, int lane)
{
    // row_shr reads 0 where the source lane is outside the 16-lane row, so a
    // segment spanning whole rows needs no lane predicate
    const int ls = lane & (SEG - 1);
    constexpr bool ROWS = SEG >= 16;
    if constexpr (SEG >= 2) { uint32_t o = dpp32<DPP_ROW_SHR(1)>(0u, v); if (ROWS || ls >= 1) v += o; }
    if constexpr (SEG >= 4) { uint32_t o = dpp32<DPP_ROW_SHR(2)>(0u, v); if (ROWS || ls >= 2) v += o; }
    if constexpr (SEG >= 8) { uint32_t o = dpp32<DPP_ROW_SHR(4)>(0u, v); if (ROWS || ls >= 4) v += o; }
    if constexpr (SEG >= 16) { uint32_t o = dpp32<DPP_ROW_SHR(8)>(0u, v); v += o; }
    if constexpr (SEG >= 32) v += dpp32<DPP_ROW_BCAST15, 0xA>(0u, v);
    if constexpr (SEG >= 64) v += dpp32<DPP_ROW_BCAST31, 0xC>(0u, v);
    return v;
}

// Segmented inclusive "last writer" scan over byte lanes: d holds data
// bytes, m marks (0xFF) the bytes this lane wrote (d is zero elsewhere).  A
// lane ends with, per byte, the data of the nearest lane at or before it in
// its segment that wrote the byte, and m the union of the marks.  Within a
// 16-lane row a row_shr source outside the row reads (0, 0), which changes
// nothing, so whole-row segments need no lane predicate.
template <int SEG>
__device__ inline void seg_scan_last_writer(uint32_t &d, uint32_t &m, int lane)
{
    const int ls = lane & (SEG - 1);
    constexpr bool ROWS = SEG >= 16;
#define SZX_LW_STEP(CTRL, DIST, RM)                                     \
    {                                                                   \
        const uint32_t od = dpp32<CTRL, RM>(0u, d);                     \
        const uint32_t om = dpp32<CTRL, RM>(0u, m);                     \
        if (ROWS || ls >= (DIST)) {                                     \
            d = (d & m) | (od & ~m);                                    \
            m |= om;                                                    \
        }                                                               \
    }
    if constexpr (SEG >= 2) SZX_LW_STEP(DPP_ROW_SHR(1), 1, 0xF)
    if constexpr (SEG >= 4) SZX_LW_STEP(DPP_ROW_SHR(2), 2, 0xF)
    if constexpr (SEG >= 8) SZX_LW_STEP(DPP_ROW_SHR(4), 4, 0xF)
    if constexpr (SEG >= 16) SZX_LW_STEP(DPP_ROW_SHR(8), 8, 0xF)
    if constexpr (SEG >= 32) SZX_LW_STEP(DPP_ROW_BCAST15, 16, 0xA)
    if constexpr (SEG >= 64) SZX_LW_STEP(DPP_ROW_BCAST31, 32, 0xC)
#undef SZX_LW_STEP
}

// Value held by lane-1 (0 in the first lane of each segment).
template <int SEG, typename V>
__device__ inline V seg_prev(V v, int lane)
{
    V o = dppv<DPP_WAVE_SHR1>((V)0, v);
    return (lane & (SEG - 1)) ? o : (V)0;
}

// Broadcast of the first / last lane of each segment.
template <int SEG, typename V>
__device__ inline V seg_first(V v)
{
    if constexpr (SEG == 1) return v;
    else if constexpr (SEG <= 32) return swzv<SWZ(0x1F & ~(SEG - 1), 0, 0)>(v);
    else return __shfl(v, 0, 64);
}
template <int SEG, typename V>
__device__ inline V seg_last(V v)
{
    if constexpr (SEG == 1) return v;
    else if constexpr (SEG <= 32) return swzv<SWZ(0x1F & ~(SEG - 1), SEG - 1, 0)>(v);
    else return __shfl(v, 63, 64);
}

}  // namespace szxamd
