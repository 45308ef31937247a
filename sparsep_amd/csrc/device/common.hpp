// Device-side helpers shared by the gfx950 SpMV kernels: value/accumulator
// traits, wave-level reductions and launch constants.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "sparsep.h"

namespace spd {

// Accumulator per value type.  Integer types accumulate in an unsigned type
// of at least 32 bits: C `val_dt acc; acc += a*b` (support/utils.h:13-34)
// is arithmetic modulo 2^bits, so accumulating modulo 2^32 (or 2^64) and
// truncating once at the end gives the identical result for any order.
template <typename T> struct Acc;
template <> struct Acc<int8_t> { using type = uint32_t; };
template <> struct Acc<int16_t> { using type = uint32_t; };
template <> struct Acc<int32_t> { using type = uint32_t; };
template <> struct Acc<int64_t> { using type = uint64_t; };
template <> struct Acc<float> { using type = float; };
template <> struct Acc<double> { using type = double; };

template <typename T> using acc_t = typename Acc<T>::type;

// a*b in the accumulator domain.  Signed operands are sign-extended first,
// then multiplied modulo 2^32/2^64 (the two's-complement product).
template <typename T> __device__ __forceinline__ acc_t<T> mul(T a, T b) {
    if constexpr (sizeof(T) <= 4 && !__is_same(T, float)) {
        return (uint32_t)((int32_t)a) * (uint32_t)((int32_t)b);
    } else if constexpr (__is_same(T, int64_t)) {
        return (uint64_t)a * (uint64_t)b;
    } else {
        return a * b;
    }
}

template <typename T> __device__ __forceinline__ acc_t<T> fma_acc(T a, T b, acc_t<T> c) {
    if constexpr (__is_same(T, float)) {
        return __builtin_fmaf(a, b, c);
    } else if constexpr (__is_same(T, double)) {
        return __builtin_fma(a, b, c);
    } else {
        return c + mul<T>(a, b);
    }
}

template <typename T> __device__ __forceinline__ T to_val(acc_t<T> a) { return (T)a; }

// ds_swizzle in bitmask mode: lane l reads lane ((l & and) | or) ^ xor inside
// each group of 32 lanes.  The offset must be a compile-time constant.
template <int XOR> __device__ __forceinline__ uint32_t swz_xor(uint32_t v) {
    constexpr int pattern = (XOR << 10) | (0 << 5) | 0x1F;
    return (uint32_t)__builtin_amdgcn_ds_swizzle((int)v, pattern);
}

template <int XOR> __device__ __forceinline__ uint32_t lane_xor_u32(uint32_t v) {
    if constexpr (XOR < 32) {
        return swz_xor<XOR>(v);
    } else {
        return (uint32_t)__shfl_xor((int)v, XOR, 64);
    }
}

template <int XOR, typename A> __device__ __forceinline__ A lane_xor(A v) {
    if constexpr (sizeof(A) == 4) {
        uint32_t u = __builtin_bit_cast(uint32_t, v);
        return __builtin_bit_cast(A, lane_xor_u32<XOR>(u));
    } else {
        uint64_t u = __builtin_bit_cast(uint64_t, v);
        uint32_t lo = lane_xor_u32<XOR>((uint32_t)u);
        uint32_t hi = lane_xor_u32<XOR>((uint32_t)(u >> 32));
        return __builtin_bit_cast(A, ((uint64_t)hi << 32) | lo);
    }
}

// Butterfly sum over groups of L lanes (L a power of two <= 64); every lane
// of the group ends with the group sum.
template <int L, typename A> __device__ __forceinline__ A group_sum(A v) {
    if constexpr (L >= 64) v += lane_xor<32>(v);
    if constexpr (L >= 32) v += lane_xor<16>(v);
    if constexpr (L >= 16) v += lane_xor<8>(v);
    if constexpr (L >= 8) v += lane_xor<4>(v);
    if constexpr (L >= 4) v += lane_xor<2>(v);
    if constexpr (L >= 2) v += lane_xor<1>(v);
    return v;
}

constexpr int WAVE = 64;

// Uncached-at-L2-for-reuse hints are not used: every matrix stream is read
// exactly once per SpMV and x is reused through L2/MALL.
template <typename T> __device__ __forceinline__ T ldg(const T *p) { return __builtin_nontemporal_load(p); }

}  // namespace spd
