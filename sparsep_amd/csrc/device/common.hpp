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

// v += lane_xor<M>(v) for every power-of-two M in [lo, hi)
template <int M, int HI, typename A> __device__ __forceinline__ A xor_sum(A v) {
    if constexpr (M < HI) {
        v += lane_xor<M>(v);
        return xor_sum<M * 2, HI>(v);
    } else {
        return v;
    }
}

constexpr int WAVE = 64;

// Uncached-at-L2-for-reuse hints are not used: every matrix stream is read
// exactly once per SpMV and x is reused through L2/MALL.
template <typename P> __device__ __forceinline__ auto ldg(P p) { return __builtin_nontemporal_load(p); }

// Global-address-space views of the pointers the kernels read from problem
// descriptors: loads and stores through them compile to global_* rather than
// flat_* instructions (flat memory operations also count against the LDS
// wait counter, which serialises them with the kernels' LDS traffic).
template <typename T> using gptr = __attribute__((address_space(1))) T *;
template <typename T> __device__ __forceinline__ gptr<const T> gp(const T *p) { return (gptr<const T>)p; }
template <typename T> __device__ __forceinline__ gptr<T> gp(T *p) { return (gptr<T>)p; }

typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));

// N contiguous values from a stream read once, in the widest loads the size
// allows (16- or 8-byte, non-temporal).  p must be aligned to the load width
// (N * sizeof(T) a multiple of it and p at a multiple of N * sizeof(T)).
template <typename T, int N> __device__ __forceinline__ void ldg_n(gptr<const T> p, T (&out)[N]) {
    constexpr int B = N * (int)sizeof(T);
    if constexpr (B % 16 == 0) {
        u32x4 w[B / 16];
#pragma unroll
        for (int i = 0; i < B / 16; i++) w[i] = __builtin_nontemporal_load((gptr<const u32x4>)p + i);
        __builtin_memcpy(out, w, B);
    } else if constexpr (B % 8 == 0) {
        u32x2 w[B / 8];
#pragma unroll
        for (int i = 0; i < B / 8; i++) w[i] = __builtin_nontemporal_load((gptr<const u32x2>)p + i);
        __builtin_memcpy(out, w, B);
    } else if constexpr (B == 4) {
        const uint32_t w = __builtin_nontemporal_load((gptr<const uint32_t>)p);
        __builtin_memcpy(out, &w, B);
    } else {
#pragma unroll
        for (int i = 0; i < N; i++) out[i] = ldg(p + i);
    }
}

// N contiguous values of a vector that is reused through the caches (x):
// one wide load when the address is aligned to the vector's size, else
// element loads.
template <typename T, int N> __device__ __forceinline__ void ld_n(gptr<const T> p, T (&out)[N]) {
    constexpr int B = N * (int)sizeof(T);
    if constexpr (B % 16 == 0) {
        if (((uintptr_t)p & 15) == 0) {
            u32x4 w[B / 16];
#pragma unroll
            for (int i = 0; i < B / 16; i++) w[i] = ((gptr<const u32x4>)p)[i];
            __builtin_memcpy(out, w, B);
            return;
        }
    } else if constexpr (B == 8) {
        if (((uintptr_t)p & 7) == 0) {
            const u32x2 w = *(gptr<const u32x2>)p;
            __builtin_memcpy(out, &w, B);
            return;
        }
    }
#pragma unroll
    for (int i = 0; i < N; i++) out[i] = p[i];
}

// y[0] += v as one device-scope atomic: the lock-based merges of the
// reference's COO/BCOO kernels (one mutex, or 32 hashed mutexes, around
// `y[row] += partial`, spmv/1D/COO-nnz/dpu/task.c:819-885).  4- and 8-byte
// values use a hardware add (integers modulo 2^bits, as the accumulation);
// 1- and 2-byte integers a compare-and-swap on the aligned 32-bit word that
// holds them, so bytes next to y[0] are written back unchanged.
template <typename T> __device__ __forceinline__ void atomic_add_y(gptr<T> p, acc_t<T> v) {
    if constexpr (__is_same(T, float) || __is_same(T, double)) {
        __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    } else if constexpr (sizeof(T) == 4) {
        __hip_atomic_fetch_add((gptr<uint32_t>)p, (uint32_t)v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    } else if constexpr (sizeof(T) == 8) {
        __hip_atomic_fetch_add((gptr<uint64_t>)p, (uint64_t)v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    } else {
        const uintptr_t a = (uintptr_t)p;
        const gptr<uint32_t> w = (gptr<uint32_t>)(a & ~(uintptr_t)3);
        const uint32_t sh = (uint32_t)(a & 3) * 8, m = (sizeof(T) == 1 ? 0xFFu : 0xFFFFu) << sh;
        uint32_t old = __hip_atomic_load(w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        for (;;) {
            const uint32_t nw = (old & ~m) | ((((old & m) >> sh) + (uint32_t)v) << sh & m);
            if (__hip_atomic_compare_exchange_strong(w, &old, nw, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                     __HIP_MEMORY_SCOPE_AGENT))
                break;
        }
    }
}

}  // namespace spd
