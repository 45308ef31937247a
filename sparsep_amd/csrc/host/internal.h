/* Internal helpers shared by the host-surface C files. */
#ifndef SP_HOST_INTERNAL_H
#define SP_HOST_INTERNAL_H

#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "sparsep.h"

#ifdef __cplusplus
extern "C" {
#endif

static inline int sp_valid_dtype(int dt) { return dt >= SP_I8 && dt <= SP_F64; }
static inline int sp_is_int(int dt) { return dt >= SP_I8 && dt <= SP_I64; }

/* Store the integral / real value `v` at vals[i] in dtype dt.  Integers are
 * converted through int64 and truncated modulo 2^bits (two's complement). */
static inline void sp_put(void *vals, int dt, uint64_t i, double v) {
    switch (dt) {
    case SP_I8: ((int8_t *)vals)[i] = (int8_t)(uint8_t)(uint64_t)(int64_t)v; break;
    case SP_I16: ((int16_t *)vals)[i] = (int16_t)(uint16_t)(uint64_t)(int64_t)v; break;
    case SP_I32: ((int32_t *)vals)[i] = (int32_t)(uint32_t)(uint64_t)(int64_t)v; break;
    case SP_I64: ((int64_t *)vals)[i] = (int64_t)v; break;
    case SP_F32: ((float *)vals)[i] = (float)v; break;
    default: ((double *)vals)[i] = v; break;
    }
}

static inline double sp_get(const void *vals, int dt, uint64_t i) {
    switch (dt) {
    case SP_I8: return ((const int8_t *)vals)[i];
    case SP_I16: return ((const int16_t *)vals)[i];
    case SP_I32: return ((const int32_t *)vals)[i];
    case SP_I64: return (double)((const int64_t *)vals)[i];
    case SP_F32: return ((const float *)vals)[i];
    default: return ((const double *)vals)[i];
    }
}

/* 1 if vals[i] is non-zero (exact test in the value's own type). */
static inline int sp_nonzero(const void *vals, int dt, uint64_t i) {
    switch (dt) {
    case SP_I8: return ((const int8_t *)vals)[i] != 0;
    case SP_I16: return ((const int16_t *)vals)[i] != 0;
    case SP_I32: return ((const int32_t *)vals)[i] != 0;
    case SP_I64: return ((const int64_t *)vals)[i] != 0;
    case SP_F32: return ((const float *)vals)[i] != 0.0f;
    default: return ((const double *)vals)[i] != 0.0;
    }
}

/* Draw one value of the distribution for dtype dt. */
double sp_draw(sp_rng *r, int dt, sp_dist d);

/* Sort a COO's entries in place by (row, col), stable; fills rows[]. */
int sp_sort_coo_entries(sp_coo *A);

/* Build a COO from sorted unique (row, col) keys packed as row<<32|col,
 * drawing values from `vals` in key order with rng seeded by vseed. */
int sp_coo_from_keys(uint32_t nrows, uint32_t ncols, const uint64_t *keys, uint64_t nnz,
                     int dtype, sp_dist vals, uint64_t vseed, sp_coo **out);

/* Radix sort of u64 keys (ascending). */
void sp_radix_sort_u64(uint64_t *keys, uint64_t n);

/* Seed used for value streams derived from a structure seed. */
static inline uint64_t sp_value_seed(uint64_t seed) { return seed ^ 0x5DEECE66DA3B9F21ULL; }

#ifdef __cplusplus
}
#endif
#endif
