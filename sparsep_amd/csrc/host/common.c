/* Status strings, dtype sizes, the project PRNG and value drawing. */
#include <math.h>

#include "internal.h"

const char *sp_strerror(int rc) {
    switch (rc) {
    case SP_OK: return "ok";
    case SP_ERR_ARG: return "invalid argument";
    case SP_ERR_NOMEM: return "out of memory (host, device or plan memory cap)";
    case SP_ERR_IO: return "i/o or parse error";
    case SP_ERR_HIP: return "HIP runtime error";
    case SP_ERR_UNSUPPORTED: return "unsupported variant, dtype or shape";
    case SP_ERR_NODEVICE: return "no HIP device";
    case SP_ERR_STATE: return "object in wrong state";
    case SP_ERR_RANGE: return "input larger than the entry point's fixed capacity";
    default: return "unknown error";
    }
}

/* Reference: byte_dt in support/utils.h:13-34. */
size_t sp_dtype_size(int dt) {
    switch (dt) {
    case SP_I8: return 1;
    case SP_I16: return 2;
    case SP_I32: return 4;
    case SP_I64: return 8;
    case SP_F32: return 4;
    case SP_F64: return 8;
    default: return 0;
    }
}

/* SplitMix64 (Steele, Lea, Flood 2014) expands the seed into the
 * xoshiro256** state (Blackman, Vigna 2018). */
static uint64_t splitmix64(uint64_t *x) {
    uint64_t z = (*x += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

void sp_rng_seed(sp_rng *r, uint64_t seed) {
    uint64_t x = seed;
    for (int i = 0; i < 4; i++) r->s[i] = splitmix64(&x);
}

static inline uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

uint64_t sp_rng_next(sp_rng *r) {
    uint64_t *s = r->s;
    uint64_t result = rotl(s[1] * 5, 7) * 9;
    uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 45);
    return result;
}

double sp_rng_uniform(sp_rng *r) { return (double)(sp_rng_next(r) >> 11) * 0x1.0p-53; }

double sp_rng_normal(sp_rng *r) {
    /* Box-Muller, one output per call (the partner deviate is discarded so
     * the stream stays one-draw-per-value). */
    double u1 = sp_rng_uniform(r);
    double u2 = sp_rng_uniform(r);
    if (u1 < 0x1.0p-60) u1 = 0x1.0p-60;
    return sqrt(-2.0 * log(u1)) * cos(6.283185307179586476925286766559 * u2);
}

/* Uniform integer in [lo, hi] (inclusive) by rejection-free multiply-shift
 * on 64 bits; the bias is below 2^-50 for the spans used here. */
static int64_t draw_int(sp_rng *r, int64_t lo, int64_t hi) {
    uint64_t span = (uint64_t)(hi - lo) + 1;
    __uint128_t m = (__uint128_t)sp_rng_next(r) * span;
    return lo + (int64_t)(uint64_t)(m >> 64);
}

double sp_draw(sp_rng *r, int dt, sp_dist d) {
    switch (d.kind) {
    case SP_DIST_CONST: return d.lo;
    case SP_DIST_NORMAL: {
        double v = d.lo + d.hi * sp_rng_normal(r);
        return sp_is_int(dt) ? nearbyint(v) : v;
    }
    default:
        if (sp_is_int(dt)) return (double)draw_int(r, (int64_t)d.lo, (int64_t)d.hi);
        return d.lo + (d.hi - d.lo) * sp_rng_uniform(r);
    }
}

int sp_gen_vector(int dtype, uint64_t n, sp_dist d, uint64_t seed, void *out) {
    if (!sp_valid_dtype(dtype) || (!out && n)) return SP_ERR_ARG;
    sp_rng r;
    sp_rng_seed(&r, seed);
    for (uint64_t i = 0; i < n; i++) sp_put(out, dtype, i, sp_draw(&r, dtype, d));
    return SP_OK;
}

/* init_vector: x[i] = (val_dt)(i % 4 + 1), spmv/1D/CSR-row/host/app.c:57-61
 * (the 2D copy at spmv/2D/DCSR/host/app.c:75 is the same). */
int sp_init_vector(int dtype, uint64_t n, void *out) {
    if (!sp_valid_dtype(dtype) || (!out && n)) return SP_ERR_ARG;
    for (uint64_t i = 0; i < n; i++) sp_put(out, dtype, i, (double)(i % 4 + 1));
    return SP_OK;
}

void sp_radix_sort_u64(uint64_t *keys, uint64_t n) {
    if (n < 2) return;
    if (n <= 256) { /* a 64K-bucket pass costs far more than insertion sort here */
        for (uint64_t i = 1; i < n; i++) {
            uint64_t v = keys[i], j = i;
            while (j > 0 && keys[j - 1] > v) { keys[j] = keys[j - 1]; j--; }
            keys[j] = v;
        }
        return;
    }
    uint64_t *tmp = (uint64_t *)malloc(n * sizeof(uint64_t));
    if (!tmp) { /* fall back to an in-place shell sort */
        for (uint64_t gap = n / 2; gap > 0; gap /= 2)
            for (uint64_t i = gap; i < n; i++) {
                uint64_t v = keys[i], j = i;
                while (j >= gap && keys[j - gap] > v) { keys[j] = keys[j - gap]; j -= gap; }
                keys[j] = v;
            }
        return;
    }
    uint64_t or_all = 0, and_all = ~0ULL;
    for (uint64_t i = 0; i < n; i++) { or_all |= keys[i]; and_all &= keys[i]; }
    uint64_t *src = keys, *dst = tmp;
    static const int BITS = 16;
    uint64_t *cnt = (uint64_t *)malloc(sizeof(uint64_t) << BITS);
    for (int shift = 0; shift < 64; shift += BITS) {
        uint64_t mask = ((1ULL << BITS) - 1) << shift;
        if (((or_all ^ and_all) & mask) == 0) continue; /* digit constant */
        memset(cnt, 0, sizeof(uint64_t) << BITS);
        for (uint64_t i = 0; i < n; i++) cnt[(src[i] >> shift) & 0xFFFF]++;
        uint64_t sum = 0;
        for (uint64_t b = 0; b < (1ULL << BITS); b++) { uint64_t c = cnt[b]; cnt[b] = sum; sum += c; }
        for (uint64_t i = 0; i < n; i++) dst[cnt[(src[i] >> shift) & 0xFFFF]++] = src[i];
        uint64_t *t = src; src = dst; dst = t;
    }
    if (src != keys) memcpy(keys, src, n * sizeof(uint64_t));
    free(cnt);
    free(tmp);
}
