/* Seeded synthetic matrix generators (the workloads the configs name). */
#include <math.h>

#include "internal.h"

#define KEY(r, c) (((uint64_t)(r) << 32) | (uint32_t)(c))

int sp_gen_stencil5(uint32_t gx, uint32_t gy, int dtype, sp_dist offdiag, double diag,
                    uint64_t seed, sp_coo **out) {
    if (!out || !sp_valid_dtype(dtype) || gx == 0 || gy == 0) return SP_ERR_ARG;
    uint64_t n = (uint64_t)gx * gy;
    if (n > 0xFFFFFFFFULL) return SP_ERR_UNSUPPORTED;
    uint64_t nnz = 5 * n - 2 * (uint64_t)gx - 2 * (uint64_t)gy;
    sp_coo *A;
    int rc = sp_coo_alloc((uint32_t)n, (uint32_t)n, (uint32_t)nnz, dtype, &A);
    if (rc) return rc;
    sp_rng r;
    sp_rng_seed(&r, seed);
    uint64_t k = 0;
    for (uint32_t y = 0; y < gy; y++)
        for (uint32_t x = 0; x < gx; x++) {
            uint32_t i = y * gx + x;
            uint32_t cols[5];
            int nc = 0;
            if (y > 0) cols[nc++] = i - gx;
            if (x > 0) cols[nc++] = i - 1;
            cols[nc++] = i;
            if (x + 1 < gx) cols[nc++] = i + 1;
            if (y + 1 < gy) cols[nc++] = i + gx;
            for (int q = 0; q < nc; q++) {
                A->rowind[k] = i;
                A->colind[k] = cols[q];
                sp_put(A->values, dtype, k, cols[q] == i ? diag : sp_draw(&r, dtype, offdiag));
                k++;
            }
            A->rows[i] = (uint32_t)nc;
        }
    *out = A;
    return SP_OK;
}

/* Open-addressing set of u64 keys (key+1 stored so that 0 means empty). */
typedef struct { uint64_t *slot; uint64_t mask; } keyset;

static int keyset_init(keyset *s, uint64_t expected) {
    uint64_t cap = 1024;
    while (cap < 2 * expected) cap <<= 1;
    s->slot = (uint64_t *)calloc(cap, sizeof(uint64_t));
    s->mask = cap - 1;
    return s->slot ? 0 : -1;
}
/* returns 1 if inserted, 0 if present */
static int keyset_insert(keyset *s, uint64_t key) {
    uint64_t v = key + 1;
    uint64_t h = (key * 0x9E3779B97F4A7C15ULL) ^ (key >> 29);
    h = (h ^ (h >> 32)) & s->mask;
    for (;;) {
        uint64_t cur = s->slot[h];
        if (cur == 0) { s->slot[h] = v; return 1; }
        if (cur == v) return 0;
        h = (h + 1) & s->mask;
    }
}

int sp_gen_rmat(uint32_t scale, uint64_t nnz_target, double a, double b, double c, int dtype,
                sp_dist vals, uint64_t seed, sp_coo **out) {
    if (!out || !sp_valid_dtype(dtype) || scale == 0 || scale > 31) return SP_ERR_ARG;
    if (a < 0 || b < 0 || c < 0 || a + b + c > 1.0) return SP_ERR_ARG;
    uint64_t n = 1ULL << scale;
    if (nnz_target > n * n || nnz_target > 0xFFFFFFFFULL) return SP_ERR_ARG;
    keyset set;
    if (keyset_init(&set, nnz_target)) return SP_ERR_NOMEM;
    uint64_t *keys = (uint64_t *)malloc(nnz_target * sizeof(uint64_t) + 8);
    if (!keys) { free(set.slot); return SP_ERR_NOMEM; }
    /* Quadrant thresholds on 32-bit uniforms: two levels per 64-bit draw. */
    const double two32 = 4294967296.0;
    uint64_t ta = (uint64_t)(a * two32), tb = (uint64_t)((a + b) * two32), tc = (uint64_t)((a + b + c) * two32);
    sp_rng r;
    sp_rng_seed(&r, seed);
    uint64_t have = 0;
    while (have < nnz_target) {
        uint32_t row = 0, col = 0;
        uint64_t bits = 0;
        for (uint32_t lvl = 0; lvl < scale; lvl++) {
            if ((lvl & 1) == 0) bits = sp_rng_next(&r);
            uint64_t u = (lvl & 1) ? (bits >> 32) : (bits & 0xFFFFFFFFu);
            uint32_t rb, cb;
            if (u < ta) { rb = 0; cb = 0; }
            else if (u < tb) { rb = 0; cb = 1; }
            else if (u < tc) { rb = 1; cb = 0; }
            else { rb = 1; cb = 1; }
            row = (row << 1) | rb;
            col = (col << 1) | cb;
        }
        uint64_t key = KEY(row, col);
        if (keyset_insert(&set, key)) keys[have++] = key;
    }
    free(set.slot);
    sp_radix_sort_u64(keys, have);
    int rc = sp_coo_from_keys((uint32_t)n, (uint32_t)n, keys, have, dtype, vals, sp_value_seed(seed), out);
    free(keys);
    return rc;
}

static int cmp_u32(const void *x, const void *y) {
    uint32_t a = *(const uint32_t *)x, b = *(const uint32_t *)y;
    return a < b ? -1 : a > b;
}

int sp_gen_stencil27(uint32_t g, uint32_t dof, int periodic, int dtype, uint64_t seed,
                     sp_coo **out) {
    if (!out || !sp_valid_dtype(dtype) || g == 0 || dof == 0) return SP_ERR_ARG;
    if (periodic && g < 3) return SP_ERR_ARG;
    uint64_t nodes = (uint64_t)g * g * g, n = nodes * dof;
    if (n > 0xFFFFFFFFULL) return SP_ERR_UNSUPPORTED;
    /* count */
    uint64_t nnz = 0;
    for (uint64_t v = 0; v < nodes; v++) {
        uint32_t x = (uint32_t)(v % g), y = (uint32_t)((v / g) % g), z = (uint32_t)(v / ((uint64_t)g * g));
        uint32_t cnt = 0;
        for (int dz = -1; dz <= 1; dz++)
            for (int dy = -1; dy <= 1; dy++)
                for (int dx = -1; dx <= 1; dx++) {
                    if (!periodic && ((int64_t)x + dx < 0 || (int64_t)x + dx >= g || (int64_t)y + dy < 0 ||
                                      (int64_t)y + dy >= g || (int64_t)z + dz < 0 || (int64_t)z + dz >= g))
                        continue;
                    cnt++;
                }
        nnz += (uint64_t)cnt * dof * dof;
    }
    if (nnz > 0xFFFFFFFFULL) return SP_ERR_UNSUPPORTED;
    sp_coo *A;
    int rc = sp_coo_alloc((uint32_t)n, (uint32_t)n, (uint32_t)nnz, dtype, &A);
    if (rc) return rc;
    sp_rng r;
    sp_rng_seed(&r, seed);
    double *rowv = (double *)malloc(27 * (size_t)dof * sizeof(double));
    if (!rowv) { sp_free_coo(A); return SP_ERR_NOMEM; }
    uint64_t k = 0;
    for (uint64_t v = 0; v < nodes; v++) {
        uint32_t x = (uint32_t)(v % g), y = (uint32_t)((v / g) % g), z = (uint32_t)(v / ((uint64_t)g * g));
        uint32_t nb[27];
        int nn = 0;
        for (int dz = -1; dz <= 1; dz++)
            for (int dy = -1; dy <= 1; dy++)
                for (int dx = -1; dx <= 1; dx++) {
                    int64_t xx = (int64_t)x + dx, yy = (int64_t)y + dy, zz = (int64_t)z + dz;
                    if (periodic) {
                        xx = (xx + g) % g; yy = (yy + g) % g; zz = (zz + g) % g;
                    } else if (xx < 0 || xx >= g || yy < 0 || yy >= g || zz < 0 || zz >= g) {
                        continue;
                    }
                    nb[nn++] = (uint32_t)((uint64_t)zz * g * g + (uint64_t)yy * g + (uint64_t)xx);
                }
        qsort(nb, (size_t)nn, sizeof(uint32_t), cmp_u32);
        for (uint32_t d = 0; d < dof; d++) {
            uint32_t row = (uint32_t)(v * dof + d);
            uint32_t len = (uint32_t)nn * dof, diag_at = 0;
            double offsum = 0.0;
            for (int q = 0; q < nn; q++)
                for (uint32_t e = 0; e < dof; e++) {
                    uint32_t idx = (uint32_t)q * dof + e;
                    uint32_t col = nb[q] * dof + e;
                    double val = sp_rng_normal(&r);
                    if (sp_is_int(dtype)) val = nearbyint(val);
                    rowv[idx] = val;
                    if (col == row) diag_at = idx; else offsum += fabs(val);
                }
            rowv[diag_at] += offsum + 1.0;
            for (uint32_t idx = 0; idx < len; idx++) {
                A->rowind[k] = row;
                A->colind[k] = nb[idx / dof] * dof + idx % dof;
                sp_put(A->values, dtype, k, rowv[idx]);
                k++;
            }
            A->rows[row] = len;
        }
    }
    free(rowv);
    *out = A;
    return SP_OK;
}

/* Draw `k` distinct columns from [lo, lo+span) into cols (sorted). */
static void draw_distinct(sp_rng *r, uint32_t lo, uint32_t span, uint32_t k, uint32_t *cols,
                          uint32_t *scratch) {
    if (k >= span) { for (uint32_t q = 0; q < span; q++) cols[q] = lo + q; return; }
    if ((uint64_t)span <= 4ULL * k) { /* partial Fisher-Yates over the window */
        for (uint32_t q = 0; q < span; q++) scratch[q] = lo + q;
        for (uint32_t q = 0; q < k; q++) {
            uint32_t j = q + (uint32_t)(((__uint128_t)sp_rng_next(r) * (span - q)) >> 64);
            uint32_t t = scratch[q]; scratch[q] = scratch[j]; scratch[j] = t;
            cols[q] = scratch[q];
        }
    } else { /* rejection with a linear duplicate check (k is small) */
        uint32_t have = 0;
        while (have < k) {
            uint32_t c = lo + (uint32_t)(((__uint128_t)sp_rng_next(r) * span) >> 64);
            int dup = 0;
            for (uint32_t q = 0; q < have; q++) if (cols[q] == c) { dup = 1; break; }
            if (!dup) cols[have++] = c;
        }
    }
    qsort(cols, k, sizeof(uint32_t), cmp_u32);
}

/* Shared row-by-row generator: window(i) gives [lo, lo+span) for row i. */
typedef void (*window_fn)(uint32_t i, const void *ctx, uint32_t *lo, uint32_t *span);

static int gen_rows(uint32_t nrows, uint32_t ncols, uint32_t klo, uint32_t khi, window_fn win,
                    const void *wctx, uint32_t maxspan, int dtype, sp_dist vals, uint64_t seed,
                    sp_coo **out) {
    if (!out || !sp_valid_dtype(dtype) || klo > khi || nrows == 0 || ncols == 0) return SP_ERR_ARG;
    sp_rng r;
    sp_rng_seed(&r, seed);
    uint32_t *cnt = (uint32_t *)malloc((size_t)nrows * sizeof(uint32_t) + 4);
    if (!cnt) return SP_ERR_NOMEM;
    uint64_t nnz = 0;
    for (uint32_t i = 0; i < nrows; i++) {
        uint32_t lo, span;
        win(i, wctx, &lo, &span);
        uint32_t k = klo + (uint32_t)(((__uint128_t)sp_rng_next(&r) * (khi - klo + 1)) >> 64);
        if (k > span) k = span;
        cnt[i] = k;
        nnz += k;
    }
    if (nnz > 0xFFFFFFFFULL) { free(cnt); return SP_ERR_UNSUPPORTED; }
    sp_coo *A;
    int rc = sp_coo_alloc(nrows, ncols, (uint32_t)nnz, dtype, &A);
    if (rc) { free(cnt); return rc; }
    uint32_t *scratch = (uint32_t *)malloc(((size_t)maxspan + 1) * sizeof(uint32_t));
    uint32_t *cols = (uint32_t *)malloc(((size_t)khi + 1) * sizeof(uint32_t));
    if (!scratch || !cols) { free(cnt); free(scratch); free(cols); sp_free_coo(A); return SP_ERR_NOMEM; }
    sp_rng vr;
    sp_rng_seed(&vr, sp_value_seed(seed));
    uint64_t k = 0;
    for (uint32_t i = 0; i < nrows; i++) {
        uint32_t lo, span;
        win(i, wctx, &lo, &span);
        draw_distinct(&r, lo, span, cnt[i], cols, scratch);
        for (uint32_t q = 0; q < cnt[i]; q++) {
            A->rowind[k] = i;
            A->colind[k] = cols[q];
            sp_put(A->values, dtype, k, sp_draw(&vr, dtype, vals));
            k++;
        }
        A->rows[i] = cnt[i];
    }
    free(cnt); free(scratch); free(cols);
    *out = A;
    return SP_OK;
}

static void win_uniform(uint32_t i, const void *ctx, uint32_t *lo, uint32_t *span) {
    (void)i;
    *lo = 0;
    *span = *(const uint32_t *)ctx;
}

int sp_gen_uniform(uint32_t nrows, uint32_t ncols, uint32_t klo, uint32_t khi, int dtype,
                   sp_dist vals, uint64_t seed, sp_coo **out) {
    uint32_t span = ncols;
    uint32_t maxspan = 4 * khi + 4 < ncols ? 4 * khi + 4 : ncols;
    return gen_rows(nrows, ncols, klo, khi, win_uniform, &span, maxspan, dtype, vals, seed, out);
}

typedef struct { uint32_t n, bw; } band_ctx;
static void win_band(uint32_t i, const void *ctx, uint32_t *lo, uint32_t *span) {
    const band_ctx *b = (const band_ctx *)ctx;
    uint32_t l = i > b->bw ? i - b->bw : 0;
    uint64_t h = (uint64_t)i + b->bw + 1;
    if (h > b->n) h = b->n;
    *lo = l;
    *span = (uint32_t)(h - l);
}

int sp_gen_banded(uint32_t nrows, uint32_t bw, uint32_t klo, uint32_t khi, int dtype,
                  sp_dist vals, uint64_t seed, sp_coo **out) {
    band_ctx b = {nrows, bw};
    return gen_rows(nrows, nrows, klo, khi, win_band, &b, 2 * bw + 1, dtype, vals, seed, out);
}

typedef struct { uint32_t n, nb; } bd_ctx;
static void win_bd(uint32_t i, const void *ctx, uint32_t *lo, uint32_t *span) {
    const bd_ctx *b = (const bd_ctx *)ctx;
    uint32_t bsz = (b->n + b->nb - 1) / b->nb;
    uint32_t blk = i / bsz;
    uint32_t l = blk * bsz;
    uint32_t h = l + bsz < b->n ? l + bsz : b->n;
    *lo = l;
    *span = h - l;
}

int sp_gen_blockdiag(uint32_t nrows, uint32_t nblocks, uint32_t klo, uint32_t khi, int dtype,
                     sp_dist vals, uint64_t seed, sp_coo **out) {
    if (nblocks == 0 || nblocks > nrows) return SP_ERR_ARG;
    bd_ctx b = {nrows, nblocks};
    uint32_t bsz = (nrows + nblocks - 1) / nblocks;
    uint32_t maxspan = 4 * khi + 4 < bsz ? 4 * khi + 4 : bsz;
    return gen_rows(nrows, nrows, klo, khi, win_bd, &b, maxspan, dtype, vals, seed, out);
}
