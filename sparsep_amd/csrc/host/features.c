/* Structural features of a sparse matrix for the learned variant selector.
 * Row and column lengths come from one pass over all entries; the
 * per-entry structure (bandwidth, diagonal distances, locality, block fill)
 * from up to 32 evenly spaced runs of at most 512 rows and about 32K
 * entries each (all rows when there are at most 16,384).  Input sorted by row is read in place, otherwise grouped by
 * a counting sort first.  Extraction costs a small fraction of building one
 * plan.  The block
 * fill ratios restate the occupancy the reference computes when it builds
 * BCOO (spmv/1D/BCOO-block/support/matrix.h:491-505), for 2x2 and 4x4
 * blocks, without building the blocks.  Feature layout (index: meaning):
 *   0-3   log sizes: rows, cols, nnz, -log density
 *   4-13  row lengths: log avg, cv, log max, log max/avg, log min, empty
 *         fraction, Gini, log p50 / p90 / p99
 *   14-29 row-length log2 histogram (fraction of rows per bin)
 *   30-34 column lengths: log avg, cv, log max, empty fraction, Gini
 *   35-45 bandwidth: mean and max |col - diag| / ncols, fraction within 32
 *         of the diagonal, 8-bin diagonal-distance histogram
 *   46-48 fill ratio of 2x2, 4x4 blocks, log blocks per 4-row block row
 *   49-52 gather locality: next-column runs, |dc| < 16, new 64-byte x lines
 *         per entry, mean |len[i] - len[i-1]| / avg
 *   53-56 skew: nnz share of top 1% and 10% rows, square flag, fraction of
 *         rows longer than 4x average
 *   57    log matrix bytes
 *   58-63 dtype one-hot (INT8 ... FP64) */
#include <math.h>

#include "internal.h"

static float lg2(double v) { return (float)log2(v + 1.0); }

/* Gini coefficient of the values 0..maxv given as a histogram cnt[v]. */
static float gini_hist(const uint64_t *cnt, uint64_t maxv, uint64_t count, uint64_t total) {
    if (!count || !total) return 0.f;
    /* G = (2 sum_i i*x_(i) ) / (n sum x) - (n+1)/n with x sorted ascending, i = 1..n */
    double acc = 0, rank = 0;
    for (uint64_t v = 0; v <= maxv; v++) {
        if (!cnt[v]) continue;
        const double c = (double)cnt[v];
        /* ranks rank+1 .. rank+c, sum = c*rank + c(c+1)/2 */
        acc += (double)v * (c * rank + c * (c + 1) / 2);
        rank += c;
    }
    const double n = (double)count;
    return (float)(2 * acc / (n * (double)total) - (n + 1) / n);
}

/* Fills p50/p90/p99 of a histogram. */
static void pct_hist(const uint64_t *cnt, uint64_t maxv, uint64_t count, double *p50, double *p90, double *p99) {
    const double q[3] = {0.5, 0.9, 0.99};
    double *o[3] = {p50, p90, p99};
    uint64_t run = 0, v = 0;
    for (int k = 0; k < 3; k++) {
        const uint64_t need = (uint64_t)ceil(q[k] * (double)count);
        while (v <= maxv && run + cnt[v] < need) run += cnt[v++];
        *o[k] = (double)(v <= maxv ? v : maxv);
    }
}

int sp_features(const sp_coo *A, float *f) {
    if (!A || !f || !sp_valid_dtype(A->dtype)) return SP_ERR_ARG;
    if (A->nnz && (!A->rowind || !A->colind)) return SP_ERR_ARG;
    memset(f, 0, SP_NFEAT * sizeof(float));
    const uint64_t n = A->nrows, m = A->ncols, nnz = A->nnz;
    const size_t sv = sp_dtype_size(A->dtype);
    f[0] = lg2((double)n) / 24.f;
    f[1] = lg2((double)m) / 24.f;
    f[2] = lg2((double)nnz) / 32.f;
    f[3] = (n && m && nnz) ? (float)(-log2((double)nnz / ((double)n * (double)m)) / 40.0) : 1.f;
    f[55] = (n == m) ? 1.f : 0.f;
    f[57] = lg2((double)nnz * (double)(sv + 4)) / 40.f;
    f[58 + (A->dtype - SP_I8)] = 1.f;
    if (!n || !m || !nnz) return SP_OK;

    uint32_t *rlen = calloc(n + 1, sizeof(uint32_t));
    uint32_t *clen = calloc(m, sizeof(uint32_t));
    uint64_t *rptr = malloc((n + 1) * sizeof(uint64_t));
    uint32_t *owned = NULL;
    const uint32_t *cols = A->colind;
    int rc = SP_ERR_NOMEM;
    uint64_t *rh = NULL, *ch = NULL;
    if (!rlen || !clen || !rptr) goto done;
    int sorted = 1;
    {
        /* counts with relaxed atomic increments over host threads (the
         * passes are memory bound; the result does not depend on order) */
        const uint32_t *ri = A->rowind, *ci = A->colind;
        int bad = 0;
#pragma omp parallel reduction(| : bad) reduction(& : sorted) if (nnz > (1u << 18))
        {
            /* row counts: runs of equal rows are counted locally and added once */
            uint32_t cur = UINT32_MAX, run = 0;
#pragma omp for schedule(static) nowait
            for (uint64_t k = 0; k < nnz; k++) {
                const uint32_t r = ri[k];
                bad |= (r >= n);
                sorted &= (k == 0 || ri[k - 1] <= r);
                if (r != cur) {
                    if (run) __atomic_fetch_add(&rlen[cur], run, __ATOMIC_RELAXED);
                    cur = r < n ? r : (uint32_t)n;
                    run = 0;
                }
                run++;
            }
            if (run) __atomic_fetch_add(&rlen[cur], run, __ATOMIC_RELAXED);
#pragma omp for schedule(static)
            for (uint64_t k = 0; k < nnz; k++) {
                const uint32_t c = ci[k];
                bad |= (c >= m);
                __atomic_fetch_add(&clen[c < m ? c : 0], 1u, __ATOMIC_RELAXED);
            }
        }
        if (bad) { rc = SP_ERR_ARG; goto done; }
    }
    rptr[0] = 0;
    for (uint64_t r = 0; r < n; r++) rptr[r + 1] = rptr[r] + rlen[r];
    if (!sorted) {
        /* stable counting sort by row (input order kept inside a row) */
        uint64_t *pos = malloc(n * sizeof(uint64_t));
        owned = malloc(nnz * sizeof(uint32_t));
        if (!pos || !owned) { free(pos); goto done; }
        memcpy(pos, rptr, n * sizeof(uint64_t));
        for (uint64_t k = 0; k < nnz; k++) owned[pos[A->rowind[k]]++] = A->colind[k];
        free(pos);
        cols = owned;
    }

    /* row-length statistics */
    uint64_t rmax = 0, rmin = UINT64_MAX, rempty = 0;
    double rs2 = 0;
    for (uint64_t r = 0; r < n; r++) {
        const uint64_t l = rlen[r];
        if (l > rmax) rmax = l;
        if (l < rmin) rmin = l;
        rempty += (l == 0);
        rs2 += (double)l * (double)l;
    }
    const double ravg = (double)nnz / (double)n;
    const double rstd = sqrt(fmax(rs2 / (double)n - ravg * ravg, 0.0));
    f[4] = lg2(ravg) / 8.f;
    f[5] = (float)(rstd / (ravg + 1.0));
    f[6] = lg2((double)rmax) / 16.f;
    f[7] = lg2((double)rmax / ravg) / 16.f;
    f[8] = lg2((double)rmin) / 8.f;
    f[9] = (float)((double)rempty / (double)n);
    rh = calloc(rmax + 1, sizeof(uint64_t));
    if (!rh) goto done;
    for (uint64_t r = 0; r < n; r++) rh[rlen[r]]++;
    f[10] = gini_hist(rh, rmax, n, nnz);
    {
        double p50, p90, p99;
        pct_hist(rh, rmax, n, &p50, &p90, &p99);
        f[11] = lg2(p50) / 8.f;
        f[12] = lg2(p90) / 8.f;
        f[13] = lg2(p99) / 8.f;
    }
    for (uint64_t r = 0; r < n; r++) {
        const uint32_t l = rlen[r];
        int bin = 0;
        if (l) {
            bin = 1;
            while (bin < 15 && (1u << bin) <= l) bin++;
        }
        f[14 + bin] += 1.f;
    }
    for (int b = 0; b < 16; b++) f[14 + b] /= (float)n;
    {
        /* nnz share of the longest 1% / 10% rows, read off the histogram */
        const uint64_t top1 = (n + 99) / 100, top10 = (n + 9) / 10;
        uint64_t taken = 0, s1 = 0, s10 = 0, big = 0;
        for (uint64_t v = rmax + 1; v-- > 0;) {
            const uint64_t c = rh[v];
            if (!c) continue;
            if (taken < top1) s1 += v * (c < top1 - taken ? c : top1 - taken);
            if (taken < top10) s10 += v * (c < top10 - taken ? c : top10 - taken);
            taken += c;
            if ((double)v > 4 * ravg) big += c;
        }
        f[53] = (float)((double)s1 / (double)nnz);
        f[54] = (float)((double)s10 / (double)nnz);
        f[56] = (float)((double)big / (double)n);
    }

    /* column-length statistics */
    uint64_t cmax = 0, cempty = 0;
    double cs2 = 0;
    for (uint64_t c = 0; c < m; c++) {
        const uint64_t l = clen[c];
        if (l > cmax) cmax = l;
        cempty += (l == 0);
        cs2 += (double)l * (double)l;
    }
    const double cavg = (double)nnz / (double)m;
    f[30] = lg2(cavg) / 8.f;
    f[31] = (float)(sqrt(fmax(cs2 / (double)m - cavg * cavg, 0.0)) / (cavg + 1.0));
    f[32] = lg2((double)cmax) / 16.f;
    f[33] = (float)((double)cempty / (double)m);
    ch = calloc(cmax + 1, sizeof(uint64_t));
    if (!ch) goto done;
    for (uint64_t c = 0; c < m; c++) ch[clen[c]]++;
    f[34] = gini_hist(ch, cmax, m, nnz);

    /* bandwidth, locality and block fill over the sampled row runs */
    {
        double dsum = 0, dmax = 0;
        uint64_t near = 0, runs = 0, close = 0, lines = 0, nb2 = 0, nb4 = 0, sent = 0, srows = 0;
        uint64_t dh[8] = {0};
        double jump = 0;
        int nomem = 0;
        for (uint64_t r = 1; r < n; r++) jump += fabs((double)rlen[r] - (double)rlen[r - 1]);
        const double slope = (double)m / (double)n;
        const uint64_t RUN = 512, NRUN = 32;
        const uint64_t nruns = n <= RUN * NRUN ? 1 : NRUN;
        /* bins: 0 | 1 | 2-3 | 4-15 | 16-255 | 256-4095 | 4096-65535 | more */
        static const uint8_t bin_of_log2[64] = {1, 2, 3, 3, 4, 4, 4, 4, 5, 5, 5, 5, 6, 6, 6, 6,
                                                7, 7, 7, 7, 7, 7, 7, 7, 7, 7, 7, 7, 7, 7, 7, 7,
                                                7, 7, 7, 7, 7, 7, 7, 7, 7, 7, 7, 7, 7, 7, 7, 7,
                                                7, 7, 7, 7, 7, 7, 7, 7, 7, 7, 7, 7, 7, 7, 7, 7};
#pragma omp parallel reduction(+ : dsum, near, runs, close, lines, nb2, nb4, sent, srows, dh[:8]) \
    reduction(max : dmax) reduction(| : nomem) if (nruns > 1)
        {
            uint64_t *set = NULL;
            uint64_t cap = 0;
#pragma omp for schedule(dynamic)
            for (uint64_t q = 0; q < nruns; q++) {
                const uint64_t r0 = nruns == 1 ? 0 : ((q * (n - RUN)) / (NRUN - 1)) & ~(uint64_t)3;
                uint64_t r1 = nruns == 1 ? n : r0 + RUN;
                /* a run stops after ~32K entries (whole 4-row groups) so hub rows
                 * do not dominate the cost */
                if (nruns > 1)
                    while (r1 > r0 + 4 && rptr[r1 - 4] - rptr[r0] >= (1u << 15)) r1 -= 4;
                const uint64_t ne = rptr[r1] - rptr[r0];
                /* distinct (block row, block column) keys of the run: open addressing,
                 * 2x2 keys tagged with bit 63 to share one table with the 4x4 keys */
                uint64_t want = 16;
                while (want < 4 * ne) want <<= 1;
                if (want > cap) {
                    free(set);
                    set = malloc(want * sizeof(uint64_t));
                    cap = set ? want : 0;
                    if (!set) { nomem = 1; continue; }
                }
                memset(set, 0xFF, want * sizeof(uint64_t));
                const uint64_t mask = want - 1;
                srows += r1 - r0;
                sent += ne;
                for (uint64_t r = r0; r < r1; r++) {
                    const int64_t dg = (int64_t)floor((double)r * slope);
                    int64_t prev = -1000;
                    for (uint64_t k = rptr[r]; k < rptr[r + 1]; k++) {
                        const int64_t c = cols[k];
                        const int64_t d = c - dg;
                        const uint64_t ud = (uint64_t)(d < 0 ? -d : d);
                        dsum += (double)ud;
                        if ((double)ud > dmax) dmax = (double)ud;
                        near += (ud <= 32);
                        dh[ud ? bin_of_log2[63 - __builtin_clzll(ud)] : 0]++;
                        runs += (c == prev + 1);
                        close += (c - prev < 16 && prev - c < 16);
                        lines += (prev < 0 || (c >> 4) != (prev >> 4));
                        prev = c;
                        const uint64_t key2 = (1ULL << 63) | ((uint64_t)(r / 2) << 32) | (uint64_t)(c / 2);
                        const uint64_t key4 = ((uint64_t)(r / 4) << 32) | (uint64_t)(c / 4);
                        for (int z = 0; z < 2; z++) {
                            const uint64_t key = z ? key4 : key2;
                            uint64_t h = (key * 0x9E3779B97F4A7C15ULL) >> 20;
                            for (;; h++) {
                                uint64_t *slot = &set[h & mask];
                                if (*slot == key) break;
                                if (*slot == UINT64_MAX) {
                                    *slot = key;
                                    if (z) nb4++;
                                    else nb2++;
                                    break;
                                }
                            }
                        }
                    }
                }
            }
            free(set);
        }
        if (nomem) goto done;
        const double se = sent ? (double)sent : 1.0;
        f[35] = (float)(dsum / se / (double)m);
        f[36] = (float)(dmax / (double)m);
        f[37] = (float)((double)near / se);
        for (int b = 0; b < 8; b++) f[38 + b] = (float)((double)dh[b] / se);
        f[46] = nb2 ? (float)((double)sent / (4.0 * (double)nb2)) : 0.f;
        f[47] = nb4 ? (float)((double)sent / (16.0 * (double)nb4)) : 0.f;
        f[48] = lg2((double)nb4 / (double)((srows + 3) / 4)) / 12.f;
        f[49] = (float)((double)runs / se);
        f[50] = (float)((double)close / se);
        f[51] = (float)((double)lines / se);
        f[52] = (float)(jump / (double)n / (ravg + 1.0));
    }
    rc = SP_OK;
done:
    free(rlen);
    free(clen);
    free(rptr);
    free(owned);
    free(rh);
    free(ch);
    return rc;
}
