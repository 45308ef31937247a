/*
 * sp_oracle.c -- TEST INFRASTRUCTURE ONLY.  The CPU oracle that the tests,
 * the smoke check and the bench's cpu_baseline leg use as the checker.  No
 * product code path links, loads or calls anything in this directory.
 *
 * It restates SparseP's CPU reference `spmv_host` for each format in the
 * reference's own accumulation type and order (val_dt arithmetic; built
 * with -fwrapv so integer overflow wraps as gcc's int conversion does, and
 * -ffp-contract=off so float sums round after every multiply and add, as
 * the reference's `gcc -O3 -std=c11` host build does on x86-64):
 *
 *   orc_spmv_csr   spmv/1D/CSR-row/host/app.c:66-77       (CSR-nnz:66 identical)
 *   orc_spmv_coo   spmv/1D/COO-row/host/app.c:70-74       (COO-nnz-rgrn:71, COO-nnz:69)
 *   orc_spmv_bcsr  spmv/1D/BCSR-block/host/app.c:70-85    (BCSR-nnz:71)
 *   orc_spmv_bcoo  spmv/1D/BCOO-block/host/app.c:72-86    (BCOO-nnz:72)
 *   orc_spmv_tiled spmv/2D/DCSR/host/app.c:84-103, DCOO:80-95, RBDCSR:88-103,
 *                  BDBCSR-block:92-114 and the other 2D spmv_host copies
 *   orc_merge_*    spmv/1D/COO-nnz/host/app.c:275-289, spmv/2D/DCSR/host/app.c:340-351
 *
 * The orc_time_* functions time these restatements for the CPU baseline.
 *
 * plus an independent high-precision reference (orc_ref_csr: long double
 * accumulation and the per-row sum of |a_ij x_j| for error bounds) and an
 * OpenMP row-parallel CSR used only as a labelled "not reference" CPU
 * figure.
 *
 * Parity status: the integer paths are pinned by modular arithmetic (a
 * numpy int64 computation reduced modulo 2^bits and dense brute force on
 * small inputs); the float paths by the long-double reference with a
 * rounding-error bound, scipy.sparse on the CPU, and closed forms
 * (identity, diagonal, stencil row sums).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#ifdef _OPENMP
#include <omp.h>
#endif

enum { I8 = 0, I16 = 1, I32 = 2, I64 = 3, F32 = 4, F64 = 5 };

#define DISPATCH(dt, FN, ...)                                         \
    switch (dt) {                                                     \
    case I8: FN##_i8(__VA_ARGS__); break;                             \
    case I16: FN##_i16(__VA_ARGS__); break;                           \
    case I32: FN##_i32(__VA_ARGS__); break;                           \
    case I64: FN##_i64(__VA_ARGS__); break;                           \
    case F32: FN##_f32(__VA_ARGS__); break;                           \
    case F64: FN##_f64(__VA_ARGS__); break;                           \
    default: return -1;                                               \
    }

/* ---- CSR: spmv/1D/CSR-row/host/app.c:66-77 ------------------------------ */
#define DEF_CSR(SUF, T)                                                                  \
    static void csr_##SUF(uint32_t n, const uint32_t *rp, const uint32_t *ci, const void *va, \
                          const void *xa, void *ya) {                                    \
        const T *v = (const T *)va, *x = (const T *)xa;                                  \
        T *y = (T *)ya;                                                                  \
        for (uint32_t r = 0; r < n; ++r) {                                               \
            T sum = 0;                                                                   \
            for (uint32_t i = rp[r]; i < rp[r + 1]; ++i) sum += x[ci[i]] * v[i];         \
            y[r] = sum;                                                                  \
        }                                                                                \
    }
DEF_CSR(i8, int8_t) DEF_CSR(i16, int16_t) DEF_CSR(i32, int32_t) DEF_CSR(i64, int64_t)
DEF_CSR(f32, float) DEF_CSR(f64, double)

int orc_spmv_csr(int dt, uint32_t n, const uint32_t *rp, const uint32_t *ci, const void *v,
                 const void *x, void *y) {
    DISPATCH(dt, csr, n, rp, ci, v, x, y);
    return 0;
}

/* ---- COO: spmv/1D/COO-row/host/app.c:70-74 (y zeroed first: calloc at :276) */
#define DEF_COO(SUF, T)                                                                  \
    static void coo_##SUF(uint32_t n, uint32_t nnz, const uint32_t *ri, const uint32_t *ci, \
                          const void *va, const void *xa, void *ya) {                    \
        const T *v = (const T *)va, *x = (const T *)xa;                                  \
        T *y = (T *)ya;                                                                  \
        memset(y, 0, (size_t)n * sizeof(T));                                             \
        for (uint32_t k = 0; k < nnz; k++) y[ri[k]] += x[ci[k]] * v[k];                  \
    }
DEF_COO(i8, int8_t) DEF_COO(i16, int16_t) DEF_COO(i32, int32_t) DEF_COO(i64, int64_t)
DEF_COO(f32, float) DEF_COO(f64, double)

int orc_spmv_coo(int dt, uint32_t n, uint32_t nnz, const uint32_t *ri, const uint32_t *ci,
                 const void *v, const void *x, void *y) {
    DISPATCH(dt, coo, n, nnz, ri, ci, v, x, y);
    return 0;
}

/* ---- BCSR: spmv/1D/BCSR-block/host/app.c:70-85 --------------------------- */
#define DEF_BCSR(SUF, T)                                                                  \
    static void bcsr_##SUF(uint32_t nrows, uint32_t ncols, uint32_t R, uint32_t C,        \
                           uint32_t nbr, const uint32_t *bp, const uint32_t *bc,          \
                           const void *va, const void *xa, void *ya) {                    \
        const T *bv = (const T *)va, *x = (const T *)xa;                                  \
        T *y = (T *)ya;                                                                   \
        memset(y, 0, (size_t)nrows * sizeof(T));                                          \
        for (uint64_t n = 0; n < nbr; n++)                                                \
            for (uint64_t i = bp[n]; i < bp[n + 1]; i++) {                                \
                uint64_t j = bc[i];                                                       \
                for (uint64_t r = 0; r < R; r++) {                                        \
                    T acc = 0;                                                            \
                    for (uint64_t c = 0; c < C; c++)                                      \
                        if (n * R + r < nrows && j * C + c < ncols)                       \
                            acc += bv[i * C * R + r * C + c] * x[j * C + c];              \
                    if (n * R + r < nrows) y[n * R + r] += acc;                           \
                }                                                                         \
            }                                                                             \
    }
DEF_BCSR(i8, int8_t) DEF_BCSR(i16, int16_t) DEF_BCSR(i32, int32_t) DEF_BCSR(i64, int64_t)
DEF_BCSR(f32, float) DEF_BCSR(f64, double)

int orc_spmv_bcsr(int dt, uint32_t nrows, uint32_t ncols, uint32_t R, uint32_t C, uint32_t nbr,
                  const uint32_t *bp, const uint32_t *bc, const void *bv, const void *x, void *y) {
    DISPATCH(dt, bcsr, nrows, ncols, R, C, nbr, bp, bc, bv, x, y);
    return 0;
}

/* ---- BCOO: spmv/1D/BCOO-block/host/app.c:72-86 --------------------------- */
#define DEF_BCOO(SUF, T)                                                                  \
    static void bcoo_##SUF(uint32_t nrows, uint32_t ncols, uint32_t R, uint32_t C,        \
                           uint64_t nb, const uint32_t *br, const uint32_t *bc,           \
                           const void *va, const void *xa, void *ya) {                    \
        const T *bv = (const T *)va, *x = (const T *)xa;                                  \
        T *y = (T *)ya;                                                                   \
        memset(y, 0, (size_t)nrows * sizeof(T));                                          \
        for (uint64_t n = 0; n < nb; n++) {                                               \
            uint64_t i = br[n], j = bc[n];                                                \
            for (uint64_t r = 0; r < R; r++) {                                            \
                T acc = 0;                                                                \
                for (uint64_t c = 0; c < C; c++)                                          \
                    if (i * R + r < nrows && j * C + c < ncols)                           \
                        acc += bv[n * C * R + r * C + c] * x[j * C + c];                  \
                if (i * R + r < nrows) y[i * R + r] += acc;                               \
            }                                                                             \
        }                                                                                 \
    }
DEF_BCOO(i8, int8_t) DEF_BCOO(i16, int16_t) DEF_BCOO(i32, int32_t) DEF_BCOO(i64, int64_t)
DEF_BCOO(f32, float) DEF_BCOO(f64, double)

int orc_spmv_bcoo(int dt, uint32_t nrows, uint32_t ncols, uint32_t R, uint32_t C, uint64_t nb,
                  const uint32_t *br, const uint32_t *bc, const void *bv, const void *x, void *y) {
    DISPATCH(dt, bcoo, nrows, ncols, R, C, nb, br, bc, bv, x, y);
    return 0;
}

/* ---- 2D tilings --------------------------------------------------------
 * Tile-major order as in every 2D spmv_host: for each tile t (row-major
 * tile order for D tilings, stripe order for RBD and BD), for each local row (or
 * entry / block), the tile's per-row sum is added to y.  Tile t covers rows
 * [row_off[t], row_off[t]+rows[t]) and columns starting at col_off[t].
 * fmt: 0 CSR (rowptr stride rows[t]+1 at ptr_off[t]), 1 COO, 2 BCSR
 * (browptr at ptr_off[t], nbr[t] block rows), 3 BCOO.  Entry / block arrays
 * are concatenated in tile order at ent_off[t]. */
#define DEF_TILED(SUF, T)                                                                       \
    static void tiled_##SUF(int fmt, uint32_t nrows, uint32_t ncols, uint32_t ntiles,           \
                            const uint32_t *row_off, const uint32_t *rows, const uint32_t *col_off, \
                            const uint32_t *col_w, const uint64_t *ptr_off, const uint64_t *ent_off, \
                            const uint32_t *ptr, const uint32_t *ri, const uint32_t *ci,        \
                            uint32_t R, uint32_t C, const void *va, const void *xa, void *ya) { \
        const T *v = (const T *)va, *x = (const T *)xa;                                         \
        T *y = (T *)ya;                                                                         \
        memset(y, 0, (size_t)nrows * sizeof(T));                                                \
        for (uint32_t t = 0; t < ntiles; t++) {                                                 \
            const uint32_t ro = row_off[t], co = col_off[t];                                    \
            const uint64_t e0 = ent_off[t], e1 = ent_off[t + 1];                                \
            if (fmt == 0) {                                                                     \
                const uint32_t *rp = ptr + ptr_off[t];                                          \
                for (uint32_t r = 0; r < rows[t]; r++) {                                        \
                    T sum = 0;                                                                  \
                    for (uint32_t k = rp[r]; k < rp[r + 1]; k++)                                \
                        sum += x[co + ci[e0 + k]] * v[e0 + k];                                  \
                    if (ro + r < nrows) y[ro + r] += sum;                                       \
                }                                                                               \
            } else if (fmt == 1) {                                                              \
                for (uint64_t k = e0; k < e1; k++) y[ro + ri[k]] += x[co + ci[k]] * v[k];       \
            } else {                                                                            \
                const uint64_t rc = (uint64_t)R * C;                                            \
                uint32_t nbrt = (rows[t] + R - 1) / R;                                          \
                for (uint32_t I = 0; I < (fmt == 2 ? nbrt : 1); I++) {                          \
                    uint64_t b0 = fmt == 2 ? e0 + ptr[ptr_off[t] + I] : e0;                     \
                    uint64_t b1 = fmt == 2 ? e0 + ptr[ptr_off[t] + I + 1] : e1;                 \
                    for (uint64_t b = b0; b < b1; b++) {                                        \
                        uint64_t bi = fmt == 2 ? I : ri[b], bj = ci[b];                         \
                        for (uint64_t r = 0; r < R; r++) {                                      \
                            T acc = 0;                                                          \
                            for (uint64_t c = 0; c < C; c++)                                    \
                                if (bi * R + r < rows[t] && bj * C + c < col_w[t] &&            \
                                    ro + bi * R + r < nrows && co + bj * C + c < ncols)         \
                                    acc += v[b * rc + r * C + c] * x[co + bj * C + c];          \
                            if (bi * R + r < rows[t] && ro + bi * R + r < nrows)                \
                                y[ro + bi * R + r] += acc;                                      \
                        }                                                                       \
                    }                                                                           \
                }                                                                               \
            }                                                                                   \
        }                                                                                       \
    }
DEF_TILED(i8, int8_t) DEF_TILED(i16, int16_t) DEF_TILED(i32, int32_t) DEF_TILED(i64, int64_t)
DEF_TILED(f32, float) DEF_TILED(f64, double)

int orc_spmv_tiled(int dt, int fmt, uint32_t nrows, uint32_t ncols, uint32_t ntiles,
                   const uint32_t *row_off, const uint32_t *rows, const uint32_t *col_off,
                   const uint32_t *col_w, const uint64_t *ptr_off, const uint64_t *ent_off,
                   const uint32_t *ptr, const uint32_t *ri, const uint32_t *ci, uint32_t R,
                   uint32_t C, const void *v, const void *x, void *y) {
    DISPATCH(dt, tiled, fmt, nrows, ncols, ntiles, row_off, rows, col_off, col_w, ptr_off, ent_off,
             ptr, ri, ci, R, C, v, x, y);
    return 0;
}

/* ---- Merges -------------------------------------------------------------
 * 2D vertical merge (spmv/2D/DCSR/host/app.c:340-351): for every row,
 * y[row] = part_0[row] + part_1[row] + ... in part order, OpenMP across
 * rows with `nthreads` threads (the reference -n flag). */
#define DEF_MERGE(SUF, T)                                                                   \
    static void merge_##SUF(uint32_t nparts, uint64_t n, uint64_t stride, void *pa, int nth) { \
        T *p = (T *)pa;                                                                     \
        _Pragma("omp parallel for num_threads(nth) schedule(static)")                       \
        for (int64_t i = 0; i < (int64_t)n; i++)                                            \
            for (uint32_t c = 1; c < nparts; c++) p[i] += p[(uint64_t)c * stride + i];      \
    }
DEF_MERGE(i8, int8_t) DEF_MERGE(i16, int16_t) DEF_MERGE(i32, int32_t) DEF_MERGE(i64, int64_t)
DEF_MERGE(f32, float) DEF_MERGE(f64, double)

int orc_merge_parts(int dt, uint32_t nparts, uint64_t n, uint64_t stride, void *parts, int nthreads) {
    if (nthreads < 1) nthreads = 1;
    DISPATCH(dt, merge, nparts, n, stride, parts, nthreads);
    return 0;
}

/* ---- High-precision reference ------------------------------------------
 * y_ref[r] = sum of a_rk * x_k in long double (64-bit significand: the
 * products of two floats are exact, those of two doubles are within 2^-64
 * relative), absb[r] = sum of |a_rk * x_k|.  Any summation order of the
 * same products in precision u satisfies |y - y_ref| <= gamma(len+1)*absb
 * with gamma(m) = m*u/(1-m*u) (Higham, Accuracy and Stability of Numerical
 * Algorithms, 2nd ed., eq. 3.5), which the tests use as the tolerance. */
static long double getv(int dt, const void *a, uint64_t i) {
    switch (dt) {
    case I8: return ((const int8_t *)a)[i];
    case I16: return ((const int16_t *)a)[i];
    case I32: return ((const int32_t *)a)[i];
    case I64: return (long double)((const int64_t *)a)[i];
    case F32: return ((const float *)a)[i];
    default: return ((const double *)a)[i];
    }
}

int orc_ref_csr(int dt, uint32_t n, const uint32_t *rp, const uint32_t *ci, const void *v,
                const void *x, double *y_ref, double *absb) {
    for (uint32_t r = 0; r < n; r++) {
        long double s = 0, a = 0;
        for (uint32_t k = rp[r]; k < rp[r + 1]; k++) {
            long double p = getv(dt, v, k) * getv(dt, x, ci[k]);
            s += p;
            a += fabsl(p);
        }
        y_ref[r] = (double)s;
        if (absb) absb[r] = (double)a;
    }
    return 0;
}

/* ---- CPU baselines -------------------------------------------------------
 * Times `iters` runs of the reference-order CSR spmv_host (single thread,
 * as the reference) and returns the median ms.  omp_threads > 0 instead
 * times an OpenMP row-parallel loop of the same arithmetic ("not
 * reference").  The run stops early once `budget_s` seconds are spent. */
static double now_s(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec + ts.tv_nsec * 1e-9;
}

static int cmp_d(const void *a, const void *b) {
    double x = *(const double *)a, y = *(const double *)b;
    return x < y ? -1 : x > y;
}

#define DEF_CSR_OMP(SUF, T)                                                                  \
    static void csr_omp_##SUF(uint32_t n, const uint32_t *rp, const uint32_t *ci, const void *va, \
                              const void *xa, void *ya, int nth) {                           \
        const T *v = (const T *)va, *x = (const T *)xa;                                      \
        T *y = (T *)ya;                                                                      \
        _Pragma("omp parallel for num_threads(nth) schedule(dynamic, 1024)")                 \
        for (int64_t r = 0; r < (int64_t)n; ++r) {                                           \
            T sum = 0;                                                                       \
            for (uint32_t i = rp[r]; i < rp[r + 1]; ++i) sum += x[ci[i]] * v[i];             \
            y[r] = sum;                                                                      \
        }                                                                                    \
    }
DEF_CSR_OMP(i8, int8_t) DEF_CSR_OMP(i16, int16_t) DEF_CSR_OMP(i32, int32_t)
DEF_CSR_OMP(i64, int64_t) DEF_CSR_OMP(f32, float) DEF_CSR_OMP(f64, double)

double orc_time_csr(int dt, uint32_t n, const uint32_t *rp, const uint32_t *ci, const void *v,
                    const void *x, void *y, int iters, int warmup, int omp_threads, double budget_s,
                    int *iters_done) {
    double t[1024];
    if (iters > 1024) iters = 1024;
    int done = 0;
    const double start = now_s();
    for (int it = -warmup; it < iters; it++) {
        double t0 = now_s();
        if (omp_threads > 0) {
            switch (dt) {
            case I8: csr_omp_i8(n, rp, ci, v, x, y, omp_threads); break;
            case I16: csr_omp_i16(n, rp, ci, v, x, y, omp_threads); break;
            case I32: csr_omp_i32(n, rp, ci, v, x, y, omp_threads); break;
            case I64: csr_omp_i64(n, rp, ci, v, x, y, omp_threads); break;
            case F32: csr_omp_f32(n, rp, ci, v, x, y, omp_threads); break;
            default: csr_omp_f64(n, rp, ci, v, x, y, omp_threads); break;
            }
        } else {
            orc_spmv_csr(dt, n, rp, ci, v, x, y);
        }
        double t1 = now_s();
        if (it >= 0) t[done++] = (t1 - t0) * 1e3;
        if (budget_s > 0 && t1 - start > budget_s && done >= 1) break;
    }
    if (iters_done) *iters_done = done;
    if (done == 0) return -1.0;
    qsort(t, (size_t)done, sizeof(double), cmp_d);
    return t[done / 2];
}

/* Times the other spmv_host restatements (single thread, as the
 * reference) and the 2D vertical merge (OpenMP over `nthreads`, the
 * reference's -n flag): median ms of `iters` runs after `warmup`, stopping
 * once `budget_s` seconds are spent.  Each run starts from the y the
 * restatement itself zeroes (the reference's calloc). */
typedef struct {
    int kind, dt, fmt;
    uint32_t nrows, ncols, n2, R, C;
    uint64_t nb;
    const uint32_t *a0, *a1, *a2, *a3, *a4, *a5, *a6, *a7, *a8;
    const uint64_t *p0, *p1;
    const void *v, *x;
    void *y;
    int nthreads;
} job_t;

static void run_job(const job_t *j) {
    switch (j->kind) {
    case 0: orc_spmv_coo(j->dt, j->nrows, j->n2, j->a0, j->a1, j->v, j->x, j->y); break;
    case 1: orc_spmv_bcsr(j->dt, j->nrows, j->ncols, j->R, j->C, j->n2, j->a0, j->a1, j->v, j->x, j->y); break;
    case 2: orc_spmv_bcoo(j->dt, j->nrows, j->ncols, j->R, j->C, j->nb, j->a0, j->a1, j->v, j->x, j->y); break;
    case 3:
        orc_spmv_tiled(j->dt, j->fmt, j->nrows, j->ncols, j->n2, j->a0, j->a1, j->a2, j->a3, j->p0, j->p1, j->a4,
                       j->a5, j->a6, j->R, j->C, j->v, j->x, j->y);
        break;
    default: orc_merge_parts(j->dt, j->n2, j->nb, (uint64_t)j->ncols, j->y, j->nthreads); break;
    }
}

static double time_job(const job_t *j, int iters, int warmup, double budget_s, int *iters_done) {
    double t[1024];
    if (iters > 1024) iters = 1024;
    int done = 0;
    const double start = now_s();
    for (int it = -warmup; it < iters; it++) {
        double t0 = now_s();
        run_job(j);
        double t1 = now_s();
        if (it >= 0) t[done++] = (t1 - t0) * 1e3;
        if (budget_s > 0 && t1 - start > budget_s && done >= 1) break;
    }
    if (iters_done) *iters_done = done;
    if (done == 0) return -1.0;
    qsort(t, (size_t)done, sizeof(double), cmp_d);
    return t[done / 2];
}

double orc_time_coo(int dt, uint32_t n, uint32_t nnz, const uint32_t *ri, const uint32_t *ci, const void *v,
                    const void *x, void *y, int iters, int warmup, double budget_s, int *iters_done) {
    job_t j = {0};
    j.kind = 0, j.dt = dt, j.nrows = n, j.n2 = nnz, j.a0 = ri, j.a1 = ci, j.v = v, j.x = x, j.y = y;
    return time_job(&j, iters, warmup, budget_s, iters_done);
}

double orc_time_bcsr(int dt, uint32_t nrows, uint32_t ncols, uint32_t R, uint32_t C, uint32_t nbr,
                     const uint32_t *bp, const uint32_t *bc, const void *bv, const void *x, void *y, int iters,
                     int warmup, double budget_s, int *iters_done) {
    job_t j = {0};
    j.kind = 1, j.dt = dt, j.nrows = nrows, j.ncols = ncols, j.R = R, j.C = C, j.n2 = nbr, j.a0 = bp, j.a1 = bc;
    j.v = bv, j.x = x, j.y = y;
    return time_job(&j, iters, warmup, budget_s, iters_done);
}

double orc_time_bcoo(int dt, uint32_t nrows, uint32_t ncols, uint32_t R, uint32_t C, uint64_t nb,
                     const uint32_t *br, const uint32_t *bc, const void *bv, const void *x, void *y, int iters,
                     int warmup, double budget_s, int *iters_done) {
    job_t j = {0};
    j.kind = 2, j.dt = dt, j.nrows = nrows, j.ncols = ncols, j.R = R, j.C = C, j.nb = nb, j.a0 = br, j.a1 = bc;
    j.v = bv, j.x = x, j.y = y;
    return time_job(&j, iters, warmup, budget_s, iters_done);
}

double orc_time_tiled(int dt, int fmt, uint32_t nrows, uint32_t ncols, uint32_t ntiles, const uint32_t *row_off,
                      const uint32_t *rows, const uint32_t *col_off, const uint32_t *col_w, const uint64_t *ptr_off,
                      const uint64_t *ent_off, const uint32_t *ptr, const uint32_t *ri, const uint32_t *ci,
                      uint32_t R, uint32_t C, const void *v, const void *x, void *y, int iters, int warmup,
                      double budget_s, int *iters_done) {
    job_t j = {0};
    j.kind = 3, j.dt = dt, j.fmt = fmt, j.nrows = nrows, j.ncols = ncols, j.n2 = ntiles;
    j.a0 = row_off, j.a1 = rows, j.a2 = col_off, j.a3 = col_w, j.p0 = ptr_off, j.p1 = ent_off;
    j.a4 = ptr, j.a5 = ri, j.a6 = ci, j.R = R, j.C = C, j.v = v, j.x = x, j.y = y;
    return time_job(&j, iters, warmup, budget_s, iters_done);
}

/* parts: nparts vectors of n values at `stride`; merged into the first. */
double orc_time_merge(int dt, uint32_t nparts, uint64_t n, uint64_t stride, void *parts, int nthreads, int iters,
                      int warmup, double budget_s, int *iters_done) {
    job_t j = {0};
    j.kind = 4, j.dt = dt, j.n2 = nparts, j.nb = n, j.ncols = (uint32_t)stride, j.y = parts;
    j.nthreads = nthreads < 1 ? 1 : nthreads;
    if (stride > UINT32_MAX) return -1.0;
    return time_job(&j, iters, warmup, budget_s, iters_done);
}

int orc_num_cpus(void) {
#ifdef _OPENMP
    return omp_get_num_procs();
#else
    return 1;
#endif
}
