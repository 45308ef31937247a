/*
 * ref_harness.c -- TEST INFRASTRUCTURE ONLY.  A thin C shim compiled
 * together with the reference's own support/matrix.h and
 * support/partition.h, where they lie under /root/reference (see
 * oracle/Makefile), into oracle/_ref/libref_<dir>_<TYPE>.so.  It exports
 * flat-array wrappers so the tests can compare this library's host surface
 * (loader, builders, partitioners) with the reference implementation
 * bit-for-bit.  Nothing here is product code.
 *
 * One kernel directory per build, selected by REF_KIND:
 *   1 CSR-row / CSR-nnz (REF_BY_NNZ)        3 BCSR-block / BCSR-nnz (REF_BY_NNZ)
 *   2 COO-row / COO-nnz-rgrn (REF_BY_NNZ)   4 BCOO-block / BCOO-nnz (REF_BY_NNZ)
 * Partition arrays are allocated larger than partition_init does, so the
 * reference's one-past writes (e.g. spmv/1D/CSR-nnz/support/partition.h:63)
 * stay inside the buffer.
 */
#include <assert.h>
#include <stdbool.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "matrix.h"
#include "partition.h"

#define EXPORT __attribute__((visibility("default")))

EXPORT int ref_val_size(void) { return (int)sizeof(val_dt); }
EXPORT void ref_srand(unsigned s) { srand(s); }

static struct partition_info_t *pinfo(uint32_t P, uint32_t T) {
    struct partition_info_t *pi = (struct partition_info_t *)calloc(1, sizeof(*pi));
    size_t big = (size_t)(P + 4) * (T + 4) + 16;
#if REF_KIND == 1 || REF_KIND == 2
    pi->row_split = (uint32_t *)calloc(P + 4, sizeof(uint32_t));
    pi->row_split_tasklet = (uint32_t *)calloc(T + 4, sizeof(uint32_t));
    (void)big;
#elif REF_KIND == 3
    pi->brow_split = (uint32_t *)calloc(P + 4, sizeof(uint32_t));
    pi->blocks_dpu = (uint32_t *)calloc(P + 4, sizeof(uint32_t));
    pi->nnzs_dpu = (uint32_t *)calloc(P + 4, sizeof(uint32_t));
    pi->brow_split_tasklet = (uint32_t *)calloc(big, sizeof(uint32_t));
#elif REF_KIND == 4
    pi->block_split = (uint32_t *)calloc(P + 4, sizeof(uint32_t));
    pi->nnzs_dpu = (uint32_t *)calloc(P + 4, sizeof(uint32_t));
    pi->block_split_tasklet = (uint32_t *)calloc(big, sizeof(uint32_t));
#endif
    return pi;
}

/* ---------------- loader ---------------- */
static struct COOMatrix *g_coo;

/* readCOOMatrix; dims = {nrows, ncols, nnz}.  path must contain a '/'. */
EXPORT int ref_load_mtx(const char *path, uint32_t *dims) {
    g_coo = readCOOMatrix(path);
    if (!g_coo) return -1;
    dims[0] = g_coo->nrows;
    dims[1] = g_coo->ncols;
    dims[2] = g_coo->nnz;
    return 0;
}

EXPORT void ref_copy_coo(uint32_t *ri, uint32_t *ci, void *vals) {
    val_dt *v = (val_dt *)vals;
    for (uint32_t k = 0; k < g_coo->nnz; k++) {
#if REF_KIND == 2
        ri[k] = g_coo->nnzs[k].rowind;
        ci[k] = g_coo->nnzs[k].colind;
        v[k] = g_coo->nnzs[k].val;
#else
        ri[k] = g_coo->rowindx[k];
        ci[k] = g_coo->colind[k];
        v[k] = g_coo->values[k];
#endif
    }
}

#if REF_KIND == 2
/* sortCOOMatrix on a copy of (ri, ci, vals); outputs sorted arrays and rows[]. */
EXPORT void ref_sort_coo(uint32_t nrows, uint32_t ncols, uint32_t nnz, const uint32_t *ri,
                         const uint32_t *ci, const void *vals, uint32_t *ri_o, uint32_t *ci_o,
                         void *v_o, uint32_t *rows_o) {
    struct COOMatrix A;
    A.nrows = nrows; A.ncols = ncols; A.nnz = nnz;
    A.rows = (uint32_t *)calloc(nrows + 1, sizeof(uint32_t));
    A.nnzs = (struct elem_t *)calloc(nnz + 8, sizeof(struct elem_t));
    for (uint32_t k = 0; k < nnz; k++) {
        A.nnzs[k].rowind = ri[k];
        A.nnzs[k].colind = ci[k];
        A.nnzs[k].val = ((const val_dt *)vals)[k];
    }
    sortCOOMatrix(&A);
    for (uint32_t k = 0; k < nnz; k++) {
        ri_o[k] = A.nnzs[k].rowind;
        ci_o[k] = A.nnzs[k].colind;
        ((val_dt *)v_o)[k] = A.nnzs[k].val;
    }
    memcpy(rows_o, A.rows, nrows * sizeof(uint32_t));
    free(A.rows);
    free(A.nnzs);
}

/* partition_by_row / partition_by_nnz on COO (uses rows[]). */
EXPORT void ref_partition(uint32_t nrows, uint32_t nnz, const uint32_t *rows, int P, uint32_t *out) {
    struct COOMatrix A;
    A.nrows = nrows; A.ncols = nrows; A.nnz = nnz;
    A.rows = (uint32_t *)rows;
    struct partition_info_t *pi = pinfo(P, 1);
#if REF_BY_NNZ
    partition_by_nnz(&A, pi, P);
#else
    partition_by_row(&A, pi, P);
#endif
    memcpy(out, pi->row_split, (P + 1) * sizeof(uint32_t));
}

EXPORT void ref_tsklt_by_row(int rows_per_dpu, int T, uint32_t *out) {
    struct partition_info_t *pi = pinfo(1, T);
    partition_tsklt_by_row(pi, rows_per_dpu, T);
    memcpy(out, pi->row_split_tasklet, (T + 1) * sizeof(uint32_t));
}

EXPORT void ref_tsklt_by_nnz(uint32_t nrows, uint32_t nnz, const uint32_t *rows, int rows_per_dpu,
                             int nnz_per_dpu, int prev_rows_dpu, int T, uint32_t *out) {
    struct COOMatrix A;
    A.nrows = nrows; A.ncols = nrows; A.nnz = nnz;
    A.rows = (uint32_t *)rows;
    struct partition_info_t *pi = pinfo(1, T);
    partition_tsklt_by_nnz(&A, pi, rows_per_dpu, nnz_per_dpu, prev_rows_dpu, T);
    memcpy(out, pi->row_split_tasklet, (T + 1) * sizeof(uint32_t));
}
#endif

#if REF_KIND == 1 || REF_KIND == 3 || REF_KIND == 4
static struct COOMatrix *mk_coo(uint32_t nrows, uint32_t ncols, uint32_t nnz, const uint32_t *ri,
                                const uint32_t *ci, const void *vals) {
    struct COOMatrix *A = (struct COOMatrix *)malloc(sizeof(*A));
    A->nrows = nrows; A->ncols = ncols; A->nnz = nnz;
    A->rowindx = (uint32_t *)malloc((nnz + 1) * sizeof(uint32_t));
    A->colind = (uint32_t *)malloc((nnz + 1) * sizeof(uint32_t));
    A->values = (val_dt *)malloc((nnz + 1) * sizeof(val_dt));
    memcpy(A->rowindx, ri, nnz * sizeof(uint32_t));
    memcpy(A->colind, ci, nnz * sizeof(uint32_t));
    memcpy(A->values, vals, nnz * sizeof(val_dt));
    return A;
}

/* coo2csr */
EXPORT void ref_coo2csr(uint32_t nrows, uint32_t ncols, uint32_t nnz, const uint32_t *ri,
                        const uint32_t *ci, const void *vals, uint32_t *rowptr, uint32_t *colind,
                        void *v_o) {
    struct COOMatrix *A = mk_coo(nrows, ncols, nnz, ri, ci, vals);
    struct CSRMatrix *C = coo2csr(A);
    memcpy(rowptr, C->rowptr, (nrows + 1) * sizeof(uint32_t));
    memcpy(colind, C->colind, nnz * sizeof(uint32_t));
    memcpy(v_o, C->values, nnz * sizeof(val_dt));
    freeCOOMatrix(A);
    freeCSRMatrix(C);
}

static struct CSRMatrix csr_view(uint32_t nrows, uint32_t nnz, const uint32_t *rowptr) {
    struct CSRMatrix C;
    C.nrows = nrows; C.ncols = nrows; C.nnz = nnz;
    C.rowptr = (uint32_t *)rowptr;
    C.colind = NULL;
    C.values = NULL;
    return C;
}
#endif

#if REF_KIND == 1
EXPORT void ref_partition(uint32_t nrows, uint32_t nnz, const uint32_t *rowptr, int P, uint32_t *out) {
    struct CSRMatrix C = csr_view(nrows, nnz, rowptr);
    struct partition_info_t *pi = pinfo(P, 1);
#if REF_BY_NNZ
    partition_by_nnz(&C, pi, P);
#else
    partition_by_row(&C, pi, P);
#endif
    memcpy(out, pi->row_split, (P + 1) * sizeof(uint32_t));
}

EXPORT void ref_tsklt_by_row(int rows_per_dpu, int T, uint32_t *out) {
    struct partition_info_t *pi = pinfo(1, T);
    partition_tsklt_by_row(pi, rows_per_dpu, T);
    memcpy(out, pi->row_split_tasklet, (T + 1) * sizeof(uint32_t));
}

EXPORT void ref_tsklt_by_nnz(uint32_t nrows, uint32_t nnz, const uint32_t *rowptr, int rows_per_dpu,
                             int nnz_per_dpu, int prev_rows_dpu, int T, uint32_t *out) {
    struct CSRMatrix C = csr_view(nrows, nnz, rowptr);
    struct partition_info_t *pi = pinfo(1, T);
    partition_tsklt_by_nnz(&C, pi, rows_per_dpu, nnz_per_dpu, prev_rows_dpu, T);
    memcpy(out, pi->row_split_tasklet, (T + 1) * sizeof(uint32_t));
}
#endif

#if REF_KIND == 3 || REF_KIND == 4
static struct BCSRMatrix *g_bcsr;

/* csr2bcsr + sortBCSRMatrix + countNNZperBlockBCSRMatrix on a CSR built
 * by the reference coo2csr; dims = {num_block_rows, num_block_cols, nb}. */
EXPORT void ref_build_bcsr(uint32_t nrows, uint32_t ncols, uint32_t nnz, const uint32_t *ri,
                           const uint32_t *ci, const void *vals, uint32_t R, uint32_t C, int sort,
                           uint64_t *dims) {
    struct COOMatrix *A = mk_coo(nrows, ncols, nnz, ri, ci, vals);
    struct CSRMatrix *S = coo2csr(A);
    g_bcsr = csr2bcsr(S, R, C);
    if (sort) sortBCSRMatrix(g_bcsr);
    countNNZperBlockBCSRMatrix(g_bcsr);
    dims[0] = g_bcsr->num_block_rows;
    dims[1] = g_bcsr->num_block_cols;
    dims[2] = g_bcsr->num_blocks;
    freeCOOMatrix(A);
    freeCSRMatrix(S);
}

EXPORT void ref_copy_bcsr(uint32_t *browptr, uint32_t *bcolind, void *bval, uint32_t *npb) {
    uint64_t rc = g_bcsr->row_block_size * g_bcsr->col_block_size;
    memcpy(browptr, g_bcsr->browptr, (g_bcsr->num_block_rows + 1) * sizeof(uint32_t));
    memcpy(bcolind, g_bcsr->bcolind, g_bcsr->num_blocks * sizeof(uint32_t));
    memcpy(bval, g_bcsr->bval, g_bcsr->num_blocks * rc * sizeof(val_dt));
    memcpy(npb, g_bcsr->nnz_per_block, g_bcsr->num_blocks * sizeof(uint32_t));
}
#endif

#if REF_KIND == 3
/* partition_by_blocks / partition_by_nnz over the built BCSR, with the
 * nnz_per_block array overridden by `npb` (so stored-entry counts can be
 * fed in), then each part's tasklet split (by_nnz selects
 * partition_tsklt_by_nnz).  Outputs brow_split[P+1], blocks[P], nnzs[P],
 * tsplit[P*(T+1)]. */
EXPORT void ref_partition_bcsr(const uint32_t *npb, uint32_t nnz, int P, int T, int tsklt_by_nnz,
                               uint32_t *brow_split, uint32_t *blocks, uint32_t *nnzs, uint32_t *tsplit) {
    memcpy(g_bcsr->nnz_per_block, npb, g_bcsr->num_blocks * sizeof(uint32_t));
    uint32_t keep = g_bcsr->nnz;
    g_bcsr->nnz = nnz;
    struct partition_info_t *pi = pinfo(P, T);
#if REF_BY_NNZ
    partition_by_nnz(g_bcsr, pi, P);
#else
    partition_by_blocks(g_bcsr, pi, P);
#endif
    memcpy(brow_split, pi->brow_split, (P + 1) * sizeof(uint32_t));
    memcpy(blocks, pi->blocks_dpu, P * sizeof(uint32_t));
    memcpy(nnzs, pi->nnzs_dpu, P * sizeof(uint32_t));
    for (int d = 0; d < P; d++) {
        if (tsklt_by_nnz) partition_tsklt_by_nnz(g_bcsr, pi, d, T, P);
        else partition_tsklt_by_block(g_bcsr, pi, d, T, P);
        memcpy(tsplit + (size_t)d * (T + 1), pi->brow_split_tasklet + (size_t)d * (T + 2),
               (T + 1) * sizeof(uint32_t));
    }
    g_bcsr->nnz = keep;
}
#endif

#if REF_KIND == 4
static struct BCOOMatrix *g_bcoo;

EXPORT void ref_bcsr2bcoo(uint32_t *browind, uint32_t *bcolind, void *bval, uint32_t *npb) {
    g_bcoo = bcsr2bcoo(g_bcsr);
    uint64_t rc = g_bcoo->row_block_size * g_bcoo->col_block_size;
    for (uint64_t i = 0; i < g_bcoo->num_blocks; i++) {
        browind[i] = g_bcoo->bind[i].rowind;
        bcolind[i] = g_bcoo->bind[i].colind;
    }
    memcpy(bval, g_bcoo->bval, g_bcoo->num_blocks * rc * sizeof(val_dt));
    memcpy(npb, g_bcoo->nnz_per_block, g_bcoo->num_blocks * sizeof(uint32_t));
}

/* partition_by_blocks (BCOO-block) / partition_by_nnz (BCOO-nnz) plus the
 * per-part tasklet splits. */
EXPORT void ref_partition_bcoo(const uint32_t *npb, uint32_t nnz, int P, int T, int tsklt_by_nnz,
                               uint32_t *block_split, uint32_t *nnzs, uint32_t *tsplit) {
    memcpy(g_bcoo->nnz_per_block, npb, g_bcoo->num_blocks * sizeof(uint32_t));
    uint32_t keep = g_bcoo->nnz;
    g_bcoo->nnz = nnz;
    struct partition_info_t *pi = pinfo(P, T);
#if REF_BY_NNZ
    partition_by_nnz(g_bcoo, pi, P);
#else
    partition_by_blocks(g_bcoo, pi, P);
#endif
    memcpy(block_split, pi->block_split, (P + 1) * sizeof(uint32_t));
    memcpy(nnzs, pi->nnzs_dpu, P * sizeof(uint32_t));
    for (int d = 0; d < P; d++) {
        if (tsklt_by_nnz) partition_tsklt_by_nnz(g_bcoo, pi, d, T, P);
        else partition_tsklt_by_block(g_bcoo, pi, d, T, P);
        memcpy(tsplit + (size_t)d * (T + 1), pi->block_split_tasklet + (size_t)d * (T + 2),
               (T + 1) * sizeof(uint32_t));
    }
    g_bcoo->nnz = keep;
}
#endif
