/*
 * ref_harness2d.c -- TEST INFRASTRUCTURE ONLY.  Flat-array wrappers around
 * the reference's 2D support code (support/matrix.h + support/partition.h
 * of one spmv/2D/<dir>, compiled where it lies, see oracle/Makefile), so
 * the tests can compare this library's tilings and stripe partitioners
 * with the reference bit-for-bit.  Not product code.
 *
 * REF2 selects the directory family:
 *   1 DCSR   2 DCOO   3 RBDCSR   4 RBDCOO   5 BDCSR   6 BDCOO
 *   7 DBCOO  8 RBDBCOO (REF_BY_NNZ: -nnz)   9 BDBCOO (REF_BY_NNZ: -nnz)
 *  10 RBDBCSR (REF_BY_NNZ: -nnz)
 * Input is a COO sorted by (row, col), as every 2D app.c sorts first.
 */
#include <assert.h>
#include <stdbool.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "matrix.h"
#if __has_include("partition.h")
#include "partition.h"
#endif

#define EXPORT __attribute__((visibility("default")))

static struct COOMatrix *mk_coo(uint32_t nrows, uint32_t ncols, uint32_t nnz, const uint32_t *ri,
                                const uint32_t *ci, const void *vals) {
    struct COOMatrix *A = (struct COOMatrix *)calloc(1, sizeof(*A));
    A->nrows = nrows; A->ncols = ncols; A->nnz = nnz;
    A->rows = (uint32_t *)calloc(nrows + 1, sizeof(uint32_t));
    A->nnzs = (struct elem_t *)calloc(nnz + 8, sizeof(struct elem_t));
    for (uint32_t k = 0; k < nnz; k++) {
        A->nnzs[k].rowind = ri[k];
        A->nnzs[k].colind = ci[k];
        A->nnzs[k].val = ((const val_dt *)vals)[k];
        A->rows[ri[k]]++;
    }
    return A;
}

EXPORT int ref2_val_size(void) { return (int)sizeof(val_dt); }

#if REF2 == 1 || REF2 == 3 || REF2 == 5
#if REF2 == 1
static struct DCSRMatrix *g;
#elif REF2 == 3
static struct RBDCSRMatrix *g;
#else
static struct BDCSRMatrix *g;
#endif
/* dims: {npartitions, rows per tile (D) or nrows, tile width or max width} */
EXPORT void ref2_build(uint32_t nrows, uint32_t ncols, uint32_t nnz, const uint32_t *ri, const uint32_t *ci,
                       const void *vals, int horz, int vert, uint32_t *dims) {
    struct COOMatrix *A = mk_coo(nrows, ncols, nnz, ri, ci, vals);
#if REF2 == 1
    g = coo2dcsr(A, horz, vert);
    dims[0] = g->npartitions; dims[1] = g->tile_height; dims[2] = g->tile_width;
#elif REF2 == 3
    g = coo2rbdcsr(A, horz, vert);
    dims[0] = g->npartitions; dims[1] = g->nrows; dims[2] = g->tile_width;
#else
    g = coo2bdcsr(A, horz, vert);
    dims[0] = g->npartitions; dims[1] = g->nrows; dims[2] = g->max_tile_width;
#endif
}
EXPORT void ref2_copy_csr(uint32_t *drowptr, uint32_t *dcolind, void *dval, uint32_t *cuts) {
#if REF2 == 1
    uint32_t th = g->tile_height;
#else
    uint32_t th = g->nrows;
#endif
    memcpy(drowptr, g->drowptr, (size_t)g->npartitions * (th + 1) * sizeof(uint32_t));
    memcpy(dcolind, g->dcolind, g->nnz * sizeof(uint32_t));
    memcpy(dval, g->dval, g->nnz * sizeof(val_dt));
#if REF2 == 5
    memcpy(cuts, g->vert_tile_widths, (g->vert_partitions + 1) * sizeof(uint32_t));
#else
    (void)cuts;
#endif
}
#if REF2 != 1
/* partition_by_nnz per stripe: row_split[vert*(horz+1)] */
EXPORT void ref2_partition(uint32_t *row_split) {
    struct partition_info_t *pi = partition_init(g, g->vert_partitions * g->horz_partitions, 1, 1);
    partition_by_nnz(g, pi);
    memcpy(row_split, pi->row_split, (size_t)g->vert_partitions * (g->horz_partitions + 1) * sizeof(uint32_t));
}
#endif
#endif

#if REF2 == 2 || REF2 == 4 || REF2 == 6
#if REF2 == 2
static struct DCOOMatrix *g;
#elif REF2 == 4
static struct RBDCOOMatrix *g;
#else
static struct BDCOOMatrix *g;
#endif
EXPORT void ref2_build(uint32_t nrows, uint32_t ncols, uint32_t nnz, const uint32_t *ri, const uint32_t *ci,
                       const void *vals, int horz, int vert, uint32_t *dims) {
    struct COOMatrix *A = mk_coo(nrows, ncols, nnz, ri, ci, vals);
#if REF2 == 2
    g = coo2dcoo(A, horz, vert);
#elif REF2 == 4
    g = coo2rbdcoo(A, horz, vert);
#else
    g = coo2bdcoo(A, horz, vert);
#endif
    dims[0] = g->npartitions; dims[1] = 0; dims[2] = 0;
}
EXPORT void ref2_copy_coo(uint32_t *rowind, uint32_t *colind, void *vals) {
    for (uint32_t k = 0; k < g->nnz; k++) {
        rowind[k] = g->nnzs[k].rowind;
        colind[k] = g->nnzs[k].colind;
        ((val_dt *)vals)[k] = g->nnzs[k].val;
    }
}
#if REF2 != 2
/* entry splits per stripe: nnz_split[vert*(horz+1)] */
EXPORT void ref2_partition(uint32_t *nnz_split) {
    struct partition_info_t *pi = partition_init(g, g->vert_partitions * g->horz_partitions, 1, 1);
    partition_by_nnz(g, pi);
    memcpy(nnz_split, pi->nnz_split, (size_t)g->vert_partitions * (g->horz_partitions + 1) * sizeof(uint32_t));
}
#endif
#endif

#if REF2 == 7 || REF2 == 8 || REF2 == 9 || REF2 == 10
#if REF2 == 7
static struct DBCOOMatrix *g;
#elif REF2 == 8
static struct RBDBCOOMatrix *g;
#elif REF2 == 9
static struct BDBCOOMatrix *g;
#else
static struct RBDBCSRMatrix *g;
#endif
/* block tiles (count by values, sorted, then as BCOO); dims: {npartitions, nblocks, num_block_rows} */
EXPORT void ref2_build(uint32_t nrows, uint32_t ncols, uint32_t nnz, const uint32_t *ri, const uint32_t *ci,
                       const void *vals, int horz, int vert, uint32_t R, uint32_t C, uint32_t *dims) {
    struct COOMatrix *A = mk_coo(nrows, ncols, nnz, ri, ci, vals);
#if REF2 == 7
    struct DCSRMatrix *d = coo2dcsr(A, horz, vert);
    struct DBCSRMatrix *b = dcsr2dbcsr(d, R, C);
    sortDBCSRMatrix(b); /* spmv/2D/DBCOO/host/app.c:139-143 order */
    countNNZperBlockDBCSRMatrix(b);
    g = dbcsr2dbcoo(b);
#elif REF2 == 8
    struct RBDCSRMatrix *d = coo2rbdcsr(A, horz, vert);
    struct RBDBCSRMatrix *b = rbdcsr2rbdbcsr(d, R, C);
    sortRBDBCSRMatrix(b); /* spmv/2D/RBDBCOO-block/host/app.c:149-152 order */
    countNNZperBlockRBDBCSRMatrix(b);
    g = rbdbcsr2rbdbcoo(b);
#elif REF2 == 9
    struct BDCSRMatrix *d = coo2bdcsr(A, horz, vert);
    struct BDBCSRMatrix *b = bdcsr2bdbcsr(d, R, C);
    sortBDBCSRMatrix(b); /* spmv/2D/BDBCOO-nnz/host/app.c:145-148 order */
    countNNZperBlockBDBCSRMatrix(b);
    g = bdbcsr2bdbcoo(b);
#else
    struct RBDCSRMatrix *d = coo2rbdcsr(A, horz, vert);
    g = rbdcsr2rbdbcsr(d, R, C);
    countNNZperBlockRBDBCSRMatrix(g);
#endif
    dims[0] = g->npartitions; dims[1] = g->nblocks; dims[2] = g->num_block_rows;
}
#if REF2 != 10
EXPORT void ref2_copy_bcoo(uint32_t *browind, uint32_t *bcolind, void *bval, uint32_t *npb, uint32_t *bpp) {
    uint64_t rc = g->row_block_size * g->col_block_size;
    for (uint32_t k = 0; k < g->nblocks; k++) {
        browind[k] = g->bind[k].rowind;
        bcolind[k] = g->bind[k].colind;
    }
    memcpy(bval, g->bval, (size_t)g->nblocks * rc * sizeof(val_dt));
    memcpy(npb, g->nnz_per_block, g->nblocks * sizeof(uint32_t));
#if REF2 == 7
    memcpy(bpp, g->blocks_per_partition, g->npartitions * sizeof(uint32_t));
#else
    memcpy(bpp, g->blocks_per_vert_partition, g->npartitions * sizeof(uint32_t));
#endif
}
#endif
#if REF2 != 7
/* stripe partition: blocks_dpu / nnzs_dpu [vert*horz] and, for BCSR
 * stripes, brow_split [vert*(horz+1)] */
EXPORT void ref2_partition(uint32_t *blocks_dpu, uint32_t *nnzs_dpu, uint32_t *brow_split) {
    struct partition_info_t *pi = partition_init(g, g->vert_partitions * g->horz_partitions, 1, 1);
#if REF_BY_NNZ
    partition_by_nnz(g, pi);
#else
    partition_by_block(g, pi);
#endif
    memcpy(blocks_dpu, pi->blocks_dpu, (size_t)g->vert_partitions * g->horz_partitions * sizeof(uint32_t));
    memcpy(nnzs_dpu, pi->nnzs_dpu, (size_t)g->vert_partitions * g->horz_partitions * sizeof(uint32_t));
#if REF2 == 10
    memcpy(brow_split, pi->brow_split, (size_t)g->vert_partitions * (g->horz_partitions + 1) * sizeof(uint32_t));
#else
    (void)brow_split;
#endif
}
#endif
#endif
