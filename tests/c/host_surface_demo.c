/* A C caller of the host surface, written the way a SparseP host program
 * would use it: load / generate, convert, partition, and check a few
 * invariants.  Prints "ok" and exits 0 on success.  With a GPU present it
 * also runs one CSR-row SpMV through the device API and compares it with a
 * plain C loop (integers, exact). */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "sparsep.h"

#define CHECK(c)                                                          \
    do {                                                                  \
        if (!(c)) {                                                       \
            fprintf(stderr, "check failed: %s (line %d)\n", #c, __LINE__); \
            return 1;                                                     \
        }                                                                 \
    } while (0)

int main(void) {
    sp_dist off = {SP_DIST_UNIFORM, -8, 8};
    sp_coo *A = NULL;
    CHECK(sp_gen_stencil5(64, 64, SP_I32, off, 4.0, 1, &A) == SP_OK);
    CHECK(A->nrows == 4096 && A->nnz == 5 * 4096 - 4 * 64);
    sp_csr *C = NULL;
    CHECK(sp_coo2csr(A, &C) == SP_OK);
    CHECK(C->rowptr[C->nrows] == C->nnz);

    uint32_t split[9];
    CHECK(sp_partition_by_nnz(C->rowptr, C->nrows, 8, split) == SP_OK);
    CHECK(split[0] == 0 && split[8] == C->nrows);
    for (int i = 0; i < 8; i++) CHECK(split[i] <= split[i + 1]);

    sp_bcsr *B = NULL;
    CHECK(sp_csr2bcsr(C, 4, 4, &B) == SP_OK);
    CHECK(sp_sort_bcsr(B) == SP_OK);
    CHECK(sp_count_nnz_per_block(B, C, 0) == SP_OK);
    uint64_t total = 0;
    for (uint64_t b = 0; b < B->num_blocks; b++) total += B->nnz_per_block[b];
    CHECK(total == C->nnz);

    float feat[SP_NFEAT];
    CHECK(sp_features(A, feat) == SP_OK);

    sp_ctx *ctx = NULL;
    int rc = sp_ctx_create(0, &ctx);
    if (rc == SP_OK) {
        int32_t *x = malloc(A->ncols * sizeof(int32_t)), *y = malloc(A->nrows * sizeof(int32_t));
        sp_dist xd = {SP_DIST_UNIFORM, -4, 4};
        CHECK(sp_gen_vector(SP_I32, A->ncols, xd, 1, x) == SP_OK);
        sp_plan *p = NULL;
        sp_opts o;
        sp_opts_default(&o);
        CHECK(sp_plan_create_csr(ctx, SP_K_CSR_ROW, C, &o, &p) == SP_OK);
        void *dx = NULL, *dy = NULL;
        CHECK(sp_dmalloc(ctx, A->ncols * 4, &dx) == SP_OK && sp_dmalloc(ctx, A->nrows * 4, &dy) == SP_OK);
        CHECK(sp_memcpy_h2d(ctx, dx, x, A->ncols * 4) == SP_OK);
        CHECK(sp_spmv(p, dx, dy) == SP_OK && sp_ctx_sync(ctx) == SP_OK);
        CHECK(sp_memcpy_d2h(ctx, y, dy, A->nrows * 4) == SP_OK);
        const int32_t *v = (const int32_t *)C->values;
        for (uint32_t r = 0; r < C->nrows; r++) {
            uint32_t acc = 0; /* the reference's wrapping int arithmetic */
            for (uint32_t k = C->rowptr[r]; k < C->rowptr[r + 1]; k++) acc += (uint32_t)v[k] * (uint32_t)x[C->colind[k]];
            CHECK(y[r] == (int32_t)acc);
        }
        sp_dfree(ctx, dx);
        sp_dfree(ctx, dy);
        sp_plan_destroy(p);
        sp_ctx_destroy(ctx);
        free(x);
        free(y);
        printf("ok (gpu)\n");
    } else {
        CHECK(rc == SP_ERR_NODEVICE || rc == SP_ERR_HIP);
        printf("ok\n");
    }
    sp_free_bcsr(B);
    sp_free_csr(C);
    sp_free_coo(A);
    return 0;
}
