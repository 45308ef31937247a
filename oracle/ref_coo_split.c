/*
 * ref_coo_split.c -- TEST INFRASTRUCTURE ONLY.  Runs the reference's own
 * COO-nnz inter-DPU split (straddling rows, spmv/1D/COO-nnz/host/app.c:118-205)
 * and its even tasklet split (the same loop body, :186-204) on flat arrays,
 * so the tests can compare sp_partition_coo_nnz and sp_partition_even with
 * them bit-for-bit.
 *
 * The split lives inside main() of app.c, which needs the absent UPMEM SDK
 * (dpu.h), so oracle/Makefile copies two spans of that file's text at build
 * time into oracle/_ref (never committed, never shipped): the struct
 * dpu_info_t definition (app.c:45-53) and the body of the DPU_FOREACH loop
 * (app.c:119-205).  This file supplies the loop around that body and the
 * variables it reads and writes; the body itself calls nothing from the SDK.
 * The reference's support/common.h (dpu_arguments_t) and support/matrix.h
 * (struct COOMatrix) are included where they lie.  Nothing here is product
 * code.
 */
#include <stdbool.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "common.h"
#include "matrix.h"

#define EXPORT __attribute__((visibility("default")))

#include "coo_nnz_info.inc"

/* rows[nrows] = entries per row (struct COOMatrix.rows after sortCOOMatrix).
 * Outputs per part i < nparts: prev_rows, rows_per_part, prev_nnz,
 * nnz_per_part, merge (dpu_info_t fields) and, for the first `tsk_parts`
 * parts, the tasklet split start_nnz / nnz_per_tasklet [NR_TASKLETS] each.
 * Returns NR_TASKLETS. */
EXPORT int ref_coo_nnz_split(uint32_t nrows, uint32_t nnz_total, const uint32_t *rows_in,
                             uint32_t nparts, uint32_t *o_prev_rows, uint32_t *o_rows,
                             uint32_t *o_prev_nnz, uint32_t *o_nnz, uint32_t *o_merge,
                             uint32_t tsk_parts, uint32_t *o_tstart, uint32_t *o_tcount) {
    /* The body reads A->rows[r-1] with r = 0 when a part gets no entries
     * (nparts > nnz); a leading zero keeps that read inside the buffer, and
     * trailing zeros keep r = nrows readable. */
    uint32_t *pad = (uint32_t *)calloc((size_t)nrows + nparts + 8, sizeof(uint32_t));
    if (!pad) return -1;
    memcpy(pad + 1, rows_in, (size_t)nrows * sizeof(uint32_t));
    struct COOMatrix Ab;
    memset(&Ab, 0, sizeof(Ab));
    Ab.nrows = nrows;
    Ab.nnz = nnz_total;
    Ab.rows = pad + 1;
    struct COOMatrix *A = &Ab;
    uint32_t nr_of_dpus = nparts;
    struct dpu_info_t *dpu_info = (struct dpu_info_t *)calloc(nparts, sizeof(struct dpu_info_t));
    dpu_arguments_t *input_args = (dpu_arguments_t *)calloc(nparts, sizeof(dpu_arguments_t));
    uint64_t max_rows_per_dpu = 0;
    uint64_t max_nnz_per_dpu = 0;
    uint32_t prev_row = 0;
    uint32_t prev_nnz = 0;
    uint32_t i;
    for (i = 0; i < nr_of_dpus; i++) {
#include "coo_nnz_split_body.inc"
    }
    (void)max_rows_per_dpu;
    (void)max_nnz_per_dpu;
    for (i = 0; i < nparts; i++) {
        o_prev_rows[i] = dpu_info[i].prev_rows_dpu;
        o_rows[i] = dpu_info[i].rows_per_dpu;
        o_prev_nnz[i] = dpu_info[i].prev_nnz_dpu;
        o_nnz[i] = dpu_info[i].nnz;
        o_merge[i] = dpu_info[i].merge;
    }
    for (i = 0; i < tsk_parts && i < nparts; i++)
        for (uint32_t t = 0; t < NR_TASKLETS; t++) {
            o_tstart[(size_t)i * NR_TASKLETS + t] = input_args[i].start_nnz[t];
            o_tcount[(size_t)i * NR_TASKLETS + t] = input_args[i].nnz_per_tasklet[t];
        }
    free(dpu_info);
    free(input_args);
    free(pad);
    return NR_TASKLETS;
}
