/* 1D partitioners (reference: support/partition.h and the inline splits of
 * spmv/1D/COO-nnz/host/app.c).  Each function reproduces the reference's
 * split arrays exactly, including its greedy-threshold rules; the one-past
 * writes some reference versions make into a P+2 array are emulated with a
 * private buffer so the caller's arrays are [nparts+1]. */
#include "internal.h"

/* partition_by_row: spmv/1D/CSR-row/support/partition.h:35-60 */
int sp_partition_by_row(uint32_t nrows, uint32_t nparts, uint32_t *row_split) {
    if (nparts == 0 || !row_split) return SP_ERR_ARG;
    if (nparts == 1) { row_split[0] = 0; row_split[1] = nrows; return SP_OK; }
    uint32_t chunks = nrows / nparts, rest = nrows % nparts, curr = 0;
    row_split[0] = 0;
    for (uint32_t i = 0; i < nparts; i++) {
        uint32_t rows = chunks + (i < rest ? 1 : 0);
        curr += rows;
        if (curr > nrows) curr = nrows;
        row_split[i + 1] = curr;
    }
    return SP_OK;
}

/* Greedy cut over `count` items with weights w(i): the shared rule of
 * partition_by_nnz (spmv/1D/CSR-nnz/support/partition.h:35-78) and its
 * twins.  cut_ok(i) adds the tasklet-level granularity condition. */
static void greedy_split(uint32_t count, uint32_t nparts, uint32_t total, const uint32_t *rowptr,
                         uint32_t base, uint32_t gran, uint32_t *split /* [nparts+2] */) {
    uint32_t per = total / nparts, curr = 0, split_cnt = 0;
    split[0] = 0;
    for (uint32_t i = 0; i < count; i++) {
        curr += rowptr[base + i + 1] - rowptr[base + i];
        if (curr >= per && ((i + 1) % gran == 0)) {
            ++split_cnt;
            if (split_cnt <= nparts) split[split_cnt] = i + 1;
            curr = 0;
        }
    }
    if (curr < per && split_cnt <= nparts) split[++split_cnt] = count;
    if (split_cnt > nparts) split[nparts] = count;
    for (uint32_t i = split_cnt + 1; i <= nparts; i++) split[i] = count;
}

/* partition_by_nnz: spmv/1D/CSR-nnz/support/partition.h:35-78 */
int sp_partition_by_nnz(const uint32_t *rowptr, uint32_t nrows, uint32_t nparts, uint32_t *row_split) {
    if (nparts == 0 || !row_split || !rowptr) return SP_ERR_ARG;
    if (nparts == 1) { row_split[0] = 0; row_split[1] = nrows; return SP_OK; }
    uint32_t *tmp = (uint32_t *)malloc(((size_t)nparts + 2) * sizeof(uint32_t));
    if (!tmp) return SP_ERR_NOMEM;
    greedy_split(nrows, nparts, rowptr[nrows] - rowptr[0], rowptr, 0, 1, tmp);
    memcpy(row_split, tmp, ((size_t)nparts + 1) * sizeof(uint32_t));
    free(tmp);
    return SP_OK;
}

/* partition_tsklt_by_row: spmv/1D/CSR-row/support/partition.h:66-88 */
int sp_partition_tsklt_by_row(uint32_t rows, uint32_t T, uint32_t gran, uint32_t *split) {
    if (T == 0 || gran == 0 || !split) return SP_ERR_ARG;
    uint64_t gt = (uint64_t)gran * T;
    uint32_t chunks = (uint32_t)(rows / gt), rest = (uint32_t)(rows % gt), curr = 0;
    split[0] = 0;
    for (uint32_t i = 0; i < T; i++) {
        uint64_t r = (uint64_t)gran * chunks + (i < rest ? gran : 0);
        uint64_t c = (uint64_t)curr + r;
        curr = c > rows ? rows : (uint32_t)c;
        split[i + 1] = curr;
    }
    return SP_OK;
}

/* partition_tsklt_by_nnz: spmv/1D/CSR-row/support/partition.h:92-139 */
int sp_partition_tsklt_by_nnz(const uint32_t *rowptr, uint32_t row_begin, uint32_t rows, uint32_t nnz,
                              uint32_t T, uint32_t gran, uint32_t *split) {
    if (T == 0 || gran == 0 || !split || !rowptr) return SP_ERR_ARG;
    uint32_t *tmp = (uint32_t *)malloc(((size_t)T + 2) * sizeof(uint32_t));
    if (!tmp) return SP_ERR_NOMEM;
    greedy_split(rows, T, nnz, rowptr, row_begin, gran, tmp);
    memcpy(split, tmp, ((size_t)T + 1) * sizeof(uint32_t));
    free(tmp);
    return SP_OK;
}

/* Exact nnz split with straddling rows: spmv/1D/COO-nnz/host/app.c:118-156 */
int sp_partition_coo_nnz(const uint32_t *rows_nnz, uint32_t nrows, uint32_t nnz, uint32_t P,
                         uint32_t *prev_rows, uint32_t *rows_per_part, uint32_t *prev_nnz_out,
                         uint32_t *nnz_per_part, uint32_t *merge) {
    if (P == 0 || !rows_nnz || !prev_rows || !rows_per_part || !prev_nnz_out || !nnz_per_part || !merge)
        return SP_ERR_ARG;
    uint32_t chunks = nnz / P, rest = nnz % P;
    uint32_t prev_row = 0, prev_nnz = 0;
    for (uint32_t i = 0; i < P; i++) {
        uint32_t per = chunks + (i < rest ? 1 : 0);
        uint32_t first;
        if (rest > 0) first = i >= rest ? rest * (chunks + 1) + (i - rest) * chunks : i * (chunks + 1);
        else first = i * chunks;
        uint32_t r = prev_row, t = prev_nnz;
        uint64_t cur = 0;
        while (cur < per) {
            if (r >= nrows) return SP_ERR_ARG; /* rows_nnz does not sum to nnz */
            cur += rows_nnz[r] - t;
            t = 0;
            r++;
        }
        prev_rows[i] = prev_row;
        rows_per_part[i] = r - prev_row;
        prev_nnz_out[i] = first;
        nnz_per_part[i] = per;
        merge[i] = prev_nnz != 0;
        if (cur == per) {
            prev_row = r;
            prev_nnz = 0;
        } else {
            prev_row = r - 1;
            prev_nnz = rows_nnz[r - 1] - (uint32_t)(cur - per);
        }
    }
    return SP_OK;
}

/* Even split: spmv/1D/COO-nnz/host/app.c:186-205 */
int sp_partition_even(uint32_t count, uint32_t T, uint32_t *start, uint32_t *cnt) {
    if (T == 0 || !start || !cnt) return SP_ERR_ARG;
    uint32_t chunks = count / T, rest = count % T;
    for (uint32_t t = 0; t < T; t++) {
        cnt[t] = chunks + (t < rest ? 1 : 0);
        if (rest > 0) start[t] = t >= rest ? rest * (chunks + 1) + (t - rest) * chunks : t * (chunks + 1);
        else start[t] = t * chunks;
    }
    return SP_OK;
}

/* partition_by_blocks / partition_by_nnz at block-row granularity:
 * spmv/1D/BCSR-block/support/partition.h:39-113 and
 * spmv/1D/BCSR-nnz/support/partition.h:39-113. */
int sp_partition_bcsr(const uint32_t *browptr, const uint32_t *npb, uint32_t nbr, uint64_t nb,
                      uint32_t nnz, uint32_t P, int by_nnz, uint32_t *brow_split,
                      uint32_t *blocks_part, uint32_t *nnzs_part) {
    if (P == 0 || !browptr || !npb || !brow_split || !blocks_part || !nnzs_part) return SP_ERR_ARG;
    if (P == 1) {
        brow_split[0] = 0; brow_split[1] = nbr;
        blocks_part[0] = (uint32_t)nb; nnzs_part[0] = nnz;
        return SP_OK;
    }
    uint32_t *bs = (uint32_t *)calloc((size_t)P + 2, sizeof(uint32_t));
    uint32_t *bp = (uint32_t *)calloc((size_t)P + 2, sizeof(uint32_t));
    uint32_t *np = (uint32_t *)calloc((size_t)P + 2, sizeof(uint32_t));
    if (!bs || !bp || !np) { free(bs); free(bp); free(np); return SP_ERR_NOMEM; }
    uint32_t per = by_nnz ? nnz / P : (uint32_t)(nb / P);
    uint32_t cb = 0, cn = 0, split_cnt = 0;
    bs[0] = 0;
    for (uint32_t i = 0; i < nbr; i++) {
        for (uint32_t j = browptr[i]; j < browptr[i + 1]; j++) { cn += npb[j]; cb++; }
        if ((by_nnz ? cn : cb) >= per) {
            ++split_cnt;
            if (split_cnt <= P) {
                bs[split_cnt] = i + 1;
                bp[split_cnt - 1] = cb;
                np[split_cnt - 1] = cn;
                cb = 0;
                cn = 0;
            }
        }
    }
    if ((by_nnz ? cn : cb) < per && split_cnt <= P) {
        bs[++split_cnt] = nbr;
        bp[split_cnt - 1] = cb;
        np[split_cnt - 1] = cn;
    }
    if (split_cnt > P) {
        bs[P] = nbr;
        bp[P - 1] += cb;
        np[P - 1] += cn;
    }
    for (uint32_t i = split_cnt + 1; i <= P; i++) { bs[i] = nbr; bp[i - 1] = 0; np[i - 1] = 0; }
    memcpy(brow_split, bs, ((size_t)P + 1) * sizeof(uint32_t));
    memcpy(blocks_part, bp, (size_t)P * sizeof(uint32_t));
    memcpy(nnzs_part, np, (size_t)P * sizeof(uint32_t));
    free(bs); free(bp); free(np);
    return SP_OK;
}

/* partition_tsklt_by_block / partition_tsklt_by_nnz:
 * spmv/1D/BCSR-block/support/partition.h:119-188 and :192-262. */
int sp_partition_tsklt_bcsr(const uint32_t *browptr, const uint32_t *npb, const uint32_t *brow_split,
                            const uint32_t *blocks_part, const uint32_t *nnzs_part, uint32_t part,
                            uint32_t T, int by_nnz, uint32_t *split) {
    if (T == 0 || !browptr || !npb || !brow_split || !split) return SP_ERR_ARG;
    uint32_t s = brow_split[part], e = brow_split[part + 1];
    if (T == 1) { split[0] = s; split[1] = e; return SP_OK; }
    uint32_t *tmp = (uint32_t *)malloc(((size_t)T + 2) * sizeof(uint32_t));
    if (!tmp) return SP_ERR_NOMEM;
    uint32_t per = by_nnz ? nnzs_part[part] / T : blocks_part[part] / T;
    uint32_t cb = 0, cn = 0, split_cnt = 0;
    tmp[0] = s;
    for (uint32_t i = s; i < e; i++) {
        for (uint32_t j = browptr[i]; j < browptr[i + 1]; j++) { cn += npb[j]; cb++; }
        if ((by_nnz ? cn : cb) >= per) {
            ++split_cnt;
            if (split_cnt <= T) { tmp[split_cnt] = i + 1; cb = 0; cn = 0; }
        }
    }
    if ((by_nnz ? cn : cb) < per && split_cnt <= T) tmp[++split_cnt] = e;
    if (split_cnt > T) tmp[T] = e;
    for (uint32_t i = split_cnt + 1; i <= T; i++) tmp[i] = e;
    memcpy(split, tmp, ((size_t)T + 1) * sizeof(uint32_t));
    free(tmp);
    return SP_OK;
}

/* BCOO: spmv/1D/BCOO-block/support/partition.h:37-74 (even blocks) and
 * spmv/1D/BCOO-nnz/support/partition.h:37-97 (greedy nnz). */
int sp_partition_bcoo(const uint32_t *npb, uint64_t nb, uint32_t nnz, uint32_t P, int by_nnz,
                      uint32_t *block_split, uint32_t *nnzs_part) {
    if (P == 0 || !npb || !block_split || !nnzs_part) return SP_ERR_ARG;
    if (P == 1) { block_split[0] = 0; block_split[1] = (uint32_t)nb; nnzs_part[0] = nnz; return SP_OK; }
    if (!by_nnz) {
        uint32_t per = (uint32_t)(nb / P), rest = (uint32_t)(nb % P);
        block_split[0] = 0;
        for (uint32_t i = 0; i < P; i++) {
            block_split[i + 1] = block_split[i] + per + (i < rest ? 1 : 0);
            uint32_t cn = 0;
            for (uint32_t j = block_split[i]; j < block_split[i + 1]; j++) cn += npb[j];
            nnzs_part[i] = cn;
        }
        return SP_OK;
    }
    uint32_t *bs = (uint32_t *)calloc((size_t)P + 2, sizeof(uint32_t));
    uint32_t *np = (uint32_t *)calloc((size_t)P + 2, sizeof(uint32_t));
    if (!bs || !np) { free(bs); free(np); return SP_ERR_NOMEM; }
    uint32_t per = nnz / P, cn = 0, split_cnt = 0;
    for (uint64_t i = 0; i < nb; i++) {
        cn += npb[i];
        if (cn >= per) {
            ++split_cnt;
            if (split_cnt <= P) { bs[split_cnt] = (uint32_t)(i + 1); np[split_cnt - 1] = cn; cn = 0; }
        }
    }
    if (cn < per && split_cnt <= P) { bs[++split_cnt] = (uint32_t)nb; np[split_cnt - 1] = cn; }
    if (split_cnt > P) { bs[P] = (uint32_t)nb; np[P - 1] += cn; }
    for (uint32_t i = split_cnt + 1; i <= P; i++) { bs[i] = (uint32_t)nb; np[i - 1] = 0; }
    memcpy(block_split, bs, ((size_t)P + 1) * sizeof(uint32_t));
    memcpy(nnzs_part, np, (size_t)P * sizeof(uint32_t));
    free(bs); free(np);
    return SP_OK;
}

/* spmv/1D/BCOO-block/support/partition.h:80-160 */
int sp_partition_tsklt_bcoo(const uint32_t *npb, const uint32_t *block_split, const uint32_t *nnzs_part,
                            uint32_t part, uint32_t T, int by_nnz, uint32_t *split) {
    if (T == 0 || !npb || !block_split || !split) return SP_ERR_ARG;
    uint32_t s = block_split[part], e = block_split[part + 1];
    if (T == 1) { split[0] = s; split[1] = e; return SP_OK; }
    if (!by_nnz) {
        uint32_t cnt = e - s, per = cnt / T, rest = cnt % T;
        split[0] = s;
        for (uint32_t i = 0; i < T; i++) split[i + 1] = split[i] + per + (i < rest ? 1 : 0);
        return SP_OK;
    }
    uint32_t *tmp = (uint32_t *)malloc(((size_t)T + 2) * sizeof(uint32_t));
    if (!tmp) return SP_ERR_NOMEM;
    uint32_t per = nnzs_part[part] / T, cn = 0, split_cnt = 0;
    tmp[0] = s;
    for (uint32_t i = s; i < e; i++) {
        cn += npb[i];
        if (cn >= per) {
            ++split_cnt;
            if (split_cnt <= T) { tmp[split_cnt] = i + 1; cn = 0; }
        }
    }
    if (cn < per && split_cnt <= T) tmp[++split_cnt] = e;
    if (split_cnt > T) tmp[T] = e;
    for (uint32_t i = split_cnt + 1; i <= T; i++) tmp[i] = e;
    memcpy(split, tmp, ((size_t)T + 1) * sizeof(uint32_t));
    free(tmp);
    return SP_OK;
}

/* find_partitions: spmv/2D/DCSR/host/app.c:66-69 */
uint32_t sp_find_partitions(uint32_t nparts, uint32_t vert) { return vert ? nparts / vert : 0; }
