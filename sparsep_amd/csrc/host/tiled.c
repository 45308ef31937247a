/* 2D tilings of SparseP: equal tiles (D*), equal-width stripes (RBD*) and
 * nnz-balanced stripes (BD*), each with CSR / COO / BCSR / BCOO tiles.
 *
 *   D  tiles : coo2dcsr   spmv/2D/DCSR/support/matrix.h:159-230
 *              coo2dcoo   spmv/2D/DCOO/support/matrix.h:158-217
 *              dcsr2dbcsr spmv/2D/DBCSR/support/matrix.h:278 (+ dbcsr2dbcoo, DBCOO:677)
 *   RB tiles : coo2rbdcsr spmv/2D/RBDCSR/support/matrix.h:159-223
 *              coo2rbdcoo spmv/2D/RBDCOO/support/matrix.h:157-216
 *              rbdcsr2rbdbcsr spmv/2D/RBDBCSR-block/support/matrix.h:270
 *   BD tiles : coo2bdcsr  spmv/2D/BDCSR/support/matrix.h:161-282 (cut points :183-229)
 *              coo2bdcoo  spmv/2D/BDCOO/support/matrix.h:158
 *              bdcsr2bdbcsr spmv/2D/BDBCSR-block/support/matrix.h:334
 * Entries keep the sorted-COO order inside a tile (the reference buckets the
 * sorted COO stably); block tiles are built per tile with the same
 * csr2bcsr/sort/count steps as the 1D formats.
 */
#include "internal.h"

void sp_free_tiled(sp_tiled *A) {
    if (!A) return;
    free(A->vert_tile_widths); free(A->nnzs_per_partition); free(A->nnz_offset);
    free(A->drowptr); free(A->drowind); free(A->dcolind); free(A->dval);
    free(A->blocks_per_partition); free(A->block_offset); free(A->browptr); free(A->browind);
    free(A->bcolind); free(A->bval); free(A->nnz_per_block);
    free(A);
}

/* Column cut points of BD tilings: greedy `curr >= nnz/vert` over the
 * per-column counts (spmv/2D/BDCSR/support/matrix.h:183-229).  The
 * reference's one-past write when the last split lands on `vert` is
 * emulated in a private buffer. */
static int bd_cuts(const sp_coo *A, uint32_t vert, uint32_t *cuts) {
    uint32_t *cnt = (uint32_t *)calloc((size_t)A->ncols + 1, sizeof(uint32_t));
    uint32_t *w = (uint32_t *)calloc((size_t)vert + 2, sizeof(uint32_t));
    if (!cnt || !w) { free(cnt); free(w); return SP_ERR_NOMEM; }
    for (uint32_t k = 0; k < A->nnz; k++) cnt[A->colind[k]]++;
    uint32_t per = A->nnz / vert, curr = 0, split_cnt = 0;
    w[0] = 0;
    for (uint32_t t = 0; t < A->ncols; t++) {
        curr += cnt[t];
        if (curr >= per) {
            ++split_cnt;
            if (split_cnt <= vert) w[split_cnt] = t + 1;
            curr = 0;
        }
    }
    if (curr < per && split_cnt <= vert) w[++split_cnt] = A->ncols;
    if (split_cnt > vert) w[vert] = A->ncols;
    for (uint32_t t = split_cnt + 1; t <= vert; t++) w[t] = A->ncols;
    memcpy(cuts, w, ((size_t)vert + 1) * sizeof(uint32_t));
    free(cnt);
    free(w);
    return SP_OK;
}

static uint32_t stripe_of(const uint32_t *cuts, uint32_t vert, uint32_t col) {
    uint32_t lo = 0, hi = vert; /* first c with cuts[c+1] > col */
    while (lo < hi) {
        uint32_t mid = (lo + hi) / 2;
        if (cuts[mid + 1] > col) hi = mid;
        else lo = mid + 1;
    }
    return lo;
}

static int is_sorted(const sp_coo *A) {
    for (uint32_t k = 1; k < A->nnz; k++) {
        if (A->rowind[k] < A->rowind[k - 1]) return 0;
        if (A->rowind[k] == A->rowind[k - 1] && A->colind[k] < A->colind[k - 1]) return 0;
    }
    return 1;
}

int sp_coo2tiled(const sp_coo *A, int tiling, int format, uint32_t horz, uint32_t vert, uint32_t R,
                 uint32_t C, int count_values, sp_tiled **out) {
    if (!A || !out || vert == 0 || horz == 0 || !sp_valid_dtype(A->dtype)) return SP_ERR_ARG;
    if (tiling < SP_TILE_D || tiling > SP_TILE_BD || format < SP_FMT_CSR || format > SP_FMT_BCOO) return SP_ERR_ARG;
    if ((format >= SP_FMT_BCSR) && (R == 0 || C == 0)) return SP_ERR_ARG;
    if (!is_sorted(A)) return SP_ERR_ARG; /* every 2D app.c sorts before tiling */
    for (uint32_t k = 0; k < A->nnz; k++)
        if (A->rowind[k] >= A->nrows || A->colind[k] >= A->ncols) return SP_ERR_ARG;
    const size_t sv = sp_dtype_size(A->dtype);
    const uint32_t n = A->nrows, m = A->ncols;
    sp_tiled *T = (sp_tiled *)calloc(1, sizeof(sp_tiled));
    if (!T) return SP_ERR_NOMEM;
    T->tiling = tiling; T->format = format; T->dtype = A->dtype;
    T->nrows = n; T->ncols = m; T->nnz = A->nnz;
    T->horz_partitions = horz; T->vert_partitions = vert;
    T->ntiles = tiling == SP_TILE_D ? horz * vert : vert;
    T->tile_height = tiling == SP_TILE_D ? (n + horz - 1) / horz : n;
    if (T->tile_height == 0) T->tile_height = 1;
    T->vert_tile_widths = (uint32_t *)calloc((size_t)vert + 2, sizeof(uint32_t));
    T->nnzs_per_partition = (uint32_t *)calloc((size_t)T->ntiles + 1, sizeof(uint32_t));
    T->nnz_offset = (uint64_t *)calloc((size_t)T->ntiles + 1, sizeof(uint64_t));
    int rc = SP_OK;
    if (!T->vert_tile_widths || !T->nnzs_per_partition || !T->nnz_offset) { rc = SP_ERR_NOMEM; goto fail; }
    if (tiling == SP_TILE_BD) {
        if ((rc = bd_cuts(A, vert, T->vert_tile_widths))) goto fail;
        T->tile_width = 0;
        for (uint32_t c = 0; c < vert; c++) {
            uint32_t w = T->vert_tile_widths[c + 1] - T->vert_tile_widths[c];
            if (w > T->tile_width) T->tile_width = w;
        }
    } else {
        T->tile_width = (m + vert - 1) / vert;
        if (T->tile_width == 0) T->tile_width = 1;
        for (uint32_t c = 0; c <= vert; c++) {
            uint64_t cut = (uint64_t)c * T->tile_width;
            T->vert_tile_widths[c] = (uint32_t)(cut < m ? cut : m);
        }
    }
    const uint32_t th = T->tile_height, nt = T->ntiles;
    /* tile of every entry, counts, offsets */
    uint32_t *tile = (uint32_t *)malloc((size_t)A->nnz * sizeof(uint32_t) + 4);
    if (!tile) { rc = SP_ERR_NOMEM; goto fail; }
    for (uint32_t k = 0; k < A->nnz; k++) {
        uint32_t c = tiling == SP_TILE_BD ? stripe_of(T->vert_tile_widths, vert, A->colind[k])
                                          : A->colind[k] / T->tile_width;
        uint32_t t = tiling == SP_TILE_D ? (A->rowind[k] / th) * vert + c : c;
        tile[k] = t;
        T->nnzs_per_partition[t]++;
    }
    for (uint32_t t = 0; t < nt; t++) T->nnz_offset[t + 1] = T->nnz_offset[t] + T->nnzs_per_partition[t];
    /* bucket entries (stable) into tile-local rows / columns */
    uint32_t *lrow = (uint32_t *)malloc((size_t)A->nnz * sizeof(uint32_t) + 4);
    uint32_t *lcol = (uint32_t *)malloc((size_t)A->nnz * sizeof(uint32_t) + 4);
    char *lval = (char *)malloc((size_t)A->nnz * sv + 8);
    uint64_t *next = (uint64_t *)malloc(((size_t)nt + 1) * sizeof(uint64_t));
    if (!lrow || !lcol || !lval || !next) {
        free(tile); free(lrow); free(lcol); free(lval); free(next);
        rc = SP_ERR_NOMEM;
        goto fail;
    }
    memcpy(next, T->nnz_offset, ((size_t)nt + 1) * sizeof(uint64_t));
    for (uint32_t k = 0; k < A->nnz; k++) {
        uint32_t t = tile[k], c = t % vert;
        uint64_t dst = next[t]++;
        lrow[dst] = tiling == SP_TILE_D ? A->rowind[k] - (t / vert) * th : A->rowind[k];
        lcol[dst] = A->colind[k] - T->vert_tile_widths[c];
        memcpy(lval + dst * sv, (const char *)A->values + (size_t)k * sv, sv);
    }
    free(tile);
    free(next);
    /* per-tile row pointers (stride th+1) */
    uint32_t *rptr = (uint32_t *)calloc((size_t)nt * (th + 1) + 1, sizeof(uint32_t));
    if (!rptr) { free(lrow); free(lcol); free(lval); rc = SP_ERR_NOMEM; goto fail; }
    for (uint32_t t = 0; t < nt; t++) {
        uint32_t *rp = rptr + (size_t)t * (th + 1);
        for (uint64_t k = T->nnz_offset[t]; k < T->nnz_offset[t + 1]; k++) rp[lrow[k] + 1]++;
        for (uint32_t i = 0; i < th; i++) rp[i + 1] += rp[i];
    }
    if (format == SP_FMT_CSR || format == SP_FMT_COO) {
        T->dcolind = lcol;
        T->dval = lval;
        if (format == SP_FMT_CSR) { T->drowptr = rptr; free(lrow); }
        else { T->drowind = lrow; free(rptr); }
        *out = T;
        return SP_OK;
    }
    /* block tiles */
    T->row_block_size = R; T->col_block_size = C;
    T->num_block_rows = (th + R - 1) / R;
    T->num_block_cols = (T->tile_width + C - 1) / C;
    T->blocks_per_partition = (uint32_t *)calloc((size_t)nt + 1, sizeof(uint32_t));
    T->block_offset = (uint64_t *)calloc((size_t)nt + 1, sizeof(uint64_t));
    T->browptr = (uint32_t *)calloc((size_t)nt * (T->num_block_rows + 1) + 1, sizeof(uint32_t));
    sp_bcsr **tb = (sp_bcsr **)calloc((size_t)nt + 1, sizeof(sp_bcsr *));
    if (!T->blocks_per_partition || !T->block_offset || !T->browptr || !tb) rc = SP_ERR_NOMEM;
    for (uint32_t t = 0; t < nt && !rc; t++) {
        const uint32_t c = t % vert;
        sp_csr view;
        view.nrows = th;
        view.ncols = T->vert_tile_widths[c + 1] - T->vert_tile_widths[c];
        view.nnz = T->nnzs_per_partition[t];
        view.dtype = A->dtype;
        view.rowptr = rptr + (size_t)t * (th + 1);
        view.colind = lcol + T->nnz_offset[t];
        view.values = lval + T->nnz_offset[t] * sv;
        if (view.ncols == 0) view.ncols = 1; /* empty stripe */
        rc = sp_build_bcsr(&view, R, C, count_values, &tb[t]);
        if (!rc) {
            T->blocks_per_partition[t] = (uint32_t)tb[t]->num_blocks;
            T->block_offset[t + 1] = T->block_offset[t] + tb[t]->num_blocks;
        }
    }
    if (!rc) {
        const uint64_t nb = T->block_offset[nt];
        const size_t bs = (size_t)R * C * sv;
        T->num_blocks = nb;
        T->bcolind = (uint32_t *)malloc(nb * sizeof(uint32_t) + 4);
        T->bval = malloc(nb * bs + 8);
        T->nnz_per_block = (uint32_t *)malloc(nb * sizeof(uint32_t) + 4);
        if (format == SP_FMT_BCOO) T->browind = (uint32_t *)malloc(nb * sizeof(uint32_t) + 4);
        if (!T->bcolind || !T->bval || !T->nnz_per_block || (format == SP_FMT_BCOO && !T->browind)) rc = SP_ERR_NOMEM;
        for (uint32_t t = 0; t < nt && !rc; t++) {
            const sp_bcsr *B = tb[t];
            const uint64_t o = T->block_offset[t];
            memcpy(T->browptr + (size_t)t * (T->num_block_rows + 1), B->browptr,
                   ((size_t)B->num_block_rows + 1) * sizeof(uint32_t));
            memcpy(T->bcolind + o, B->bcolind, B->num_blocks * sizeof(uint32_t));
            memcpy((char *)T->bval + o * bs, B->bval, B->num_blocks * bs);
            memcpy(T->nnz_per_block + o, B->nnz_per_block, B->num_blocks * sizeof(uint32_t));
            if (format == SP_FMT_BCOO)
                for (uint32_t I = 0; I < B->num_block_rows; I++)
                    for (uint32_t k = B->browptr[I]; k < B->browptr[I + 1]; k++) T->browind[o + k] = I;
        }
    }
    if (tb) {
        for (uint32_t t = 0; t < nt; t++) sp_free_bcsr(tb[t]);
        free(tb);
    }
    free(rptr); free(lrow); free(lcol); free(lval);
    if (rc) goto fail;
    *out = T;
    return SP_OK;
fail:
    sp_free_tiled(T);
    return rc;
}

/* Per-stripe splits of RB/BD tilings into `horz` parts (ex DPUs per
 * vertical partition).  split[c*(horz+1) + i]:
 *   CSR : row splits, greedy by entries
 *         (partition_by_nnz, spmv/2D/RBDCSR/support/partition.h:45-89)
 *   BCSR: block-row splits, greedy by blocks (by_nnz = 0,
 *         spmv/2D/RBDBCSR-block/support/partition.h:50-123) or by non-zeros
 *         (spmv/2D/RBDBCSR-nnz/support/partition.h:51)
 *   COO : entry splits (nnz_split, spmv/2D/RBDCOO/support/partition.h:45-118)
 *   BCOO: block splits, greedy by blocks or by non-zeros
 *         (spmv/2D/RBDBCOO-block/support/partition.h:50-140,
 *          spmv/2D/RBDBCOO-nnz/support/partition.h:50)
 * Values are stripe-local.  For D tilings every tile is its own part and
 * the split is the tile's full range (split[t*(horz+1)..] is not used). */
int sp_partition_tiled(const sp_tiled *A, uint32_t horz, int by_nnz, uint32_t *split) {
    if (!A || !split || horz == 0) return SP_ERR_ARG;
    if (A->tiling == SP_TILE_D) return SP_ERR_ARG;
    const uint32_t vert = A->vert_partitions;
    uint32_t *tmp = (uint32_t *)calloc((size_t)horz + 2, sizeof(uint32_t));
    if (!tmp) return SP_ERR_NOMEM;
    for (uint32_t c = 0; c < vert; c++) {
        uint32_t count, per, curr = 0, split_cnt = 0;
        memset(tmp, 0, ((size_t)horz + 2) * sizeof(uint32_t));
        if (A->format == SP_FMT_CSR || A->format == SP_FMT_BCSR) {
            const int blk = A->format == SP_FMT_BCSR;
            const uint32_t nr = blk ? A->num_block_rows : A->nrows;
            const uint32_t *rp = blk ? A->browptr + (size_t)c * (A->num_block_rows + 1)
                                     : A->drowptr + (size_t)c * (A->tile_height + 1);
            const uint32_t *npb = blk ? A->nnz_per_block + A->block_offset[c] : NULL;
            uint32_t tot = 0;
            if (blk && by_nnz) for (uint32_t b = 0; b < A->blocks_per_partition[c]; b++) tot += npb[b];
            else tot = rp[nr];
            per = tot / horz;
            for (uint32_t i = 0; i < nr; i++) {
                if (blk && by_nnz) for (uint32_t b = rp[i]; b < rp[i + 1]; b++) curr += npb[b];
                else curr += rp[i + 1] - rp[i];
                if (curr >= per) {
                    ++split_cnt;
                    if (split_cnt <= horz) { tmp[split_cnt] = i + 1; curr = 0; }
                    else if (!blk) curr = 0;
                }
            }
            count = nr;
        } else {
            const int blk = A->format == SP_FMT_BCOO;
            const uint32_t cnt = blk ? A->blocks_per_partition[c] : A->nnzs_per_partition[c];
            const uint32_t *npb = blk ? A->nnz_per_block + A->block_offset[c] : NULL;
            uint32_t tot = cnt;
            if (blk && by_nnz) { tot = 0; for (uint32_t b = 0; b < cnt; b++) tot += npb[b]; }
            per = tot / horz;
            for (uint32_t i = 0; i < cnt; i++) {
                curr += (blk && by_nnz) ? npb[i] : 1;
                if (curr >= per) {
                    ++split_cnt;
                    if (split_cnt <= horz) { tmp[split_cnt] = i + 1; curr = 0; }
                    else if (!blk) curr = 0;
                }
            }
            count = cnt;
        }
        if (curr < per && split_cnt <= horz) tmp[++split_cnt] = count;
        if (split_cnt > horz) tmp[horz] = count;
        for (uint32_t i = split_cnt + 1; i <= horz; i++) tmp[i] = count;
        memcpy(split + (size_t)c * (horz + 1), tmp, ((size_t)horz + 1) * sizeof(uint32_t));
    }
    free(tmp);
    return SP_OK;
}
