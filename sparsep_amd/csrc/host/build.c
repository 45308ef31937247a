/* COO / CSR / BCSR / BCOO builders (reference: support/matrix.h). */
#include <stdio.h>

#include "internal.h"

#define ALLOC(p, n, T) ((p) = (T *)malloc((size_t)(n) * sizeof(T) + 1))
#define CALLOC(p, n, T) ((p) = (T *)calloc((size_t)(n) + 1, sizeof(T)))

int sp_coo_alloc(uint32_t nrows, uint32_t ncols, uint32_t nnz, int dtype, sp_coo **out) {
    if (!out || !sp_valid_dtype(dtype)) return SP_ERR_ARG;
    sp_coo *A = (sp_coo *)calloc(1, sizeof(sp_coo));
    if (!A) return SP_ERR_NOMEM;
    A->nrows = nrows; A->ncols = ncols; A->nnz = nnz; A->dtype = dtype;
    ALLOC(A->rowind, nnz, uint32_t);
    ALLOC(A->colind, nnz, uint32_t);
    A->values = malloc((size_t)nnz * sp_dtype_size(dtype) + 8);
    CALLOC(A->rows, nrows, uint32_t);
    if (!A->rowind || !A->colind || !A->values || !A->rows) { sp_free_coo(A); return SP_ERR_NOMEM; }
    *out = A;
    return SP_OK;
}

void sp_free_coo(sp_coo *A) {
    if (!A) return;
    free(A->rowind); free(A->colind); free(A->values); free(A->rows); free(A);
}
void sp_free_csr(sp_csr *A) {
    if (!A) return;
    free(A->rowptr); free(A->colind); free(A->values); free(A);
}
void sp_free_bcsr(sp_bcsr *A) {
    if (!A) return;
    free(A->browptr); free(A->bcolind); free(A->bval); free(A->nnz_per_block); free(A);
}
void sp_free_bcoo(sp_bcoo *A) {
    if (!A) return;
    free(A->browind); free(A->bcolind); free(A->bval); free(A->nnz_per_block); free(A);
}

int sp_coo_from_keys(uint32_t nrows, uint32_t ncols, const uint64_t *keys, uint64_t nnz,
                     int dtype, sp_dist vals, uint64_t vseed, sp_coo **out) {
    if (nnz > 0xFFFFFFFFULL) return SP_ERR_UNSUPPORTED;
    sp_coo *A;
    int rc = sp_coo_alloc(nrows, ncols, (uint32_t)nnz, dtype, &A);
    if (rc) return rc;
    sp_rng r;
    sp_rng_seed(&r, vseed);
    for (uint64_t k = 0; k < nnz; k++) {
        uint32_t row = (uint32_t)(keys[k] >> 32), col = (uint32_t)keys[k];
        A->rowind[k] = row;
        A->colind[k] = col;
        A->rows[row]++;
        sp_put(A->values, dtype, k, sp_draw(&r, dtype, vals));
    }
    *out = A;
    return SP_OK;
}

/* Stable sort of the entries of A by (row, col): counting sort by row
 * (stable) then a stable merge sort by column inside each row. */
int sp_sort_coo_entries(sp_coo *A) {
    if (!A) return SP_ERR_ARG;
    const uint32_t n = A->nrows, nnz = A->nnz;
    const size_t sv = sp_dtype_size(A->dtype);
    for (uint32_t k = 0; k < nnz; k++)
        if (A->rowind[k] >= n || A->colind[k] >= A->ncols) return SP_ERR_ARG;
    uint32_t *ptr = (uint32_t *)calloc((size_t)n + 2, sizeof(uint32_t));
    uint32_t *perm = (uint32_t *)malloc((size_t)nnz * sizeof(uint32_t) + 4);
    uint32_t *tmp = (uint32_t *)malloc((size_t)nnz * sizeof(uint32_t) + 4);
    if (!ptr || !perm || !tmp) { free(ptr); free(perm); free(tmp); return SP_ERR_NOMEM; }
    for (uint32_t k = 0; k < nnz; k++) ptr[A->rowind[k] + 1]++;
    for (uint32_t i = 0; i < n; i++) ptr[i + 1] += ptr[i];
    {
        uint32_t *next = (uint32_t *)malloc((size_t)n * sizeof(uint32_t) + 4);
        if (!next) { free(ptr); free(perm); free(tmp); return SP_ERR_NOMEM; }
        memcpy(next, ptr, (size_t)n * sizeof(uint32_t));
        for (uint32_t k = 0; k < nnz; k++) perm[next[A->rowind[k]]++] = k;
        free(next);
    }
    const uint32_t *col = A->colind;
    for (uint32_t i = 0; i < n; i++) {
        uint32_t b = ptr[i], e = ptr[i + 1], len = e - b;
        if (len < 2) continue;
        if (len <= 32) { /* insertion sort (stable) */
            for (uint32_t a = b + 1; a < e; a++) {
                uint32_t v = perm[a], j = a;
                while (j > b && col[perm[j - 1]] > col[v]) { perm[j] = perm[j - 1]; j--; }
                perm[j] = v;
            }
            continue;
        }
        /* bottom-up merge sort (stable) */
        uint32_t *src = perm + b, *dst = tmp + b;
        for (uint32_t w = 1; w < len; w *= 2) {
            for (uint32_t lo = 0; lo < len; lo += 2 * w) {
                uint32_t mid = lo + w < len ? lo + w : len;
                uint32_t hi = lo + 2 * w < len ? lo + 2 * w : len;
                uint32_t p = lo, q = mid, o = lo;
                while (p < mid && q < hi) dst[o++] = (col[src[q]] < col[src[p]]) ? src[q++] : src[p++];
                while (p < mid) dst[o++] = src[p++];
                while (q < hi) dst[o++] = src[q++];
            }
            uint32_t *t = src; src = dst; dst = t;
        }
        if (src != perm + b) memcpy(perm + b, src, (size_t)len * sizeof(uint32_t));
    }
    /* apply the permutation */
    uint32_t *nr = tmp;
    for (uint32_t k = 0; k < nnz; k++) nr[k] = A->colind[perm[k]];
    memcpy(A->colind, nr, (size_t)nnz * sizeof(uint32_t));
    char *nv = (char *)malloc((size_t)nnz * sv + 8);
    if (!nv) { free(ptr); free(perm); free(tmp); return SP_ERR_NOMEM; }
    for (uint32_t k = 0; k < nnz; k++) memcpy(nv + (size_t)k * sv, (char *)A->values + (size_t)perm[k] * sv, sv);
    free(A->values);
    A->values = nv;
    for (uint32_t i = 0; i < n; i++)
        for (uint32_t k = ptr[i]; k < ptr[i + 1]; k++) A->rowind[k] = i;
    if (!A->rows) A->rows = (uint32_t *)calloc((size_t)n + 1, sizeof(uint32_t));
    if (!A->rows) { free(ptr); free(perm); free(tmp); return SP_ERR_NOMEM; }
    for (uint32_t i = 0; i < n; i++) A->rows[i] = ptr[i + 1] - ptr[i];
    free(ptr); free(perm); free(tmp);
    return SP_OK;
}

/* Reference: sortCOOMatrix (spmv/1D/COO-row/support/matrix.h:105-125). */
int sp_sort_coo(sp_coo *A) { return sp_sort_coo_entries(A); }

/* Reference: coo2csr (spmv/1D/CSR-row/support/matrix.h:113-152). */
int sp_coo2csr(const sp_coo *A, sp_csr **out) {
    if (!A || !out || !sp_valid_dtype(A->dtype)) return SP_ERR_ARG;
    const size_t sv = sp_dtype_size(A->dtype);
    sp_csr *C = (sp_csr *)calloc(1, sizeof(sp_csr));
    if (!C) return SP_ERR_NOMEM;
    C->nrows = A->nrows; C->ncols = A->ncols; C->nnz = A->nnz; C->dtype = A->dtype;
    CALLOC(C->rowptr, (size_t)A->nrows + 1, uint32_t);
    ALLOC(C->colind, A->nnz, uint32_t);
    C->values = calloc((size_t)A->nnz + 8, sv);
    if (!C->rowptr || !C->colind || !C->values) { sp_free_csr(C); return SP_ERR_NOMEM; }
    for (uint32_t k = 0; k < A->nnz; k++) {
        if (A->rowind[k] >= A->nrows) { sp_free_csr(C); return SP_ERR_ARG; }
        C->rowptr[A->rowind[k] + 1]++;
    }
    for (uint32_t i = 0; i < A->nrows; i++) C->rowptr[i + 1] += C->rowptr[i];
    uint32_t *next = (uint32_t *)malloc((size_t)A->nrows * sizeof(uint32_t) + 4);
    if (!next) { sp_free_csr(C); return SP_ERR_NOMEM; }
    memcpy(next, C->rowptr, (size_t)A->nrows * sizeof(uint32_t));
    for (uint32_t k = 0; k < A->nnz; k++) {
        uint32_t dst = next[A->rowind[k]]++;
        C->colind[dst] = A->colind[k];
        memcpy((char *)C->values + (size_t)dst * sv, (const char *)A->values + (size_t)k * sv, sv);
    }
    free(next);
    *out = C;
    return SP_OK;
}

/* Reference: csr2bcsr (spmv/1D/BCSR-block/support/matrix.h:202-400).  Same
 * two passes (count blocks per block row, then copy in first-seen order),
 * but a block row clears only the block columns it touched, so the cost is
 * O(nnz + num_block_rows) instead of O(num_block_rows * num_block_cols).
 * nnz_per_block receives stored-entry counts; sp_count_nnz_per_block
 * switches it to the reference's value count. */
int sp_csr2bcsr(const sp_csr *A, uint32_t r, uint32_t c, sp_bcsr **out) {
    if (!A || !out || r == 0 || c == 0 || !sp_valid_dtype(A->dtype)) return SP_ERR_ARG;
    const size_t sv = sp_dtype_size(A->dtype);
    const uint32_t nbr = (uint32_t)(((uint64_t)A->nrows + r - 1) / r);
    const uint32_t nbc = (uint32_t)(((uint64_t)A->ncols + c - 1) / c);
    sp_bcsr *B = (sp_bcsr *)calloc(1, sizeof(sp_bcsr));
    if (!B) return SP_ERR_NOMEM;
    B->row_block_size = r; B->col_block_size = c;
    B->num_block_rows = nbr; B->num_block_cols = nbc;
    B->num_rows_left = A->nrows % r;
    B->nrows = A->nrows; B->ncols = A->ncols; B->nnz = A->nnz; B->dtype = A->dtype;
    uint32_t *pos = (uint32_t *)malloc((size_t)nbc * sizeof(uint32_t) + 4);
    uint32_t *touched = (uint32_t *)malloc(((size_t)A->nnz < nbc ? (size_t)A->nnz : nbc) * sizeof(uint32_t) + 4);
    CALLOC(B->browptr, (size_t)nbr + 1, uint32_t);
    if (!pos || !touched || !B->browptr) { free(pos); free(touched); sp_free_bcsr(B); return SP_ERR_NOMEM; }
    memset(pos, 0xFF, (size_t)nbc * sizeof(uint32_t));
    /* Phase I: count distinct block columns per block row. */
    uint64_t nb = 0;
    for (uint32_t I = 0; I < nbr; I++) {
        uint32_t nt = 0;
        uint32_t i0 = I * r, i1 = (uint64_t)i0 + r < A->nrows ? i0 + r : A->nrows;
        for (uint32_t i = i0; i < i1; i++)
            for (uint32_t k = A->rowptr[i]; k < A->rowptr[i + 1]; k++) {
                uint32_t J = A->colind[k] / c;
                if (pos[J] == 0xFFFFFFFFu) { pos[J] = 0; touched[nt++] = J; }
            }
        for (uint32_t t = 0; t < nt; t++) pos[touched[t]] = 0xFFFFFFFFu;
        nb += nt;
        if (nb > 0xFFFFFFFFULL) { free(pos); free(touched); sp_free_bcsr(B); return SP_ERR_UNSUPPORTED; }
        B->browptr[I + 1] = (uint32_t)nb;
    }
    B->num_blocks = nb;
    const size_t rc = (size_t)r * c;
    ALLOC(B->bcolind, nb, uint32_t);
    B->bval = calloc((size_t)(nb + 1) * rc, sv);
    CALLOC(B->nnz_per_block, nb, uint32_t);
    if (!B->bcolind || !B->bval || !B->nnz_per_block) { free(pos); free(touched); sp_free_bcsr(B); return SP_ERR_NOMEM; }
    /* Phase II: copy blocks in first-seen order. */
    for (uint32_t I = 0; I < nbr; I++) {
        uint32_t nt = 0, base = B->browptr[I];
        uint32_t i0 = I * r, i1 = (uint64_t)i0 + r < A->nrows ? i0 + r : A->nrows;
        for (uint32_t i = i0; i < i1; i++)
            for (uint32_t k = A->rowptr[i]; k < A->rowptr[i + 1]; k++) {
                uint32_t j = A->colind[k], J = j / c;
                if (pos[J] == 0xFFFFFFFFu) { pos[J] = base + nt; B->bcolind[base + nt] = J; touched[nt++] = J; }
                uint64_t slot = (uint64_t)pos[J] * rc + (uint64_t)(i - i0) * c + (j - J * c);
                memcpy((char *)B->bval + slot * sv, (const char *)A->values + (size_t)k * sv, sv);
                B->nnz_per_block[pos[J]]++;
            }
        for (uint32_t t = 0; t < nt; t++) pos[touched[t]] = 0xFFFFFFFFu;
    }
    free(pos); free(touched);
    *out = B;
    return SP_OK;
}

/* Reference: sortBCSRMatrix (spmv/1D/BCSR-block/support/matrix.h:460-469).
 * Block columns are unique in a block row, so the result equals the
 * reference quicksort's. */
int sp_sort_bcsr(sp_bcsr *B) {
    if (!B) return SP_ERR_ARG;
    const size_t sv = sp_dtype_size(B->dtype), bs = (size_t)B->row_block_size * B->col_block_size * sv;
    uint64_t maxlen = 0;
    for (uint32_t I = 0; I < B->num_block_rows; I++) {
        uint64_t len = B->browptr[I + 1] - B->browptr[I];
        if (len > maxlen) maxlen = len;
    }
    uint64_t *key = (uint64_t *)malloc(maxlen * sizeof(uint64_t) + 8);
    uint32_t *ci = (uint32_t *)malloc(maxlen * sizeof(uint32_t) + 4);
    uint32_t *cn = (uint32_t *)malloc(maxlen * sizeof(uint32_t) + 4);
    char *cv = (char *)malloc(maxlen * bs + 8);
    if (!key || !ci || !cn || !cv) { free(key); free(ci); free(cn); free(cv); return SP_ERR_NOMEM; }
    for (uint32_t I = 0; I < B->num_block_rows; I++) {
        uint32_t b = B->browptr[I], e = B->browptr[I + 1], len = e - b;
        int sorted = 1;
        for (uint32_t k = b + 1; k < e; k++) if (B->bcolind[k - 1] > B->bcolind[k]) { sorted = 0; break; }
        if (sorted) continue;
        for (uint32_t k = 0; k < len; k++) key[k] = ((uint64_t)B->bcolind[b + k] << 32) | k;
        sp_radix_sort_u64(key, len);
        for (uint32_t k = 0; k < len; k++) {
            uint32_t s = b + (uint32_t)(key[k] & 0xFFFFFFFFu);
            ci[k] = B->bcolind[s];
            cn[k] = B->nnz_per_block[s];
            memcpy(cv + (size_t)k * bs, (char *)B->bval + (size_t)s * bs, bs);
        }
        memcpy(B->bcolind + b, ci, (size_t)len * sizeof(uint32_t));
        memcpy(B->nnz_per_block + b, cn, (size_t)len * sizeof(uint32_t));
        memcpy((char *)B->bval + (size_t)b * bs, cv, (size_t)len * bs);
    }
    free(key); free(ci); free(cn); free(cv);
    return SP_OK;
}

/* Reference: countNNZperBlockBCSRMatrix (spmv/1D/BCSR-block/support/matrix.h:474-485). */
int sp_count_nnz_per_block(sp_bcsr *B, const sp_csr *A, int count_values) {
    if (!B) return SP_ERR_ARG;
    const size_t rc = (size_t)B->row_block_size * B->col_block_size;
    if (count_values) {
        for (uint64_t i = 0; i < B->num_blocks; i++) {
            uint32_t cnt = 0;
            for (size_t q = 0; q < rc; q++) cnt += (uint32_t)sp_nonzero(B->bval, B->dtype, i * rc + q);
            B->nnz_per_block[i] = cnt;
        }
        return SP_OK;
    }
    if (!A || A->nrows != B->nrows) return SP_ERR_ARG;
    /* stored entries: map each entry to its block via a per-block-row table */
    uint32_t *pos = (uint32_t *)malloc((size_t)B->num_block_cols * sizeof(uint32_t) + 4);
    if (!pos) return SP_ERR_NOMEM;
    memset(B->nnz_per_block, 0, B->num_blocks * sizeof(uint32_t));
    for (uint32_t I = 0; I < B->num_block_rows; I++) {
        for (uint32_t k = B->browptr[I]; k < B->browptr[I + 1]; k++) pos[B->bcolind[k]] = k;
        uint32_t i0 = I * B->row_block_size;
        uint32_t i1 = i0 + B->row_block_size < B->nrows ? i0 + B->row_block_size : B->nrows;
        for (uint32_t i = i0; i < i1; i++)
            for (uint32_t k = A->rowptr[i]; k < A->rowptr[i + 1]; k++)
                B->nnz_per_block[pos[A->colind[k] / B->col_block_size]]++;
    }
    free(pos);
    return SP_OK;
}

/* Reference: bcsr2bcoo (spmv/1D/BCOO-block/support/matrix.h:549-580). */
int sp_bcsr2bcoo(const sp_bcsr *B, sp_bcoo **out) {
    if (!B || !out) return SP_ERR_ARG;
    const size_t bs = (size_t)B->row_block_size * B->col_block_size * sp_dtype_size(B->dtype);
    sp_bcoo *O = (sp_bcoo *)calloc(1, sizeof(sp_bcoo));
    if (!O) return SP_ERR_NOMEM;
    O->row_block_size = B->row_block_size; O->col_block_size = B->col_block_size;
    O->num_block_rows = B->num_block_rows; O->num_block_cols = B->num_block_cols;
    O->num_blocks = B->num_blocks; O->num_rows_left = B->num_rows_left;
    O->nrows = B->nrows; O->ncols = B->ncols; O->nnz = B->nnz; O->dtype = B->dtype;
    ALLOC(O->browind, B->num_blocks, uint32_t);
    ALLOC(O->bcolind, B->num_blocks, uint32_t);
    O->bval = malloc((size_t)(B->num_blocks + 1) * bs);
    ALLOC(O->nnz_per_block, B->num_blocks, uint32_t);
    if (!O->browind || !O->bcolind || !O->bval || !O->nnz_per_block) { sp_free_bcoo(O); return SP_ERR_NOMEM; }
    for (uint32_t I = 0; I < B->num_block_rows; I++)
        for (uint32_t k = B->browptr[I]; k < B->browptr[I + 1]; k++) O->browind[k] = I;
    memcpy(O->bcolind, B->bcolind, B->num_blocks * sizeof(uint32_t));
    memcpy(O->bval, B->bval, (size_t)(B->num_blocks + 1) * bs);
    memcpy(O->nnz_per_block, B->nnz_per_block, B->num_blocks * sizeof(uint32_t));
    *out = O;
    return SP_OK;
}

int sp_build_bcsr(const sp_csr *A, uint32_t r, uint32_t c, int count_values, sp_bcsr **out) {
    sp_bcsr *B = NULL;
    int rc = sp_csr2bcsr(A, r, c, &B);
    if (rc) return rc;
    rc = sp_sort_bcsr(B);
    if (!rc && count_values) rc = sp_count_nnz_per_block(B, A, 1);
    if (rc) { sp_free_bcsr(B); return rc; }
    *out = B;
    return SP_OK;
}
