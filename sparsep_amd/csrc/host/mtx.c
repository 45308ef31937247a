/* Matrix Market loader (reference: readCOOMatrix,
 * spmv/1D/CSR-row/support/matrix.h:44-96). */
#include <errno.h>
#include <stdio.h>

#include "internal.h"

int sp_read_mtx(const char *path, int dtype, int flags, sp_coo **out) {
    if (!path || !out || !sp_valid_dtype(dtype)) return SP_ERR_ARG;
    FILE *fp = fopen(path, "r");
    if (!fp) return SP_ERR_IO;
    char line[1024];
    int have_header = 0;
    uint64_t nrows = 0, ncols = 0, nnz = 0, i = 0;
    sp_coo *A = NULL;
    int rc = SP_OK;
    while (fgets(line, sizeof line, fp)) {
        char *p = line;
        while (*p == ' ' || *p == '\t') p++;
        if (*p == '%') continue;
        if (*p == '\n' || *p == '\r' || *p == 0) continue;
        if (!have_header) {
            char *end;
            errno = 0;
            nrows = strtoull(p, &end, 10);
            if (end == p) { rc = SP_ERR_IO; break; }
            p = end;
            ncols = strtoull(p, &end, 10);
            if (end == p) { rc = SP_ERR_IO; break; }
            p = end;
            nnz = strtoull(p, &end, 10);
            if (end == p || errno) { rc = SP_ERR_IO; break; }
            if (nrows > 0xFFFFFFF0ULL || ncols > 0xFFFFFFF0ULL || nnz > 0xFFFFFFFFULL) { rc = SP_ERR_UNSUPPORTED; break; }
            uint64_t pr = nrows, pc = ncols;
            if (flags & SP_MTX_PAD) { /* matrix.h:68-73: pad to 8/byte_dt */
                uint64_t g = 8 / sp_dtype_size(dtype);
                if (pr % g) pr += g - pr % g;
                if (pc % g) pc += g - pc % g;
            }
            rc = sp_coo_alloc((uint32_t)pr, (uint32_t)pc, (uint32_t)nnz, dtype, &A);
            if (rc) break;
            have_header = 1;
            continue;
        }
        if (i >= nnz) { rc = SP_ERR_IO; break; }
        char *end;
        uint64_t r = strtoull(p, &end, 10);
        if (end == p) { rc = SP_ERR_IO; break; }
        p = end;
        uint64_t c = strtoull(p, &end, 10);
        if (end == p) { rc = SP_ERR_IO; break; }
        p = end;
        if (r < 1 || r > nrows || c < 1 || c > ncols) { rc = SP_ERR_IO; break; }
        double v = 1.0;
        if (flags & SP_MTX_COMPAT) {
            v = (double)(rand() % 4 + 1); /* matrix.h:83 */
        } else {
            double fv = strtod(p, &end);
            if (end != p) v = fv;
        }
        A->rowind[i] = (uint32_t)(r - 1);
        A->colind[i] = (uint32_t)(c - 1);
        sp_put(A->values, dtype, i, v);
        A->rows[r - 1]++;
        i++;
    }
    fclose(fp);
    if (!rc && !have_header) rc = SP_ERR_IO;
    if (!rc && i != nnz) rc = SP_ERR_IO;
    if (!rc && (flags & SP_MTX_SORT)) rc = sp_sort_coo(A);
    if (rc) { sp_free_coo(A); return rc; }
    *out = A;
    return SP_OK;
}
