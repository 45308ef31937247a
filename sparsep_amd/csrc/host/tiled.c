/* 2D tilings (D*, RBD*, BD*): built in a later step. */
#include "internal.h"

void sp_free_tiled(sp_tiled *A) { (void)A; }
int sp_coo2tiled(const sp_coo *A, int tiling, int format, uint32_t horz, uint32_t vert, uint32_t r,
                 uint32_t c, int count_values, sp_tiled **out) {
    (void)A; (void)tiling; (void)format; (void)horz; (void)vert; (void)r; (void)c; (void)count_values; (void)out;
    return SP_ERR_UNSUPPORTED;
}
int sp_partition_tiled_rows(const sp_tiled *A, uint32_t horz, uint32_t *row_split) {
    (void)A; (void)horz; (void)row_split;
    return SP_ERR_UNSUPPORTED;
}
