/*
 * sparsep.h -- C ABI of the MI355X (gfx950) SpMV library.
 *
 * Two halves:
 *   1. The host surface: Matrix Market loader, seeded generators, the
 *      COO/CSR/BCSR/BCOO builders, the 2D tilings (D*, RBD*, BD*) and the
 *      1D/2D partitioners.  These are plain C11 on the CPU and keep the
 *      semantics of SparseP's support/matrix.h and support/partition.h
 *      (each declaration cites the reference function it replaces).
 *   2. The device path: a context (HIP device + stream), a plan per kernel
 *      variant (the 25 SparseP kernel directories), stream-ordered SpMV/SpMM,
 *      a batched persistent launch, timing and the variant selector.  It
 *      replaces the UPMEM host library calls and every dpu/task.c.
 *
 * Conventions: every function returns 0 (SP_OK) or a negative sp_status;
 * nothing aborts.  The caller owns host matrices and vectors; a plan owns
 * its device copy of the matrix; a context owns its stream.  A context is
 * not thread-safe: use one per host thread.  Values are untyped pointers
 * tagged by sp_dtype; integers use two's-complement wraparound exactly as
 * C `val_dt` arithmetic does under gcc (support/utils.h:13-34).
 */
#ifndef SPARSEP_H
#define SPARSEP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------------ */
/* Status codes and data types                                              */
/* ------------------------------------------------------------------------ */

typedef enum sp_status {
    SP_OK = 0,
    SP_ERR_ARG = -1,          /* invalid argument or shape                 */
    SP_ERR_NOMEM = -2,        /* host or device allocation failed, or the  */
                              /* plan exceeds sp_opts.mem_cap_bytes        */
    SP_ERR_IO = -3,           /* file could not be opened or parsed        */
    SP_ERR_HIP = -4,          /* HIP runtime error                         */
    SP_ERR_UNSUPPORTED = -5,  /* variant/dtype/shape not supported         */
    SP_ERR_NODEVICE = -6,     /* no HIP device                             */
    SP_ERR_STATE = -7,        /* object in the wrong state                 */
    SP_ERR_RANGE = -8         /* input larger than a fixed-capacity entry  */
                              /* point holds                               */
} sp_status;

const char *sp_strerror(int rc);

/* Reference: support/utils.h:13-34 (val_dt / byte_dt per TYPE macro). */
typedef enum sp_dtype {
    SP_I8 = 0, SP_I16 = 1, SP_I32 = 2, SP_I64 = 3, SP_F32 = 4, SP_F64 = 5
} sp_dtype;

size_t sp_dtype_size(int dtype);

/* ------------------------------------------------------------------------ */
/* Matrix structures (reference: support/matrix.h of each kernel directory) */
/* ------------------------------------------------------------------------ */

/* COO, struct-of-arrays.  Reference COOMatrix in two layouts: SoA
 * (spmv/1D/CSR-row/support/matrix.h:19-26) and AoS elem_t with per-row
 * counts `rows` (spmv/1D/COO-row/support/matrix.h:19-25).  Both are kept in
 * one struct: `rows` is filled by sp_sort_coo / the generators. */
typedef struct sp_coo {
    uint32_t nrows, ncols, nnz;
    int32_t dtype;
    uint32_t *rowind;    /* [nnz] */
    uint32_t *colind;    /* [nnz] */
    void *values;        /* [nnz] of dtype */
    uint32_t *rows;      /* [nrows] nnz per row, or NULL */
} sp_coo;

/* Reference: spmv/1D/CSR-row/support/matrix.h:31-38 */
typedef struct sp_csr {
    uint32_t nrows, ncols, nnz;
    int32_t dtype;
    uint32_t *rowptr;    /* [nrows+1] */
    uint32_t *colind;    /* [nnz] */
    void *values;        /* [nnz] */
} sp_csr;

/* Reference: spmv/1D/BCSR-block/support/matrix.h:43-57 */
typedef struct sp_bcsr {
    uint32_t row_block_size, col_block_size;
    uint32_t num_block_rows, num_block_cols;
    uint64_t num_blocks;
    uint32_t num_rows_left;
    uint32_t nrows, ncols, nnz;
    int32_t dtype;
    uint32_t *browptr;       /* [num_block_rows+1] */
    uint32_t *bcolind;       /* [num_blocks] */
    void *bval;              /* [num_blocks * r * c], row-major in a block */
    uint32_t *nnz_per_block; /* [num_blocks] */
} sp_bcsr;

/* Reference: spmv/1D/BCOO-block/support/matrix.h (BCOOMatrix; bind_t at
 * spmv/1D/BCOO-nnz/support/utils.h:39-42).  bind_t is kept as two arrays. */
typedef struct sp_bcoo {
    uint32_t row_block_size, col_block_size;
    uint32_t num_block_rows, num_block_cols;
    uint64_t num_blocks;
    uint32_t num_rows_left;
    uint32_t nrows, ncols, nnz;
    int32_t dtype;
    uint32_t *browind;       /* [num_blocks] bind[i].rowind */
    uint32_t *bcolind;       /* [num_blocks] bind[i].colind */
    void *bval;              /* [num_blocks * r * c] */
    uint32_t *nnz_per_block; /* [num_blocks] */
} sp_bcoo;

/* 2D tilings.  One struct covers the twelve 2D formats:
 *   tiling D  (equal tiles, local rows)  spmv/2D/D{CSR,COO,BCSR,BCOO}
 *   tiling RB (equal-width stripes)      spmv/2D/RBD{CSR,COO,BCSR,BCOO}
 *   tiling BD (nnz-balanced stripes)     spmv/2D/BD{CSR,COO,BCSR,BCOO}
 * Tile t covers rows [row_off(t), row_off(t)+tile_rows) and columns
 * [vert_tile_widths[c], vert_tile_widths[c+1]) with c = t % vert.  For D,
 * tile t = r*vert + c and row_off = r*tile_height; for RB/BD there are
 * `vert` tiles with row_off = 0 and tile_rows = nrows.  Row and column
 * indices inside a tile are local, as in the reference builders. */
typedef enum sp_tiling { SP_TILE_D = 0, SP_TILE_RB = 1, SP_TILE_BD = 2 } sp_tiling;
typedef enum sp_format { SP_FMT_CSR = 0, SP_FMT_COO = 1, SP_FMT_BCSR = 2, SP_FMT_BCOO = 3 } sp_format;

typedef struct sp_tiled {
    int32_t tiling, format, dtype;
    uint32_t nrows, ncols, nnz;
    uint32_t horz_partitions, vert_partitions;
    uint32_t ntiles;             /* D: horz*vert; RB/BD: vert            */
    uint32_t tile_height;        /* rows per tile (D); nrows (RB/BD)     */
    uint32_t tile_width;         /* D/RB: ceil(ncols/vert); BD: max width */
    uint32_t *vert_tile_widths;  /* [vert+1] column cut points           */
    uint32_t *nnzs_per_partition;/* [ntiles] stored entries per tile    */
    uint64_t *nnz_offset;        /* [ntiles+1] entry offsets            */
    /* CSR: rowptr per tile, stride tile_height+1 (reference drowptr).
     * COO: drowind holds local rows.                                   */
    uint32_t *drowptr;           /* CSR: [ntiles*(tile_height+1)]       */
    uint32_t *drowind;           /* COO: [nnz]                          */
    uint32_t *dcolind;           /* CSR/COO: [nnz] local columns        */
    void *dval;                  /* CSR/COO: [nnz]                      */
    /* Block formats (r x c blocks, tile-local block indices).          */
    uint32_t row_block_size, col_block_size;
    uint32_t num_block_rows;     /* per tile: ceil(tile_height / r)     */
    uint32_t num_block_cols;     /* per tile: ceil(tile_width / c)      */
    uint64_t num_blocks;
    uint32_t *blocks_per_partition; /* [ntiles]                        */
    uint64_t *block_offset;      /* [ntiles+1]                          */
    uint32_t *browptr;           /* BCSR: [ntiles*(num_block_rows+1)]   */
    uint32_t *browind;           /* BCOO: [num_blocks]                  */
    uint32_t *bcolind;           /* [num_blocks]                        */
    void *bval;                  /* [num_blocks*r*c]                    */
    uint32_t *nnz_per_block;     /* [num_blocks]                        */
} sp_tiled;

/* ------------------------------------------------------------------------ */
/* Loader, generators, PRNG                                                 */
/* ------------------------------------------------------------------------ */

/* Loader flags.  SP_MTX_COMPAT reproduces readCOOMatrix exactly: values are
 * `rand()%4+1` from the process's glibc rand() (file values ignored), and
 * with SP_MTX_PAD nrows/ncols are padded up to a multiple of 8/sizeof(val)
 * (spmv/1D/CSR-row/support/matrix.h:44-96; the 1D B* directories do not pad,
 * spmv/1D/BCSR-block/support/matrix.h:63-110). */
enum {
    SP_MTX_PAD = 1,        /* pad dims as the padded loaders do          */
    SP_MTX_COMPAT = 2,     /* values = rand()%4+1, file values ignored   */
    SP_MTX_SORT = 4        /* sort by (row, col) and fill `rows`         */
};

/* Reference: readCOOMatrix (spmv/1D/CSR-row/support/matrix.h:44).  Unlike
 * the reference, a missing file or a malformed header returns SP_ERR_IO. */
int sp_read_mtx(const char *path, int dtype, int flags, sp_coo **out);

/* xoshiro256** seeded through SplitMix64: the project PRNG.  Every
 * generator and every test input is drawn from it. */
typedef struct sp_rng { uint64_t s[4]; } sp_rng;
void sp_rng_seed(sp_rng *r, uint64_t seed);
uint64_t sp_rng_next(sp_rng *r);
double sp_rng_uniform(sp_rng *r);          /* [0, 1) with 53 bits      */
double sp_rng_normal(sp_rng *r);           /* N(0,1), Box-Muller       */

/* Value distribution for generated values. */
typedef enum sp_dist_kind {
    SP_DIST_UNIFORM = 0,   /* ints: uniform integer in [lo, hi]; floats: [lo, hi) */
    SP_DIST_NORMAL = 1,    /* N(lo, hi^2): lo = mean, hi = std                   */
    SP_DIST_CONST = 2      /* every value = lo                                   */
} sp_dist_kind;
typedef struct sp_dist { int32_t kind; double lo, hi; } sp_dist;

/* Fill `n` values of dtype from the distribution with the given seed. */
int sp_gen_vector(int dtype, uint64_t n, sp_dist d, uint64_t seed, void *out);

/* Reference: init_vector (spmv/1D/CSR-row/host/app.c:57-61): the compat
 * input vector x[i] = i % 4 + 1. */
int sp_init_vector(int dtype, uint64_t n, void *out);

/* 2D 5-point stencil on a gx*gy grid: row i = y*gx + x; diagonal `diag`,
 * the (up to 4) neighbours drawn from `offdiag`.  Sorted, unique. */
int sp_gen_stencil5(uint32_t gx, uint32_t gy, int dtype, sp_dist offdiag,
                    double diag, uint64_t seed, sp_coo **out);

/* R-MAT over 2^scale x 2^scale with quadrant probabilities (a,b,c,1-a-b-c),
 * sampled until nnz_target distinct (i,j); kept in first-seen order of the
 * sample stream, then sorted by (row, col).  Values from `vals`. */
int sp_gen_rmat(uint32_t scale, uint64_t nnz_target, double a, double b, double c,
                int dtype, sp_dist vals, uint64_t seed, sp_coo **out);

/* 3D 27-point stencil on a g^3 grid with `dof` coupled unknowns per node:
 * dense dof x dof blocks, entries N(0,1); each row's diagonal is raised by
 * (sum of |off-diagonal entries of the row|) + 1.  periodic != 0 wraps the
 * grid so every row has 27*dof entries.  Row = node*dof + d. */
int sp_gen_stencil27(uint32_t g, uint32_t dof, int periodic, int dtype,
                     uint64_t seed, sp_coo **out);

/* Uniform random: each row gets a count drawn uniformly from
 * [nnz_row_lo, nnz_row_hi] of distinct uniform columns. */
int sp_gen_uniform(uint32_t nrows, uint32_t ncols, uint32_t nnz_row_lo,
                   uint32_t nnz_row_hi, int dtype, sp_dist vals, uint64_t seed,
                   sp_coo **out);

/* Banded: row i draws a count from [nnz_row_lo, nnz_row_hi] of distinct
 * columns in [i-bw, i+bw] (clipped). */
int sp_gen_banded(uint32_t nrows, uint32_t bw, uint32_t nnz_row_lo,
                  uint32_t nnz_row_hi, int dtype, sp_dist vals, uint64_t seed,
                  sp_coo **out);

/* Block-diagonal: `nblocks` diagonal blocks of equal size, each row draws a
 * count from [nnz_row_lo, nnz_row_hi] distinct columns inside its block. */
int sp_gen_blockdiag(uint32_t nrows, uint32_t nblocks, uint32_t nnz_row_lo,
                     uint32_t nnz_row_hi, int dtype, sp_dist vals, uint64_t seed,
                     sp_coo **out);

/* ------------------------------------------------------------------------ */
/* Builders (reference: support/matrix.h)                                   */
/* ------------------------------------------------------------------------ */

/* Allocate an empty COO of the given shape (arrays uninitialised). */
int sp_coo_alloc(uint32_t nrows, uint32_t ncols, uint32_t nnz, int dtype, sp_coo **out);
void sp_free_coo(sp_coo *A);           /* freeCOOMatrix   */
void sp_free_csr(sp_csr *A);           /* freeCSRMatrix   */
void sp_free_bcsr(sp_bcsr *A);         /* freeBCSRMatrix  */
void sp_free_bcoo(sp_bcoo *A);         /* freeBCOOMatrix  */
void sp_free_tiled(sp_tiled *A);

/* Reference: sortCOOMatrix + comparator (spmv/1D/COO-row/support/matrix.h:86-125).
 * Sorts by (row, col) and fills `rows`.  Stable for equal keys. */
int sp_sort_coo(sp_coo *A);

/* Reference: coo2csr (spmv/1D/CSR-row/support/matrix.h:113-152): counting
 * sort by row that keeps input order within a row. */
int sp_coo2csr(const sp_coo *A, sp_csr **out);

/* Reference: csr2bcsr (spmv/1D/BCSR-block/support/matrix.h:202-400): blocks
 * of a block row are emitted in first-seen order (row-major scan of the
 * block row); duplicate (i,j) keep the last value.  Linear time: only the
 * block columns touched by a block row are cleared. */
int sp_csr2bcsr(const sp_csr *A, uint32_t row_block_size, uint32_t col_block_size,
                sp_bcsr **out);

/* Reference: sortBCSRMatrix (spmv/1D/BCSR-block/support/matrix.h:460-469):
 * blocks of each block row sorted by block column. */
int sp_sort_bcsr(sp_bcsr *A);

/* Reference: countNNZperBlockBCSRMatrix (spmv/1D/BCSR-block/support/matrix.h:474-485).
 * count_values != 0 counts non-zero VALUES as the reference does;
 * count_values == 0 counts stored entries (what the partitioners' total-nnz
 * sanity checks assume when values can be zero).  `csr` is needed only for
 * the stored-entry count. */
int sp_count_nnz_per_block(sp_bcsr *A, const sp_csr *csr, int count_values);

/* Reference: bcsr2bcoo (spmv/1D/BCOO-block/support/matrix.h:549). */
int sp_bcsr2bcoo(const sp_bcsr *A, sp_bcoo **out);

/* Convenience: csr2bcsr + sort + count, as every B* app.c does. */
int sp_build_bcsr(const sp_csr *A, uint32_t r, uint32_t c, int count_values, sp_bcsr **out);

/* 2D builders.  Input COO must be sorted (sp_sort_coo), as every 2D app.c
 * sorts before building (spmv/2D/RBDCSR/host/app.c:129-130).
 *   sp_coo2dcsr   coo2dcsr   spmv/2D/DCSR/support/matrix.h:159-230
 *   sp_coo2dcoo   coo2dcoo   spmv/2D/DCOO/support/matrix.h:158-217
 *   sp_coo2rbdcsr coo2rbdcsr spmv/2D/RBDCSR/support/matrix.h:159-223
 *   sp_coo2rbdcoo coo2rbdcoo spmv/2D/RBDCOO/support/matrix.h:157
 *   sp_coo2bdcsr  coo2bdcsr  spmv/2D/BDCSR/support/matrix.h:161-282
 *   sp_coo2bdcoo  coo2bdcoo  spmv/2D/BDCOO/support/matrix.h:158
 * sp_tiled2block turns a CSR tiling into its BCSR (as_bcoo = 0) or BCOO
 * (as_bcoo = 1) form: dcsr2dbcsr (spmv/2D/DBCSR/support/matrix.h:278),
 * rbdcsr2rbdbcsr (spmv/2D/RBDBCSR-block/support/matrix.h:270),
 * bdcsr2bdbcsr (spmv/2D/BDBCSR-block/support/matrix.h:334), each followed
 * by the block sort and the nnz-per-block count, and *2*bcoo. */
int sp_coo2tiled(const sp_coo *A, int tiling, int format, uint32_t horz_partitions,
                 uint32_t vert_partitions, uint32_t row_block_size,
                 uint32_t col_block_size, int count_values, sp_tiled **out);

/* ------------------------------------------------------------------------ */
/* Partitioners (reference: support/partition.h; inline splits in app.c)   */
/* All split arrays are caller-allocated.                                   */
/* ------------------------------------------------------------------------ */

/* partition_by_row (spmv/1D/CSR-row/support/partition.h:35-60; COO twin
 * spmv/1D/COO-row/support/partition.h:35): even row chunks, remainder to
 * the first parts.  row_split[nparts+1]. */
int sp_partition_by_row(uint32_t nrows, uint32_t nparts, uint32_t *row_split);

/* partition_by_nnz, row granular (spmv/1D/CSR-nnz/support/partition.h:35-78;
 * COO twin spmv/1D/COO-nnz-rgrn/support/partition.h:35): greedy
 * `curr >= nnz/nparts` cuts.  rowptr[nrows+1].  row_split[nparts+1]. */
int sp_partition_by_nnz(const uint32_t *rowptr, uint32_t nrows, uint32_t nparts,
                        uint32_t *row_split);

/* partition_tsklt_by_row (spmv/1D/CSR-row/support/partition.h:66-88 with
 * granularity 8/sizeof(val); COO uses granularity 1,
 * spmv/1D/COO-row/support/partition.h:67).  split[ntasklets+1], local rows. */
int sp_partition_tsklt_by_row(uint32_t rows, uint32_t ntasklets, uint32_t granularity,
                              uint32_t *split);

/* partition_tsklt_by_nnz (spmv/1D/CSR-row/support/partition.h:92-139;
 * COO twin spmv/1D/COO-row/support/partition.h:93): rows [row_begin,
 * row_begin+rows) of rowptr, nnz = their entry count. */
int sp_partition_tsklt_by_nnz(const uint32_t *rowptr, uint32_t row_begin, uint32_t rows,
                              uint32_t nnz, uint32_t ntasklets, uint32_t granularity,
                              uint32_t *split);

/* Exact nnz split with straddling rows (spmv/1D/COO-nnz/host/app.c:118-156).
 * rows_nnz[nrows] = nnz per row.  Outputs per part: first row, row count,
 * first entry, entry count and the merge flag (part starts mid-row). */
int sp_partition_coo_nnz(const uint32_t *rows_nnz, uint32_t nrows, uint32_t nnz,
                         uint32_t nparts, uint32_t *prev_rows, uint32_t *rows_per_part,
                         uint32_t *prev_nnz, uint32_t *nnz_per_part, uint32_t *merge);

/* Even split of `count` items into ntasklets chunks (remainder to the first
 * chunks): the COO-nnz / DCOO tasklet split (spmv/1D/COO-nnz/host/app.c:186-205;
 * spmv/2D/DCOO/host/app.c:189-207).  start[ntasklets], count_out[ntasklets]. */
int sp_partition_even(uint32_t count, uint32_t ntasklets, uint32_t *start, uint32_t *count_out);

/* BCSR block-row-granular splits (spmv/1D/BCSR-block/support/partition.h:39-113,
 * BCSR-nnz/support/partition.h:39): by_nnz = 0 balances blocks, 1 balances
 * nnz_per_block.  brow_split[nparts+1], blocks_part[nparts], nnzs_part[nparts]. */
int sp_partition_bcsr(const uint32_t *browptr, const uint32_t *nnz_per_block,
                      uint32_t num_block_rows, uint64_t num_blocks, uint32_t nnz,
                      uint32_t nparts, int by_nnz, uint32_t *brow_split,
                      uint32_t *blocks_part, uint32_t *nnzs_part);

/* BCSR tasklet split of part `part` (partition_tsklt_by_block / _by_nnz,
 * spmv/1D/BCSR-block/support/partition.h:119-188,192-262): split[ntasklets+1]
 * in global block-row indices. */
int sp_partition_tsklt_bcsr(const uint32_t *browptr, const uint32_t *nnz_per_block,
                            const uint32_t *brow_split, const uint32_t *blocks_part,
                            const uint32_t *nnzs_part, uint32_t part, uint32_t ntasklets,
                            int by_nnz, uint32_t *split);

/* BCOO block splits: by_nnz = 0 is the exact even block split
 * (spmv/1D/BCOO-block/support/partition.h:37-74), by_nnz = 1 the greedy
 * nnz split at block granularity (spmv/1D/BCOO-nnz/support/partition.h:37-97).
 * block_split[nparts+1], nnzs_part[nparts]. */
int sp_partition_bcoo(const uint32_t *nnz_per_block, uint64_t num_blocks, uint32_t nnz,
                      uint32_t nparts, int by_nnz, uint32_t *block_split,
                      uint32_t *nnzs_part);

/* BCOO tasklet split of part `part` (spmv/1D/BCOO-block/support/partition.h:80-160):
 * by_nnz = 0 even blocks, 1 greedy nnz.  split[ntasklets+1], global blocks. */
int sp_partition_tsklt_bcoo(const uint32_t *nnz_per_block, const uint32_t *block_split,
                            const uint32_t *nnzs_part, uint32_t part, uint32_t ntasklets,
                            int by_nnz, uint32_t *split);

/* find_partitions (spmv/2D/DCSR/host/app.c:66-69): horz = n / vert. */
uint32_t sp_find_partitions(uint32_t nparts, uint32_t vert);

/* RBD/BD per-stripe splits into `horz` parts, split[c*(horz+1) + i],
 * stripe-local: row splits (CSR; spmv/2D/RBDCSR/support/partition.h:45-89),
 * block-row splits (BCSR; by blocks spmv/2D/RBDBCSR-block/support/partition.h:50,
 * by non-zeros spmv/2D/RBDBCSR-nnz/support/partition.h:51), entry splits
 * (COO; nnz_split of spmv/2D/RBDCOO/support/partition.h:45-118) or block
 * splits (BCOO; spmv/2D/RBDBCOO-block/support/partition.h:50,
 * spmv/2D/RBDBCOO-nnz/support/partition.h:50).  by_nnz selects the
 * non-zero-balanced rule for the block formats.  SP_ERR_ARG for D tilings
 * (each D tile is its own part). */
int sp_partition_tiled(const sp_tiled *A, uint32_t horz, int by_nnz, uint32_t *split);

/* ------------------------------------------------------------------------ */
/* Device path                                                              */
/* ------------------------------------------------------------------------ */

typedef struct sp_ctx sp_ctx;
typedef struct sp_plan sp_plan;

/* The 25 SparseP kernel directories, in README tree order. */
typedef enum sp_kernel_id {
    SP_K_CSR_ROW = 0, SP_K_CSR_NNZ = 1, SP_K_COO_ROW = 2, SP_K_COO_NNZ_RGRN = 3,
    SP_K_COO_NNZ = 4, SP_K_BCSR_BLOCK = 5, SP_K_BCSR_NNZ = 6, SP_K_BCOO_BLOCK = 7,
    SP_K_BCOO_NNZ = 8, SP_K_DCSR = 9, SP_K_DCOO = 10, SP_K_DBCSR = 11, SP_K_DBCOO = 12,
    SP_K_RBDCSR = 13, SP_K_RBDCOO = 14, SP_K_RBDBCSR_BLOCK = 15, SP_K_RBDBCSR_NNZ = 16,
    SP_K_RBDBCOO_BLOCK = 17, SP_K_RBDBCOO_NNZ = 18, SP_K_BDCSR = 19, SP_K_BDCOO = 20,
    SP_K_BDBCSR_BLOCK = 21, SP_K_BDBCSR_NNZ = 22, SP_K_BDBCOO_BLOCK = 23,
    SP_K_BDBCOO_NNZ = 24, SP_K_COUNT = 25
} sp_kernel_id;

const char *sp_kernel_name(int id);

/* GPU algorithm inside a variant. */
typedef enum sp_algo {
    SP_ALGO_AUTO = 0,
    /* row granular: `lanes_per_row` lanes per row (1D CSR with
     * items_per_lane 4: lanes of 4-entry chunks; AUTO takes that for even
     * rows of 16 entries or more) */
    SP_ALGO_VECTOR = 1,
    /* entry balanced: fixed entry tiles, rows split.  BCSR ids 5 and 6:
     * merge-path tiles over block-row ends and blocks (AUTO takes them when
     * the longest block row exceeds 4 * average + 64 blocks). */
    SP_ALGO_SEGMENTED = 2,
    /* SEGMENTED with every tile's entries stored in column order, so the
     * lanes of one gather instruction share cache lines; products return to
     * row order in LDS, giving the same sums in the same order.  1D CSR
     * and COO ids (0-4); needs ncols <= 2^32 / (256 * items_per_lane).
     * AUTO takes it in place of SEGMENTED wherever the columns fit (COO:
     * any algo but SEGMENTED does). */
    SP_ALGO_SEGMENTED_SORTED = 3,
    /* SEGMENTED_SORTED with the most frequent columns' x entries staged in
     * LDS (up to 112 KB; 96 KB for 8-byte accumulators): one 1024-thread
     * workgroup per CU, tiles of 4096 merge items.  CSR ids 0 and 1 only;
     * needs ncols <= 2^20.  Several right-hand sides run one at a time. */
    SP_ALGO_SEGMENTED_HOT = 4,
    /* SEGMENTED_SORTED with the rows that cross tiles merged by device
     * atomics instead of carries: the reference's CGLOCK / FGLOCK builds
     * (spmv/1D/COO-nnz/Makefile:7-9: one or 32 hashed mutexes around
     * y[row] += partial, spmv/1D/COO-nnz/dpu/task.c:819-885; likewise
     * spmv/1D/BCOO-nnz/dpu/task.c).  Its LOCKFREE default is the carry
     * merge of SEGMENTED / SEGMENTED_SORTED.  The COO and BCOO ids, 1D
     * (2-4, 7, 8) and 2D (10, 12, 14, 17, 18, 20, 23, 24, without
     * lds_slice).  The BCSR ids (5, 6, 11, 15, 16, 21, 22) accept it and
     * run their default plan: the reference's BCSR lock builds
     * (spmv/1D/BCSR-nnz/Makefile:9) guard the 8-byte MRAM write granules
     * that neighbouring tasklets share, and no BCSR kernel here shares a
     * block row between workgroups without a carry merge.  The CSR ids
     * return SP_ERR_UNSUPPORTED (the reference has no lock build for them).
     * Integer results are exact; floating-point sums of a row spread over
     * three or more tiles can differ in the last bits between runs.  Rows
     * the atomics touch are zeroed by a small kernel before the main one.
     * Batched launches (sp_batch_*) merge these plans with carries. */
    SP_ALGO_SEGMENTED_ATOMIC = 5
} sp_algo;

/* Tasklet-level balance (the reference BLNC_TSKLT_* knobs). */
typedef enum sp_balance { SP_BAL_ROW = 0, SP_BAL_NNZ = 1, SP_BAL_BLOCK = 2 } sp_balance;

typedef struct sp_opts {
    uint32_t nparts;          /* ex NR_DPUS: parts (workgroup groups); 0 = auto  */
    uint32_t nsub;            /* ex NR_TASKLETS: sub-parts per part; 0 = auto   */
    int32_t sub_balance;      /* sp_balance for the sub-part split              */
    int32_t algo;             /* sp_algo                                        */
    uint32_t lanes_per_row;   /* VECTOR: 1,2,4,...,64; 0 = auto                 */
    uint32_t items_per_lane;  /* SEGMENTED: entries per lane per tile; 0 = auto; *
                               * VECTOR (1D and 2D CSR): 4 = each lane loads    *
                               * 4-entry chunks (lanes_per_row lanes of chunks) */
    uint32_t vert;            /* 2D vertical partitions (-v); 0 = auto          */
    uint32_t row_block_size;  /* B*: block rows (-r), default 4                 */
    uint32_t col_block_size;  /* B*: block cols (-c), default 4                 */
    uint32_t lds_slice;       /* 2D CSR/COO: stage each tile's x slice in LDS   *
                               * when every stripe has at most this many        *
                               * columns (and 32 KB); wider: SP_ERR_UNSUPPORTED. *
                               * 0 = off (gathers from global memory)           */
    int32_t count_values;     /* B*: nnz_per_block counts non-zero values       */
    uint64_t mem_cap_bytes;   /* fail with SP_ERR_NOMEM above this; 0 = none    */
} sp_opts;

void sp_opts_default(sp_opts *o);

/* Context: HIP device + a stream (blocking w.r.t. the null stream).
 * Replaces dpu_alloc/dpu_load/dpu_get_nr_dpus (spmv/1D/CSR-row/host/app.c:91-93). */
int sp_ctx_create(int hip_device, sp_ctx **out);
int sp_ctx_destroy(sp_ctx *ctx);
/* Run subsequent work on a caller stream (hipStream_t as void*); NULL = own. */
int sp_ctx_set_stream(sp_ctx *ctx, void *hip_stream);
void *sp_ctx_stream(sp_ctx *ctx);
int sp_ctx_sync(sp_ctx *ctx);
/* Device properties: CUs, LDS per workgroup, HBM bytes, clock, arch name. */
int sp_ctx_info(sp_ctx *ctx, int *num_cus, int *lds_per_block, uint64_t *hbm_bytes,
                char *arch, size_t arch_len);

/* Device memory helpers for C callers (stream-ordered on the ctx stream). */
int sp_dmalloc(sp_ctx *ctx, size_t bytes, void **dptr);
int sp_dfree(sp_ctx *ctx, void *dptr);
int sp_memcpy_h2d(sp_ctx *ctx, void *dst, const void *src, size_t bytes);
int sp_memcpy_d2h(sp_ctx *ctx, void *dst, const void *src, size_t bytes);

/* Build the variant's format from a sorted COO, run its partitioners and
 * upload the matrix.  Replaces the per-DPU partition + argument / matrix
 * pushes (spmv/1D/CSR-row/host/app.c:104-252). */
int sp_plan_create(sp_ctx *ctx, int kernel_id, const sp_coo *A, const sp_opts *opts,
                   sp_plan **out);
/* Same, from a CSR (CSR variants only; saves the COO->CSR build). */
int sp_plan_create_csr(sp_ctx *ctx, int kernel_id, const sp_csr *A, const sp_opts *opts,
                       sp_plan **out);
int sp_plan_destroy(sp_plan *plan);

/* Plan facts: dtype, rows, cols, nnz, algorithmic bytes per SpMV (x and y
 * counted once), device bytes held, and the resolved options. */
typedef struct sp_plan_info {
    int32_t kernel_id, dtype, algo;
    uint32_t nrows, ncols, nnz;
    uint64_t num_blocks;
    uint64_t algo_bytes;      /* algorithmic HBM bytes of one SpMV        */
    uint64_t device_bytes;    /* device memory owned by the plan          */
    uint32_t nparts, nsub, lanes_per_row, items_per_lane, vert, ntiles;
    uint32_t row_block_size, col_block_size;
    uint32_t launches;        /* kernels per SpMV                          */
} sp_plan_info;
int sp_plan_get_info(const sp_plan *plan, sp_plan_info *info);

/* y = A x on the ctx stream; x, y are device pointers of the plan dtype.
 * Replaces x push + dpu_launch + y pull + host merge
 * (spmv/1D/CSR-row/host/app.c:257-287, spmv/2D/DCSR/host/app.c:340-351).
 * A plan keeps per-call device state (carry slots, partial vectors, the
 * staged x of hot-x plans), so calls on one plan are ordered on its
 * context's stream; run two plans to overlap work on two streams. */
int sp_spmv(sp_plan *plan, const void *x, void *y);

/* Y = A X for k right-hand sides; X is ncols x k row-major with leading
 * dimension ldx >= k, Y is nrows x k with ldy >= k (device pointers). */
int sp_spmm(sp_plan *plan, int k, const void *X, int ldx, void *Y, int ldy);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): y = A x for a CSR matrix whose rowptr[nrows+1],
 * colind[nnz] and values[nnz] are already in device memory, with no plan.
 * It runs the entry-balanced merge-path kernel of a CSR-nnz plan over
 * row-ordered 2048-item tiles and allocates nothing: tile coordinates and
 * cross-tile carries live in a library-owned static device buffer of about
 * 1.6 MB per device, held until the process ends.  Limit: nrows + nnz <=
 * SP_NOALLOC_MAX_ITEMS, else SP_ERR_RANGE.  Stream-ordered on the context
 * stream.  Every call on one GPU shares that buffer, so calls that overlap
 * on one GPU (other streams or contexts) race. */
#define SP_NOALLOC_MAX_ITEMS 134217728u
int sp_spmv_csr_noalloc(sp_ctx *ctx, int dtype, uint32_t nrows, uint32_t ncols, uint32_t nnz,
                        const uint32_t *rowptr, const uint32_t *colind, const void *values,
                        const void *x, void *y);

/* Time `iters` SpMVs with HIP events recorded on the ctx stream around each
 * launch group.  flush_bytes > 0 overwrites a scratch buffer of that size
 * between iterations (outside the events) to evict L2 and MALL.
 * ms[iters] receives each iteration's time. */
int sp_time_spmv(sp_plan *plan, const void *x, void *y, int iters, uint64_t flush_bytes,
                 float *ms);
int sp_time_spmm(sp_plan *plan, int k, const void *X, int ldx, void *Y, int ldy,
                 int iters, uint64_t flush_bytes, float *ms);
/* As sp_time_spmv, and ms_dominant[iters] receives the duration of the
 * variant's dominant (first, matrix-streaming) kernel alone, measured with
 * an event recorded right after it on the same stream. */
int sp_time_spmv_split(sp_plan *plan, const void *x, void *y, int iters, uint64_t flush_bytes,
                       float *ms, float *ms_dominant);

/* Batched SpMV / SpMM (no counterpart in the reference, which runs one
 * matrix per launch): the problems of `nprob` 1D plans (same dtype) with
 * k[i] right-hand sides each run as persistent launches, one per
 * (kernel kind, tile shape, RHS width) class, sweeping every problem's work
 * units; with SP_BATCH_GRAPH they are captured once into a hipGraph and
 * replayed.  X[i] is ncols_i x k[i] row-major (leading dimension k[i]),
 * Y[i] nrows_i x k[i]; both device pointers, given per call.  A batch
 * keeps device state: calls on one batch must not overlap. */
typedef struct sp_batch sp_batch;
enum { SP_BATCH_GRAPH = 1 };
int sp_batch_create(sp_ctx *ctx, int nprob, sp_plan *const *plans, const int *k, int flags,
                    sp_batch **out);
int sp_batch_run(sp_batch *b, const void *const *X, void *const *Y);
int sp_batch_time(sp_batch *b, const void *const *X, void *const *Y, int iters,
                  uint64_t flush_bytes, float *ms);
int sp_batch_info(const sp_batch *b, uint32_t *ngroups, uint64_t *algo_bytes);
int sp_batch_destroy(sp_batch *b);
/* create + run + destroy */
int sp_batch_run_once(sp_ctx *ctx, int nprob, sp_plan *const *plans, const void *const *X,
                      void *const *Y, const int *k);
/* The unbatched counterpart, for comparison: one sp_spmm per problem back
 * to back on the ctx stream, `iters` times; ms[i] = HIP-event time of the
 * whole sequence (L2/MALL flushed with flush_bytes before each). */
int sp_time_plans(sp_ctx *ctx, int nprob, sp_plan *const *plans, const void *const *X,
                  void *const *Y, const int *k, int iters, uint64_t flush_bytes, float *ms);

/* Exhaustive autotuner (no counterpart in the reference).  A point is one
 * kernel id with one option set; sp_tune_space enumerates the candidate
 * points for matrix A (kernel_id < 0: all 25 ids), sp_tune_run builds and
 * times each point (median of `iters` SpMVs with HIP events, L2/MALL
 * flushed with flush_bytes between runs) and fills ms / build_ms / status
 * (0 ok, else the sp_status of plan creation), sp_tune_best returns the
 * index of the fastest point (of kernel_id, or of all when < 0), and
 * sp_tune_db_append appends the points as JSON lines to a timing database. */
typedef struct sp_tune_point {
    int32_t kernel_id;
    sp_opts opts;
    float ms;          /* median SpMV time, -1 when not measured */
    float build_ms;    /* plan creation time */
    int32_t status;
    uint64_t algo_bytes;
} sp_tune_point;
int sp_tune_space(const sp_coo *A, int kernel_id, sp_tune_point *pts, int cap, int *count);
int sp_tune_run(sp_ctx *ctx, const sp_coo *A, sp_tune_point *pts, int n, int iters,
                uint64_t flush_bytes);
/* sp_tune_run for Y = A X with k right-hand sides (X and Y row-major, ld = k):
 * the points are timed with sp_spmm, so a multi-RHS workload is tuned for its
 * own width (the kernels differ in how many right-hand sides one pass over
 * the matrix takes). */
int sp_tune_run_spmm(sp_ctx *ctx, const sp_coo *A, sp_tune_point *pts, int n, int iters,
                     uint64_t flush_bytes, int k);
int sp_tune_best(const sp_tune_point *pts, int n, int kernel_id);
int sp_tune_db_append(const char *path, const char *tag, const sp_tune_point *pts, int n);

/* Conjugate gradient on a plan's matrix (square, symmetric positive
 * definite; FP32 or FP64): solves A x = b for device vectors b, x (x holds
 * the initial guess on entry).  Stops when ||b - A x|| / ||b|| <= rtol or
 * after maxit iterations; *iters and *relres report where.  One plan (and
 * its tuned kernel) serves every iteration; dot products reduce in a fixed
 * order.  SP_ERR_STATE when p'Ap <= 0 (matrix not positive definite).  No
 * counterpart in the reference: the iterative-solver caller its README
 * names as the motivating use (README.md:7). */
int sp_cg(sp_plan *A, const void *b, void *x, int maxit, double rtol, int *iters, double *relres);

/* Bandwidth probe: time a device copy of `bytes` (read + write), returns
 * ms per copy; the measured stream-copy peak is 2*bytes/ms. */
int sp_time_copy(sp_ctx *ctx, uint64_t bytes, int iters, float *ms);

/* Bandwidth probe: time a read-only stream over `bytes` (16-byte
 * non-temporal loads, nothing written), returns ms per pass; the measured
 * read peak is bytes/ms.  SpMV streams are almost all reads, so this is the
 * ceiling the SpMV kernels are compared with beside the copy. */
int sp_time_read(sp_ctx *ctx, uint64_t bytes, int iters, float *ms);

/* ---------------------------------------------------------------------
 * Learned variant selector (no counterpart in the reference; its only
 * precursor is the block fill ratio computed in
 * spmv/1D/BCOO-block/support/matrix.h:491-505).
 * ------------------------------------------------------------------- */

/* Cheap structural features of A (SP_NFEAT floats, all O(nnz) host work):
 * size logs, row-length moments and log2 histogram, column-length moments,
 * bandwidth and diagonal-distance histogram, skew (Gini, max/avg), 2x2 and
 * 4x4 block fill ratios, gather locality, dtype one-hot.  A may be unsorted. */
enum { SP_NFEAT = 64 };
int sp_features(const sp_coo *A, float *feat);

/* MLP in -> h1 -> h2 -> out with ReLU, BF16 weights (rounded copies of FP32
 * master weights) and FP32 activations / accumulation, trained with Adam by
 * hand-written HIP kernels.  Parameter layout (FP32, flat):
 * W1[h1][in] b1[h1] W2[h2][h1] b2[h2] W3[out][h2] b3[out].
 * Loss = mean over all B*out outputs of (M*(Y-T))^2 (M = 1 when NULL).
 * X, T, M, Y are host arrays, row-major [B][in] / [B][out]. */
typedef struct sp_mlp sp_mlp;
int sp_mlp_create(sp_ctx *ctx, int in, int h1, int h2, int out, uint64_t seed, sp_mlp **m);
int sp_mlp_destroy(sp_mlp *m);
size_t sp_mlp_num_params(const sp_mlp *m);
int sp_mlp_set_params(sp_mlp *m, const float *params, size_t n);  /* resets Adam step count */
int sp_mlp_get_params(sp_mlp *m, float *params, size_t n);
int sp_mlp_predict(sp_mlp *m, const float *X, int B, float *Y);
/* Forward + backward of one batch: loss (may be NULL) and the gradient with
 * respect to the FP32 parameters (grad may be NULL; kept on the device for
 * sp_mlp_adam either way). */
int sp_mlp_loss_grad(sp_mlp *m, const float *X, const float *T, const float *M, int B,
                     float *loss, float *grad);
int sp_mlp_adam(sp_mlp *m, float lr, float beta1, float beta2, float eps);
/* Shuffled mini-batch Adam (beta 0.9/0.999, eps 1e-8) over n samples;
 * final_loss = mean batch loss of the last epoch. */
int sp_mlp_train(sp_mlp *m, const float *X, const float *T, const float *M, int n, int epochs,
                 int batch, float lr, uint64_t seed, float *final_loss);

/* The selector: an MLP SP_NFEAT -> h1 -> h2 -> SP_K_COUNT over standardised
 * features ((f - mean) / std) whose outputs score each kernel id (trained as
 * log2 of the id's default-configuration time over the exhaustive best, so
 * lower is better).  sp_select picks the lowest score with default options;
 * when the gap to the second-lowest score is below `margin` and
 * fallback_k > 0, it instead tunes every point of the fallback_k
 * lowest-scored ids (sp_tune_run with iters / flush_bytes) and returns the
 * fastest (*fell_back = 1).  scores (SP_K_COUNT floats) may be NULL.
 * Model files are written by sp_selector_save (magic "SPSEL001"). */
typedef struct sp_selector sp_selector;
int sp_selector_create(sp_ctx *ctx, const int *sizes, const float *params, size_t nparams,
                       const float *feat_mean, const float *feat_std, float margin,
                       sp_selector **s);
int sp_selector_load(sp_ctx *ctx, const char *path, sp_selector **s);
int sp_selector_save(const sp_selector *s, const char *path);
int sp_selector_destroy(sp_selector *s);
int sp_select(sp_selector *s, const sp_coo *A, int fallback_k, int iters, uint64_t flush_bytes,
              sp_tune_point *choice, float *scores, int *fell_back);

#ifdef __cplusplus
}
#endif
#endif /* SPARSEP_H */
