// Internal plan / context structures shared by the runtime and the kernel
// launchers.  Not part of the C ABI.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

#include "sparsep.h"

struct sp_ctx {
    int device = 0;
    hipStream_t own_stream = nullptr;
    hipStream_t stream = nullptr;
    int num_cus = 0;
    int lds_per_block = 0;
    uint64_t hbm_bytes = 0;
    char arch[64] = {0};
    void *flush_buf = nullptr;
    uint64_t flush_bytes = 0;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    // When set, launchers record it right after the dominant kernel so the
    // timing loop can report that kernel's duration on its own.
    hipEvent_t ev_mid = nullptr;
    bool mark_mid = false;
};

// Record ctx->ev_mid after the dominant kernel when a timing loop asks.
inline void mark_dominant(sp_ctx *c) {
    if (c->mark_mid) (void)hipEventRecord(c->ev_mid, c->stream);
}

// Device-resident matrix of one plan.  Which arrays are set depends on the
// format; counts are host copies.
struct sp_dev_matrix {
    int format = SP_FMT_CSR;  // sp_format
    uint32_t nrows = 0, ncols = 0, nnz = 0;
    // element formats
    uint32_t *rowptr = nullptr;   // CSR [nrows+1]
    uint32_t *rowind = nullptr;   // COO [nnz]
    uint32_t *colind = nullptr;   // [nnz]
    void *vals = nullptr;         // [nnz]
    // block formats
    uint32_t r = 0, c = 0, nbr = 0, nbc = 0;
    uint64_t nb = 0;
    uint32_t *browptr = nullptr;  // BCSR [nbr+1]
    uint32_t *browind = nullptr;  // BCOO [nb]
    uint32_t *bcolind = nullptr;  // [nb]
    void *bvals = nullptr;        // [nb*r*c]
};

struct sp_plan {
    sp_ctx *ctx = nullptr;
    int id = 0, dtype = 0, algo = 0;
    sp_opts opts{};
    uint32_t nrows = 0, ncols = 0, nnz = 0;
    uint64_t nb = 0;
    uint64_t algo_bytes = 0;
    uint64_t device_bytes = 0;
    uint32_t nparts = 0, nsub = 0, lanes = 0, ipt = 0, vert = 0, ntiles = 0;
    uint32_t launches = 1;

    sp_dev_matrix m;
    // work decomposition
    uint32_t *d_part = nullptr;      // row / block-row / block / entry splits [nparts+1]
    uint32_t *d_tile_row = nullptr;  // segmented: tile start rows [ntiles+1]
    uint32_t *d_tile_pos = nullptr;  // segmented: tile start entries [ntiles+1]
    uint32_t *d_carry_row = nullptr; // [ntiles]
    void *d_carry_val = nullptr;     // [ntiles * k_max] accumulator values
    uint32_t carry_k = 0;            // right-hand sides the carry buffer holds

    // 2D tiled plans: per-tile sub-plans run into a partial-y buffer.
    std::vector<sp_plan *> sub;
    std::vector<uint32_t> sub_row_off, sub_col_off, sub_rows;
    void *d_ypart = nullptr;         // [vert][nrows] partials
    uint64_t ypart_elems = 0;

    std::vector<void *> allocs;
};

// Allocation helpers (record in plan->allocs, respect mem_cap_bytes).
int plan_alloc(sp_plan *p, size_t bytes, void **out);
int plan_upload(sp_plan *p, const void *host, size_t bytes, void **out);

// Launchers, per format.  Return SP_OK or an error; enqueue on p->ctx->stream.
int csr_setup(sp_plan *p, const sp_csr *A);
int csr_run(sp_plan *p, const void *x, void *y, int k, int ldx, int ldy);
int coo_setup(sp_plan *p, const sp_coo *A);
int coo_run(sp_plan *p, const void *x, void *y, int k, int ldx, int ldy);
int bcsr_setup(sp_plan *p, const sp_bcsr *A);
int bcsr_run(sp_plan *p, const void *x, void *y, int k, int ldx, int ldy);
int bcoo_setup(sp_plan *p, const sp_bcoo *A);
int bcoo_run(sp_plan *p, const void *x, void *y, int k, int ldx, int ldy);

// Utility kernels (util.hip).
int launch_fill_zero(hipStream_t s, void *ptr, uint64_t bytes);
int launch_flush(hipStream_t s, void *buf, uint64_t bytes, uint32_t salt);
int launch_copy(hipStream_t s, const void *src, void *dst, uint64_t bytes);
int launch_reduce_parts(hipStream_t s, int dtype, const void *ypart, uint32_t vert, uint64_t stride,
                        void *y, uint64_t n, int k, int ldy);

#define SP_HIP_TRY(expr)                                  \
    do {                                                  \
        hipError_t _e = (expr);                           \
        if (_e != hipSuccess) return SP_ERR_HIP;          \
    } while (0)
