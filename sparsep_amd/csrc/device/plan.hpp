// Internal plan / context structures shared by the runtime, the plan
// builders and the launchers.  Not part of the C ABI.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

#include "sparsep.h"
#include "work.hpp"

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
    // When set, launchers record ev_mid right after the dominant kernel so
    // the timing loop can report that kernel's duration on its own.
    hipEvent_t ev_mid = nullptr;
    bool mark_mid = false;
};

// Record ctx->ev_mid after the dominant kernel when a timing loop asks.
inline void mark_dominant(sp_ctx *c) {
    if (c->mark_mid) (void)hipEventRecord(c->ev_mid, c->stream);
}

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

    Work work;                // the main launch group
    // host copies of the work arrays (batches concatenate them)
    std::vector<spd::Prob> h_probs;
    std::vector<uint32_t> h_unit_pid, h_c0, h_c1, h_c2, h_zpid, h_zrow;
    // 2D tilings: the work writes per-stripe partial rows into ypart
    // ([vert][nrows][k]); a reduction adds them into y in stripe order.
    bool tiled = false;
    void *d_ypart = nullptr;
    uint64_t ypart_bytes = 0;
    uint32_t ypart_k = 0;     // right-hand sides ypart currently holds

    std::vector<void *> allocs;
};

// Allocation helpers (record in plan->allocs, respect mem_cap_bytes).
int plan_alloc(sp_plan *p, size_t bytes, void **out);
int plan_upload(sp_plan *p, const void *host, size_t bytes, void **out);

// Allocation-free entry point (noalloc.hip): look up this device's static
// scratch buffer; the device must be current.
int sp_noalloc_setup(int dev);

// Plan builders (plans.hip).
int build_csr_plan(sp_plan *p, const sp_csr *A);
int build_coo_plan(sp_plan *p, const sp_coo *A);
int build_bcsr_plan(sp_plan *p, const sp_bcsr *B);
int build_bcoo_plan(sp_plan *p, const sp_bcoo *B);
int build_tiled_plan(sp_plan *p, const sp_coo *A);
int run_plan(sp_plan *p, const void *x, void *y, int k, int ldx, int ldy);

// Utility kernels (util.hip).
int launch_fill_zero(hipStream_t s, void *ptr, uint64_t bytes);
int launch_flush(hipStream_t s, void *buf, uint64_t bytes, uint32_t salt);
int launch_copy(hipStream_t s, const void *src, void *dst, uint64_t bytes);
int launch_read(hipStream_t s, const void *src, uint64_t bytes, uint32_t *sink);
int launch_reduce_parts(hipStream_t s, int dtype, const void *ypart, uint32_t vert, uint64_t stride,
                        void *y, uint64_t n, int k, int ldy);

#define SP_HIP_TRY(expr)                                  \
    do {                                                  \
        hipError_t _e = (expr);                           \
        if (_e != hipSuccess) return SP_ERR_HIP;          \
    } while (0)
