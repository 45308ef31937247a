// Context, plan and timing entry points of the C ABI.  This replaces the
// UPMEM host-library orchestration of every host/app.c: dpu_alloc/dpu_load
// (context), the per-DPU partition + argument and matrix pushes (plan
// creation), dpu_launch + x push + y pull + host merge (sp_spmv).
#include <string.h>

#include <algorithm>
#include <vector>

#include "plan.hpp"

static const char *const kNames[SP_K_COUNT] = {
    "CSR-row",       "CSR-nnz",       "COO-row",       "COO-nnz-rgrn",  "COO-nnz",
    "BCSR-block",    "BCSR-nnz",      "BCOO-block",    "BCOO-nnz",      "DCSR",
    "DCOO",          "DBCSR",         "DBCOO",         "RBDCSR",        "RBDCOO",
    "RBDBCSR-block", "RBDBCSR-nnz",   "RBDBCOO-block", "RBDBCOO-nnz",   "BDCSR",
    "BDCOO",         "BDBCSR-block",  "BDBCSR-nnz",    "BDBCOO-block",  "BDBCOO-nnz"};

extern "C" const char *sp_kernel_name(int id) {
    return (id >= 0 && id < SP_K_COUNT) ? kNames[id] : "unknown";
}

extern "C" void sp_opts_default(sp_opts *o) {
    if (!o) return;
    memset(o, 0, sizeof(*o));
    o->sub_balance = SP_BAL_NNZ;
    o->algo = SP_ALGO_AUTO;
    o->row_block_size = 4;
    o->col_block_size = 4;
}

// ---------------------------------------------------------------------------
// Context
// ---------------------------------------------------------------------------

extern "C" int sp_ctx_create(int dev, sp_ctx **out) {
    if (!out) return SP_ERR_ARG;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return SP_ERR_NODEVICE;
    if (dev < 0 || dev >= n) return SP_ERR_ARG;
    SP_HIP_TRY(hipSetDevice(dev));
    sp_ctx *c = new sp_ctx();
    c->device = dev;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, dev) != hipSuccess) { delete c; return SP_ERR_HIP; }
    c->num_cus = prop.multiProcessorCount;
    c->lds_per_block = (int)prop.sharedMemPerBlock;
    c->hbm_bytes = prop.totalGlobalMem;
    strncpy(c->arch, prop.gcnArchName, sizeof(c->arch) - 1);
    if (hipStreamCreate(&c->own_stream) != hipSuccess) { delete c; return SP_ERR_HIP; }
    c->stream = c->own_stream;
    // the allocation-free entry point's scratch address on this device (a
    // failure here is retried by that entry point's first call)
    (void)sp_noalloc_setup(dev);
    if (hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess ||
        hipEventCreate(&c->ev_mid) != hipSuccess) {
        hipStreamDestroy(c->own_stream);
        delete c;
        return SP_ERR_HIP;
    }
    *out = c;
    return SP_OK;
}

extern "C" int sp_ctx_destroy(sp_ctx *c) {
    if (!c) return SP_ERR_ARG;
    hipSetDevice(c->device);
    hipStreamSynchronize(c->stream);
    if (c->flush_buf) hipFree(c->flush_buf);
    hipEventDestroy(c->ev0);
    hipEventDestroy(c->ev1);
    hipEventDestroy(c->ev_mid);
    hipStreamDestroy(c->own_stream);
    delete c;
    return SP_OK;
}

extern "C" int sp_ctx_set_stream(sp_ctx *c, void *s) {
    if (!c) return SP_ERR_ARG;
    c->stream = s ? (hipStream_t)s : c->own_stream;
    return SP_OK;
}

extern "C" void *sp_ctx_stream(sp_ctx *c) { return c ? (void *)c->stream : nullptr; }

extern "C" int sp_ctx_sync(sp_ctx *c) {
    if (!c) return SP_ERR_ARG;
    SP_HIP_TRY(hipStreamSynchronize(c->stream));
    return SP_OK;
}

extern "C" int sp_ctx_info(sp_ctx *c, int *cus, int *lds, uint64_t *hbm, char *arch, size_t len) {
    if (!c) return SP_ERR_ARG;
    if (cus) *cus = c->num_cus;
    if (lds) *lds = c->lds_per_block;
    if (hbm) *hbm = c->hbm_bytes;
    if (arch && len) {
        strncpy(arch, c->arch, len - 1);
        arch[len - 1] = 0;
    }
    return SP_OK;
}

extern "C" int sp_dmalloc(sp_ctx *c, size_t bytes, void **p) {
    if (!c || !p) return SP_ERR_ARG;
    hipSetDevice(c->device);
    if (hipMalloc(p, bytes ? bytes : 1) != hipSuccess) return SP_ERR_NOMEM;
    return SP_OK;
}
extern "C" int sp_dfree(sp_ctx *c, void *p) {
    if (!c) return SP_ERR_ARG;
    if (p) SP_HIP_TRY(hipFree(p));
    return SP_OK;
}
extern "C" int sp_memcpy_h2d(sp_ctx *c, void *dst, const void *src, size_t bytes) {
    if (!c) return SP_ERR_ARG;
    SP_HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c->stream));
    SP_HIP_TRY(hipStreamSynchronize(c->stream));
    return SP_OK;
}
extern "C" int sp_memcpy_d2h(sp_ctx *c, void *dst, const void *src, size_t bytes) {
    if (!c) return SP_ERR_ARG;
    SP_HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->stream));
    SP_HIP_TRY(hipStreamSynchronize(c->stream));
    return SP_OK;
}

// ---------------------------------------------------------------------------
// Plan
// ---------------------------------------------------------------------------

// Every plan array gets at least 32 bytes of slack past its end, so kernels may
// read whole 4-entry chunks of an index or value array (up to 3 entries, 24
// bytes, past the last one) without leaving the allocation.
int plan_alloc(sp_plan *p, size_t bytes, void **out) {
    bytes = (bytes + 32 + 255) & ~(size_t)255;
    if (p->opts.mem_cap_bytes && p->device_bytes + bytes > p->opts.mem_cap_bytes) return SP_ERR_NOMEM;
    void *d = nullptr;
    if (hipMalloc(&d, bytes) != hipSuccess) return SP_ERR_NOMEM;
    p->allocs.push_back(d);
    p->device_bytes += bytes;
    *out = d;
    return SP_OK;
}

int plan_upload(sp_plan *p, const void *host, size_t bytes, void **out) {
    int rc = plan_alloc(p, bytes, out);
    if (rc) return rc;
    if (bytes) SP_HIP_TRY(hipMemcpy(*out, host, bytes, hipMemcpyHostToDevice));
    return SP_OK;
}

static void plan_free(sp_plan *p) {
    if (!p) return;
    for (void *d : p->allocs) (void)hipFree(d);
    delete p;
}

static bool is_csr(int id) { return id == SP_K_CSR_ROW || id == SP_K_CSR_NNZ; }
static bool is_coo(int id) { return id >= SP_K_COO_ROW && id <= SP_K_COO_NNZ; }
static bool is_bcsr(int id) { return id == SP_K_BCSR_BLOCK || id == SP_K_BCSR_NNZ; }
static bool is_bcoo(int id) { return id == SP_K_BCOO_BLOCK || id == SP_K_BCOO_NNZ; }

static uint64_t csr_bytes(uint64_t n, uint64_t m, uint64_t nnz, uint64_t sv) {
    return nnz * (sv + 4) + (n + 1) * 4 + (m + n) * sv;
}

// The BCSR variants, 1D and 2D.  Their reference builds also take CGLOCK /
// FGLOCK (spmv/1D/BCSR-nnz/Makefile:9), where the locks guard the 8-byte MRAM
// write granules that neighbouring tasklets' block rows share.  No BCSR
// kernel here shares a block row between workgroups without a carry merge,
// so a lock build maps to the default plan.
static bool is_block_csr_id(int id) {
    return id == SP_K_BCSR_BLOCK || id == SP_K_BCSR_NNZ || id == SP_K_DBCSR || id == SP_K_RBDBCSR_BLOCK ||
           id == SP_K_RBDBCSR_NNZ || id == SP_K_BDBCSR_BLOCK || id == SP_K_BDBCSR_NNZ;
}

static sp_plan *new_plan(sp_ctx *ctx, int id, int dtype, const sp_opts *opts) {
    sp_plan *p = new sp_plan();
    p->ctx = ctx;
    p->id = id;
    p->dtype = dtype;
    if (opts) p->opts = *opts;
    else sp_opts_default(&p->opts);
    if (p->opts.row_block_size == 0) p->opts.row_block_size = 4;
    if (p->opts.col_block_size == 0) p->opts.col_block_size = 4;
    if (p->opts.algo == SP_ALGO_SEGMENTED_ATOMIC && is_block_csr_id(id)) p->opts.algo = SP_ALGO_AUTO;
    p->algo = p->opts.algo;
    return p;
}

// The algorithm asked for exists and applies to the kernel id.
static int check_algo(int id, const sp_opts *opts) {
    const int algo = opts ? opts->algo : SP_ALGO_AUTO;
    if (algo < SP_ALGO_AUTO || algo > SP_ALGO_SEGMENTED_ATOMIC) return SP_ERR_ARG;
    // the variants whose reference kernels merge shared rows under locks (COO and BCOO
    // formats, 1D and 2D; BCSR ids take the request and run their default plan)
    static const int atomic_ids[] = {SP_K_COO_ROW, SP_K_COO_NNZ_RGRN, SP_K_COO_NNZ, SP_K_BCOO_BLOCK, SP_K_BCOO_NNZ,
                                     SP_K_DCOO, SP_K_DBCOO, SP_K_RBDCOO, SP_K_RBDBCOO_BLOCK, SP_K_RBDBCOO_NNZ,
                                     SP_K_BDCOO, SP_K_BDBCOO_BLOCK, SP_K_BDBCOO_NNZ};
    if (algo == SP_ALGO_SEGMENTED_ATOMIC && !is_block_csr_id(id) &&
        std::find(std::begin(atomic_ids), std::end(atomic_ids), id) == std::end(atomic_ids))
        return SP_ERR_UNSUPPORTED;
    return SP_OK;
}

extern "C" int sp_plan_create_csr(sp_ctx *ctx, int id, const sp_csr *A, const sp_opts *opts,
                                  sp_plan **out) {
    if (!ctx || !A || !out || !is_csr(id)) return SP_ERR_ARG;
    if (int rc = check_algo(id, opts)) return rc;
    if (sp_dtype_size(A->dtype) == 0) return SP_ERR_ARG;
    (void)hipSetDevice(ctx->device);
    sp_plan *p = new_plan(ctx, id, A->dtype, opts);
    p->nrows = A->nrows;
    p->ncols = A->ncols;
    p->nnz = A->nnz;
    p->algo_bytes = csr_bytes(A->nrows, A->ncols, A->nnz, sp_dtype_size(A->dtype));
    int rc = build_csr_plan(p, A);
    if (rc) { plan_free(p); return rc; }
    *out = p;
    return SP_OK;
}

extern "C" int sp_plan_create(sp_ctx *ctx, int id, const sp_coo *A, const sp_opts *opts, sp_plan **out) {
    if (!ctx || !A || !out || id < 0 || id >= SP_K_COUNT) return SP_ERR_ARG;
    if (int rc = check_algo(id, opts)) return rc;
    if (sp_dtype_size(A->dtype) == 0) return SP_ERR_ARG;
    if (A->nnz && (!A->rowind || !A->colind || !A->values)) return SP_ERR_ARG;
    for (uint32_t k = 0; k < A->nnz; k++)
        if (A->rowind[k] >= A->nrows || A->colind[k] >= A->ncols) return SP_ERR_ARG;
    (void)hipSetDevice(ctx->device);
    if (is_csr(id)) {
        sp_csr *C = nullptr;
        int rc = sp_coo2csr(A, &C);
        if (rc) return rc;
        rc = sp_plan_create_csr(ctx, id, C, opts, out);
        sp_free_csr(C);
        return rc;
    }
    sp_plan *p = new_plan(ctx, id, A->dtype, opts);
    p->nrows = A->nrows;
    p->ncols = A->ncols;
    p->nnz = A->nnz;
    const uint64_t sv = sp_dtype_size(A->dtype);
    int rc = SP_OK;
    if (is_coo(id)) {
        p->algo_bytes = (uint64_t)A->nnz * (8 + sv) + ((uint64_t)A->nrows + A->ncols) * sv;
        rc = build_coo_plan(p, A);
    } else if (is_bcsr(id) || is_bcoo(id)) {
        sp_csr *C = nullptr;
        sp_bcsr *B = nullptr;
        rc = sp_coo2csr(A, &C);
        if (!rc) rc = sp_build_bcsr(C, p->opts.row_block_size, p->opts.col_block_size, 0, &B);
        if (!rc) rc = sp_count_nnz_per_block(B, C, p->opts.count_values ? 1 : 0);
        sp_free_csr(C);
        if (!rc) {
            const uint64_t rc2 = (uint64_t)B->row_block_size * B->col_block_size;
            p->nb = B->num_blocks;
            if (is_bcsr(id)) {
                p->algo_bytes = B->num_blocks * (rc2 * sv + 4) + ((uint64_t)B->num_block_rows + 1) * 4 +
                                ((uint64_t)A->nrows + A->ncols) * sv;
                rc = build_bcsr_plan(p, B);
            } else {
                sp_bcoo *O = nullptr;
                rc = sp_bcsr2bcoo(B, &O);
                p->algo_bytes = B->num_blocks * (rc2 * sv + 8) + ((uint64_t)A->nrows + A->ncols) * sv;
                if (!rc) rc = build_bcoo_plan(p, O);
                sp_free_bcoo(O);
            }
        }
        sp_free_bcsr(B);
    } else {
        rc = build_tiled_plan(p, A);
    }
    if (rc) { plan_free(p); return rc; }
    *out = p;
    return SP_OK;
}

extern "C" int sp_plan_destroy(sp_plan *p) {
    if (!p) return SP_ERR_ARG;
    (void)hipSetDevice(p->ctx->device);
    (void)hipStreamSynchronize(p->ctx->stream);
    plan_free(p);
    return SP_OK;
}

extern "C" int sp_plan_get_info(const sp_plan *p, sp_plan_info *i) {
    if (!p || !i) return SP_ERR_ARG;
    memset(i, 0, sizeof(*i));
    i->kernel_id = p->id;
    i->dtype = p->dtype;
    i->algo = p->algo;
    i->nrows = p->nrows;
    i->ncols = p->ncols;
    i->nnz = p->nnz;
    i->num_blocks = p->nb;
    i->algo_bytes = p->algo_bytes;
    i->device_bytes = p->device_bytes;
    i->nparts = p->nparts;
    i->nsub = p->nsub;
    i->lanes_per_row = p->lanes;
    i->items_per_lane = p->ipt;
    i->vert = p->vert;
    i->ntiles = p->ntiles;
    i->row_block_size = p->opts.row_block_size;
    i->col_block_size = p->opts.col_block_size;
    i->launches = p->launches + (p->tiled ? 1 : 0);
    return SP_OK;
}

static int plan_run(sp_plan *p, const void *x, void *y, int k, int ldx, int ldy) {
    return run_plan(p, x, y, k, ldx, ldy);
}

extern "C" int sp_spmv(sp_plan *p, const void *x, void *y) {
    if (!p || (!x && p->ncols) || (!y && p->nrows)) return SP_ERR_ARG;
    return plan_run(p, x, y, 1, 1, 1);
}

extern "C" int sp_spmm(sp_plan *p, int k, const void *X, int ldx, void *Y, int ldy) {
    if (!p || k <= 0 || ldx < k || ldy < k || (!X && p->ncols) || (!Y && p->nrows)) return SP_ERR_ARG;
    return plan_run(p, X, Y, k, ldx, ldy);
}

// ---------------------------------------------------------------------------
// Timing
// ---------------------------------------------------------------------------

static int ensure_flush(sp_ctx *c, uint64_t bytes) {
    if (bytes == 0) return SP_OK;
    if (c->flush_bytes >= bytes) return SP_OK;
    if (c->flush_buf) hipFree(c->flush_buf);
    c->flush_buf = nullptr;
    c->flush_bytes = 0;
    if (hipMalloc(&c->flush_buf, bytes) != hipSuccess) return SP_ERR_NOMEM;
    c->flush_bytes = bytes;
    return SP_OK;
}

static int time_loop(sp_plan *p, int k, const void *x, int ldx, void *y, int ldy, int iters,
                     uint64_t flush_bytes, float *ms, float *ms_dominant) {
    if (!p || iters <= 0 || !ms) return SP_ERR_ARG;
    sp_ctx *c = p->ctx;
    hipSetDevice(c->device);
    int rc = ensure_flush(c, flush_bytes);
    if (rc) return rc;
    for (int it = 0; it < iters; it++) {
        if (flush_bytes) {
            rc = launch_flush(c->stream, c->flush_buf, flush_bytes, 0x9E3779B9u * (uint32_t)(it + 1));
            if (rc) return rc;
        }
        SP_HIP_TRY(hipEventRecord(c->ev0, c->stream));
        c->mark_mid = ms_dominant != nullptr;
        rc = plan_run(p, x, y, k, ldx, ldy);
        c->mark_mid = false;
        if (rc) return rc;
        SP_HIP_TRY(hipEventRecord(c->ev1, c->stream));
        SP_HIP_TRY(hipEventSynchronize(c->ev1));
        float t = 0.f;
        SP_HIP_TRY(hipEventElapsedTime(&t, c->ev0, c->ev1));
        ms[it] = t;
        if (ms_dominant) {
            float td = 0.f;
            SP_HIP_TRY(hipEventElapsedTime(&td, c->ev0, c->ev_mid));
            ms_dominant[it] = td;
        }
    }
    return SP_OK;
}

extern "C" int sp_time_spmv(sp_plan *p, const void *x, void *y, int iters, uint64_t flush_bytes, float *ms) {
    return time_loop(p, 1, x, 1, y, 1, iters, flush_bytes, ms, nullptr);
}

extern "C" int sp_time_spmv_split(sp_plan *p, const void *x, void *y, int iters, uint64_t flush_bytes,
                                  float *ms, float *ms_dominant) {
    if (!ms_dominant) return SP_ERR_ARG;
    return time_loop(p, 1, x, 1, y, 1, iters, flush_bytes, ms, ms_dominant);
}

extern "C" int sp_time_spmm(sp_plan *p, int k, const void *X, int ldx, void *Y, int ldy, int iters,
                            uint64_t flush_bytes, float *ms) {
    if (k <= 0 || ldx < k || ldy < k) return SP_ERR_ARG;
    return time_loop(p, k, X, ldx, Y, ldy, iters, flush_bytes, ms, nullptr);
}

extern "C" int sp_time_copy(sp_ctx *c, uint64_t bytes, int iters, float *ms) {
    if (!c || !ms || iters <= 0 || bytes < 16) return SP_ERR_ARG;
    hipSetDevice(c->device);
    void *a = nullptr, *b = nullptr;
    if (hipMalloc(&a, bytes) != hipSuccess) return SP_ERR_NOMEM;
    if (hipMalloc(&b, bytes) != hipSuccess) { hipFree(a); return SP_ERR_NOMEM; }
    int rc = launch_fill_zero(c->stream, a, bytes);
    for (int it = 0; it < iters && !rc; it++) {
        if (hipEventRecord(c->ev0, c->stream) != hipSuccess) { rc = SP_ERR_HIP; break; }
        rc = launch_copy(c->stream, a, b, bytes);
        if (rc) break;
        if (hipEventRecord(c->ev1, c->stream) != hipSuccess || hipEventSynchronize(c->ev1) != hipSuccess) {
            rc = SP_ERR_HIP;
            break;
        }
        float t = 0.f;
        hipEventElapsedTime(&t, c->ev0, c->ev1);
        ms[it] = t;
    }
    hipStreamSynchronize(c->stream);
    hipFree(a);
    hipFree(b);
    return rc;
}

extern "C" int sp_time_read(sp_ctx *c, uint64_t bytes, int iters, float *ms) {
    if (!c || !ms || iters <= 0 || bytes < 16) return SP_ERR_ARG;
    hipSetDevice(c->device);
    void *a = nullptr, *sink = nullptr;
    if (hipMalloc(&a, bytes) != hipSuccess) return SP_ERR_NOMEM;
    if (hipMalloc(&sink, 256) != hipSuccess) { hipFree(a); return SP_ERR_NOMEM; }
    int rc = launch_fill_zero(c->stream, a, bytes);
    for (int it = 0; it < iters && !rc; it++) {
        if (hipEventRecord(c->ev0, c->stream) != hipSuccess) { rc = SP_ERR_HIP; break; }
        rc = launch_read(c->stream, a, bytes, (uint32_t *)sink);
        if (rc) break;
        if (hipEventRecord(c->ev1, c->stream) != hipSuccess || hipEventSynchronize(c->ev1) != hipSuccess) {
            rc = SP_ERR_HIP;
            break;
        }
        float t = 0.f;
        hipEventElapsedTime(&t, c->ev0, c->ev1);
        ms[it] = t;
    }
    hipStreamSynchronize(c->stream);
    hipFree(a);
    hipFree(sink);
    return rc;
}
