// Exhaustive autotuner: enumerates the variant space (the 25 kernel ids and
// their GPU knobs), builds and times every point with HIP events, and keeps
// a JSON-lines timing database.  The learned selector is trained on these
// timings.  No counterpart in the reference.
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <thread>
#include <vector>

#include "plan.hpp"

namespace {

struct Pt {
    int id;
    sp_opts o;
};

}  // namespace

// Candidate points of one kernel id for matrix A.  The first point of every
// id is its default configuration (sp_opts_default: automatic variant and
// knob choice), the one the learned selector runs; the others vary one knob
// each (row statistics decide the lane counts worth trying).
static void space_for(const sp_coo *A, int id, std::vector<Pt> &v) {
    sp_opts base;
    sp_opts_default(&base);
    v.push_back(Pt{id, base});
    const double avg = A->nrows ? (double)A->nnz / A->nrows : 1.0;
    uint32_t Lc = 1;
    while (Lc < 64 && Lc * 2 <= avg * 1.5) Lc *= 2;
    if (id == SP_K_CSR_ROW || id == SP_K_CSR_NNZ) {
        for (uint32_t L : {Lc / 2, Lc, Lc * 2}) {
            if (L < 1 || L > 64) continue;
            Pt p{id, base};
            p.o.algo = SP_ALGO_VECTOR;
            p.o.lanes_per_row = L;
            v.push_back(p);
        }
        // the row kernel with 4-entry chunk loads, lanes around the average row's chunks
        uint32_t Lq = 4;
        while (Lq < 64 && 4.0 * Lq < avg) Lq *= 2;
        for (uint32_t L : {Lq / 2, Lq, Lq * 2}) {
            if (L < 4 || L > 64) continue;
            Pt p{id, base};
            p.o.algo = SP_ALGO_VECTOR;
            p.o.lanes_per_row = L;
            p.o.items_per_lane = 4;
            v.push_back(p);
        }
        for (uint32_t I : {4u, 8u, 16u}) {
            Pt p{id, base};
            p.o.algo = SP_ALGO_SEGMENTED;
            p.o.items_per_lane = I;
            v.push_back(p);
        }
        if (A->ncols <= (1u << 20)) {  // hot-x tiles (4096 items, col << 12 fits 32 bits)
            Pt p{id, base};
            p.o.algo = SP_ALGO_SEGMENTED_HOT;
            v.push_back(p);
        }
        // column-ordered tiles where col << log2(256 * I) fits 32 bits
        for (uint32_t I : {4u, 8u, 16u}) {
            const uint32_t pb = I == 4 ? 10 : I == 8 ? 11 : 12;
            if (A->ncols > (1ull << (32 - pb))) continue;
            Pt p{id, base};
            p.o.algo = SP_ALGO_SEGMENTED_SORTED;
            p.o.items_per_lane = I;
            v.push_back(p);
        }
    } else if (id >= SP_K_COO_ROW && id <= SP_K_COO_NNZ) {
        for (uint32_t I : {8u, 16u}) {
            Pt p{id, base};
            p.o.items_per_lane = I;
            v.push_back(p);
        }
        Pt p{id, base};  // row-ordered tiles (the default takes column order where it fits)
        p.o.algo = SP_ALGO_SEGMENTED;
        v.push_back(p);
    } else if (id >= SP_K_BCSR_BLOCK && id <= SP_K_BCOO_NNZ) {
        Pt p{id, base};
        p.o.row_block_size = p.o.col_block_size = 2;
        v.push_back(p);
        if (id == SP_K_BCSR_BLOCK || id == SP_K_BCSR_NNZ) {  // entry-balanced BCSR (automatic only when skewed)
            Pt q{id, base};
            q.o.algo = SP_ALGO_SEGMENTED;
            v.push_back(q);
        }
    } else {
        for (uint32_t vert : {4u, 8u}) {
            Pt p{id, base};
            p.o.vert = vert;
            v.push_back(p);
        }
        // 2D CSR/COO: x slices in LDS, with the fewest stripes (up to 64) whose
        // even width fits 32 KB of values
        const bool csr_coo = id == SP_K_DCSR || id == SP_K_DCOO || id == SP_K_RBDCSR || id == SP_K_RBDCOO ||
                             id == SP_K_BDCSR || id == SP_K_BDCOO;
        const uint32_t cap = 32768u / (uint32_t)sp_dtype_size(A->dtype);
        uint32_t vs = 2;
        while (vs < 64 && ((uint64_t)A->ncols + vs - 1) / vs > cap) vs *= 2;
        if (csr_coo && ((uint64_t)A->ncols + vs - 1) / vs <= cap) {
            Pt p{id, base};
            p.o.vert = vs;
            p.o.lds_slice = cap;
            v.push_back(p);
        }
    }
}

extern "C" int sp_tune_space(const sp_coo *A, int kernel_id, sp_tune_point *pts, int cap, int *count) {
    if (!A || !count) return SP_ERR_ARG;
    std::vector<Pt> v;
    for (int id = 0; id < SP_K_COUNT; id++)
        if (kernel_id < 0 || kernel_id == id) space_for(A, id, v);
    *count = (int)v.size();
    if (!pts) return SP_OK;
    if (cap < (int)v.size()) return SP_ERR_ARG;
    for (size_t i = 0; i < v.size(); i++) {
        memset(&pts[i], 0, sizeof(pts[i]));
        pts[i].kernel_id = v[i].id;
        pts[i].opts = v[i].o;
        pts[i].ms = -1.f;
        pts[i].status = 1;  // not run
    }
    return SP_OK;
}

// Builds run on host threads (plan creation is host work plus synchronous
// uploads, safe to overlap), up to `kBatch` plans at a time; the timed SpMVs
// (k = 1) or SpMMs with k right-hand sides then run one plan after another on
// the context stream, so timings never overlap.
static int tune_run(sp_ctx *ctx, const sp_coo *A, sp_tune_point *pts, int n, int iters, uint64_t flush_bytes,
                    int k) {
    if (!ctx || !A || !pts || n < 0 || iters <= 0 || k <= 0) return SP_ERR_ARG;
    (void)hipSetDevice(ctx->device);
    const size_t sv = sp_dtype_size(A->dtype);
    void *x = nullptr, *y = nullptr;
    if (hipMalloc(&x, (size_t)std::max<uint32_t>(A->ncols, 1) * k * sv) != hipSuccess) return SP_ERR_NOMEM;
    if (hipMalloc(&y, (size_t)std::max<uint32_t>(A->nrows, 1) * k * sv) != hipSuccess) { (void)hipFree(x); return SP_ERR_NOMEM; }
    {
        std::vector<char> hx((size_t)std::max<uint32_t>(A->ncols, 1) * k * sv);
        sp_dist d;
        d.kind = SP_DIST_UNIFORM;
        d.lo = -1;
        d.hi = 1;
        sp_gen_vector(A->dtype, (uint64_t)A->ncols * k, d, 12345, hx.data());
        (void)hipMemcpy(x, hx.data(), hx.size(), hipMemcpyHostToDevice);
    }
    constexpr int kBatch = 16;
    const int nthreads = (int)std::max(1u, std::min<unsigned>(kBatch, std::thread::hardware_concurrency()));
    std::vector<float> ms(iters);
    for (int b0 = 0; b0 < n; b0 += kBatch) {
        const int b1 = std::min(n, b0 + kBatch);
        std::vector<sp_plan *> plans(b1 - b0, nullptr);
        std::vector<int> rcs(b1 - b0, 0);
        std::atomic<int> next(b0);
        auto worker = [&]() {
            (void)hipSetDevice(ctx->device);
            for (int i; (i = next.fetch_add(1)) < b1;) {
                sp_tune_point &pt = pts[i];
                const auto t0 = std::chrono::steady_clock::now();
                rcs[i - b0] = sp_plan_create(ctx, pt.kernel_id, A, &pt.opts, &plans[i - b0]);
                const auto t1 = std::chrono::steady_clock::now();
                pt.build_ms = (float)std::chrono::duration<double, std::milli>(t1 - t0).count();
            }
        };
        std::vector<std::thread> pool;
        for (int t = 0; t < std::min(nthreads, b1 - b0); t++) pool.emplace_back(worker);
        for (auto &t : pool) t.join();
        for (int i = b0; i < b1; i++) {
            sp_tune_point &pt = pts[i];
            sp_plan *p = plans[i - b0];
            int rc = rcs[i - b0];
            if (rc) {
                pt.status = rc;
                pt.ms = -1.f;
                continue;
            }
            if (k == 1) {
                rc = sp_time_spmv(p, x, y, 1, flush_bytes, ms.data());  // warm-up
                if (!rc) rc = sp_time_spmv(p, x, y, iters, flush_bytes, ms.data());
            } else {
                rc = sp_time_spmm(p, k, x, k, y, k, 1, flush_bytes, ms.data());
                if (!rc) rc = sp_time_spmm(p, k, x, k, y, k, iters, flush_bytes, ms.data());
            }
            if (!rc) {
                std::vector<float> s(ms);
                std::sort(s.begin(), s.end());
                pt.ms = s[s.size() / 2];
                sp_plan_info info;
                sp_plan_get_info(p, &info);
                pt.algo_bytes = info.algo_bytes;
            } else {
                pt.ms = -1.f;
            }
            pt.status = rc;
        }
        for (sp_plan *p : plans)
            if (p) sp_plan_destroy(p);
    }
    (void)hipFree(x);
    (void)hipFree(y);
    return SP_OK;
}

extern "C" int sp_tune_run(sp_ctx *ctx, const sp_coo *A, sp_tune_point *pts, int n, int iters,
                           uint64_t flush_bytes) {
    return tune_run(ctx, A, pts, n, iters, flush_bytes, 1);
}

extern "C" int sp_tune_run_spmm(sp_ctx *ctx, const sp_coo *A, sp_tune_point *pts, int n, int iters,
                                uint64_t flush_bytes, int k) {
    return tune_run(ctx, A, pts, n, iters, flush_bytes, k);
}

extern "C" int sp_tune_best(const sp_tune_point *pts, int n, int kernel_id) {
    int best = -1;
    for (int i = 0; i < n; i++) {
        if (pts[i].status != 0 || pts[i].ms < 0) continue;
        if (kernel_id >= 0 && pts[i].kernel_id != kernel_id) continue;
        if (best < 0 || pts[i].ms < pts[best].ms) best = i;
    }
    return best;
}

extern "C" int sp_tune_db_append(const char *path, const char *tag, const sp_tune_point *pts, int n) {
    if (!path || !pts) return SP_ERR_ARG;
    FILE *f = fopen(path, "a");
    if (!f) return SP_ERR_IO;
    for (int i = 0; i < n; i++) {
        const sp_tune_point &p = pts[i];
        fprintf(f,
                "{\"matrix\": \"%s\", \"kernel\": %d, \"name\": \"%s\", \"algo\": %d, \"lanes\": %u, \"ipt\": %u, "
                "\"vert\": %u, \"r\": %u, \"c\": %u, \"nparts\": %u, \"nsub\": %u, \"ms\": %.6f, "
                "\"build_ms\": %.3f, \"status\": %d}\n",
                tag ? tag : "", p.kernel_id, sp_kernel_name(p.kernel_id), p.opts.algo, p.opts.lanes_per_row,
                p.opts.items_per_lane, p.opts.vert, p.opts.row_block_size, p.opts.col_block_size, p.opts.nparts,
                p.opts.nsub, p.ms, p.build_ms, p.status);
    }
    fclose(f);
    return SP_OK;
}
