// Batched SpMV / SpMM: many independent 1D plans run as one hipGraph launch.
//
// The problems of all plans are concatenated into launch groups keyed by
// (kernel kind, tile shape, right-hand-side chunk width); each group is one
// persistent launch whose workgroups sweep every problem's work units, with
// x / y taken from a per-call device pointer table.  Groups are independent
// (each writes its own problems' columns), so they run on up to four
// parallel branches.  The groups and their carry fix-ups are captured once
// into a hipGraph and replayed.  There is
// no counterpart in the reference (which runs one matrix per launch).
#include <string.h>

#include <algorithm>
#include <map>
#include <tuple>
#include <vector>

#include "plan.hpp"

template <typename T>
int run_tables_t(const Work &w, sp_ctx *ctx, const uintptr_t *xtab, const uintptr_t *ytab, int kc);
extern template int run_tables_t<int8_t>(const Work &, sp_ctx *, const uintptr_t *, const uintptr_t *, int);
extern template int run_tables_t<int16_t>(const Work &, sp_ctx *, const uintptr_t *, const uintptr_t *, int);
extern template int run_tables_t<int32_t>(const Work &, sp_ctx *, const uintptr_t *, const uintptr_t *, int);
extern template int run_tables_t<int64_t>(const Work &, sp_ctx *, const uintptr_t *, const uintptr_t *, int);
extern template int run_tables_t<float>(const Work &, sp_ctx *, const uintptr_t *, const uintptr_t *, int);
extern template int run_tables_t<double>(const Work &, sp_ctx *, const uintptr_t *, const uintptr_t *, int);

namespace {

struct Slot {        // one (problem, RHS chunk, sub-problem) of a group
    uint32_t prob;   // batch problem index
    int64_t xel, yel; // element offsets from X[prob] / Y[prob]
};

struct Group {
    Work w;
    int kc = 1;
    std::vector<Slot> slots;
    uint32_t tab_off = 0;  // first table entry (x at tab_off, y at tab_off + nslots)
};

int run_group(const Group &g, sp_ctx *ctx, const uintptr_t *d_tab) {
    const uintptr_t *xt = d_tab + g.tab_off, *yt = d_tab + g.tab_off + g.slots.size();
    switch (g.w.dtype) {
    case SP_I8: return run_tables_t<int8_t>(g.w, ctx, xt, yt, g.kc);
    case SP_I16: return run_tables_t<int16_t>(g.w, ctx, xt, yt, g.kc);
    case SP_I32: return run_tables_t<int32_t>(g.w, ctx, xt, yt, g.kc);
    case SP_I64: return run_tables_t<int64_t>(g.w, ctx, xt, yt, g.kc);
    case SP_F32: return run_tables_t<float>(g.w, ctx, xt, yt, g.kc);
    case SP_F64: return run_tables_t<double>(g.w, ctx, xt, yt, g.kc);
    default: return SP_ERR_ARG;
    }
}

}  // namespace

struct sp_batch {
    sp_ctx *ctx = nullptr;
    int dtype = 0, flags = 0;
    uint32_t nprob = 0;
    std::vector<Group> groups;
    uint32_t ntab = 0;
    uintptr_t *d_tab = nullptr;
    uintptr_t *h_tab = nullptr;  // pinned
    hipEvent_t tab_done = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    hipStream_t captured_on = nullptr;
    uint64_t algo_bytes = 0;
    std::vector<void *> allocs;
    // independent groups run on parallel branches (streams forked from and
    // joined back into the context stream), so one group's tail overlaps the
    // others' work; branch[i] runs the groups with group_branch == i
    std::vector<hipStream_t> branch;
    std::vector<hipEvent_t> join;
    hipEvent_t fork = nullptr;
    std::vector<int> group_branch;
};

static int balloc(sp_batch *b, size_t bytes, void **out) {
    if (hipMalloc(out, bytes ? bytes : 16) != hipSuccess) return SP_ERR_NOMEM;
    b->allocs.push_back(*out);
    return SP_OK;
}

template <typename V> static int bupload(sp_batch *b, const std::vector<V> &v, V **out) {
    int rc = balloc(b, v.size() * sizeof(V), (void **)out);
    if (rc) return rc;
    if (!v.empty()) SP_HIP_TRY(hipMemcpy(*out, v.data(), v.size() * sizeof(V), hipMemcpyHostToDevice));
    return SP_OK;
}

static void batch_free(sp_batch *b) {
    if (!b) return;
    for (hipStream_t st : b->branch) (void)hipStreamDestroy(st);
    for (hipEvent_t ev : b->join) (void)hipEventDestroy(ev);
    if (b->fork) (void)hipEventDestroy(b->fork);
    if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    if (b->tab_done) (void)hipEventDestroy(b->tab_done);
    if (b->h_tab) (void)hipHostFree(b->h_tab);
    for (void *d : b->allocs) (void)hipFree(d);
    delete b;
}

extern "C" int sp_batch_create(sp_ctx *ctx, int nprob, sp_plan *const *plans, const int *k, int flags,
                               sp_batch **out) {
    if (!ctx || nprob <= 0 || !plans || !k || !out) return SP_ERR_ARG;
    const int dt = plans[0]->dtype;
    for (int i = 0; i < nprob; i++) {
        if (!plans[i] || plans[i]->ctx != ctx || plans[i]->dtype != dt || k[i] <= 0) return SP_ERR_ARG;
        if (plans[i]->tiled) return SP_ERR_UNSUPPORTED;
    }
    (void)hipSetDevice(ctx->device);
    sp_batch *b = new sp_batch();
    b->ctx = ctx;
    b->dtype = dt;
    b->flags = flags;
    b->nprob = (uint32_t)nprob;
    // group key: kind, ipt, lanes, R, C, chunk width
    std::map<std::tuple<int, uint32_t, uint32_t, uint32_t, uint32_t, int>, size_t> gidx;
    struct Acc {
        std::vector<spd::Prob> probs;
        std::vector<uint32_t> unit_pid, c0, c1, c2, zpid, zrow;
    };
    std::vector<Acc> acc;
    for (int i = 0; i < nprob; i++) {
        const sp_plan *p = plans[i];
        const Work &w = p->work;
        b->algo_bytes += p->algo_bytes + (uint64_t)(k[i] - 1) * ((uint64_t)p->nrows + p->ncols) * sp_dtype_size(dt);
        const int kmax = max_rhs_chunk(w);
        for (int done = 0; done < k[i];) {
            const int left = k[i] - done;
            int kc = left >= 32 ? 32 : left >= 16 ? 16 : left >= 8 ? 8 : left >= 4 ? 4 : left >= 2 ? 2 : 1;
            while (kc > kmax) kc >>= 1;
            const auto key = std::make_tuple(w.kind, w.ipt, w.lanes, w.R, w.C, kc);
            auto it = gidx.find(key);
            size_t gi;
            if (it == gidx.end()) {
                gi = b->groups.size();
                gidx[key] = gi;
                b->groups.emplace_back();
                acc.emplace_back();
                Group &g = b->groups.back();
                g.kc = kc;
                g.w.kind = w.kind;
                g.w.dtype = dt;
                g.w.ipt = w.ipt;
                g.w.lanes = w.lanes;
                g.w.R = w.R;
                g.w.C = w.C;
                g.w.chunk = w.chunk;
            } else {
                gi = it->second;
            }
            Group &g = b->groups[gi];
            g.w.chunk = g.w.chunk && w.chunk;
            Acc &a = acc[gi];
            const uint32_t pid0 = (uint32_t)a.probs.size();
            for (const spd::Prob &P0 : p->h_probs) {
                spd::Prob P = P0;
                P.ldx = k[i];
                P.ldy = k[i];
                a.probs.push_back(P);
                g.slots.push_back({(uint32_t)i, P0.xoff * k[i] + done, P0.yoff * k[i] + done});
            }
            for (uint32_t u : p->h_unit_pid) a.unit_pid.push_back(u + pid0);
            a.c0.insert(a.c0.end(), p->h_c0.begin(), p->h_c0.end());
            a.c1.insert(a.c1.end(), p->h_c1.begin(), p->h_c1.end());
            a.c2.insert(a.c2.end(), p->h_c2.begin(), p->h_c2.end());
            g.w.hcol_max = std::max(g.w.hcol_max, w.hcol_max);
            for (size_t z = 0; z < p->h_zrow.size(); z++) {
                a.zpid.push_back(p->h_zpid[z] + pid0);
                a.zrow.push_back(p->h_zrow[z]);
            }
            done += kc;
        }
    }
    int rc = SP_OK;
    uint32_t tab = 0;
    for (size_t gi = 0; gi < b->groups.size() && !rc; gi++) {
        Group &g = b->groups[gi];
        Acc &a = acc[gi];
        Work &w = g.w;
        w.nprob = (uint32_t)a.probs.size();
        w.nunits = (uint32_t)a.unit_pid.size();
        w.nzero = (uint32_t)a.zrow.size();
        // persistent grid: enough workgroups to fill every CU several times over
        w.grid_cap = (uint32_t)std::max(1, ctx->num_cus) * (w.kind == WK_CSR_HOT ? 1 : 8);
        if (w.kind == WK_CSR_HOT) {
            // every slot stages its own x (a plan may appear in several slots)
            for (spd::Prob &P : a.probs) {
                void *xh = nullptr;
                if ((rc = balloc(b, ((size_t)P.nhcol * sp_dtype_size(dt) + 15) / 16 * 16, &xh))) break;
                P.xh = xh;
            }
            if (rc) break;
        }
        if ((rc = bupload(b, a.probs, &w.d_prob))) break;
        if ((rc = bupload(b, a.unit_pid, &w.d_unit_pid))) break;
        if ((rc = bupload(b, a.c0, &w.d_c0))) break;
        if (!a.c1.empty() && (rc = bupload(b, a.c1, &w.d_c1))) break;
        if (!a.c2.empty() && (rc = bupload(b, a.c2, &w.d_c2))) break;
        if (w.nzero && ((rc = bupload(b, a.zpid, &w.d_zpid)) || (rc = bupload(b, a.zrow, &w.d_zrow)))) break;
        const bool carries = w.kind == WK_CSR_MERGE || w.kind == WK_COO_SEG || w.kind == WK_BCOO_SEG || w.kind == WK_CSR_HOT ||
                             w.kind == WK_BCSR_SEG;
        if (carries) {
            const uint32_t S = (w.kind == WK_BCOO_SEG || w.kind == WK_BCSR_SEG) ? w.R : 1;
            const size_t n = (size_t)std::max<uint32_t>(w.nunits, 1) * S;
            const size_t asz = (dt == SP_I64 || dt == SP_F64) ? 8 : 4;
            w.carry_k = (uint32_t)g.kc;
            if ((rc = balloc(b, n * 4, (void **)&w.d_carry_row))) break;
            if ((rc = balloc(b, n * g.kc * asz, &w.d_carry_val))) break;
        }
        if (w.kind == WK_CSR_MERGE && !w.nzero && w.nunits) {
            // the rows crossing tiles as runs (unit t carries the row at its end coordinate),
            // as a single plan lists them: each run's carries are added with independent loads
            // instead of a walk over the units that stops at the run's end
            std::vector<spd::Run> runs;
            for (uint32_t t = 0; t < w.nunits; t++) {
                const uint32_t pid = a.unit_pid[t], r = a.c0[(size_t)t + pid + 1];
                const spd::Prob &P = a.probs[pid];
                if (r >= P.nrows || r >= P.nrows_elem) continue;
                if (!runs.empty()) {
                    spd::Run &q = runs.back();
                    if (q.pid == pid && q.row == r && q.t0 + q.n == t) { q.n++; continue; }
                }
                runs.push_back(spd::Run{pid, t, 1, r, P.yoff + (int64_t)r, P.ldy, 0});
            }
            w.nruns = (uint32_t)runs.size();
            if ((rc = bupload(b, runs, &w.d_runs))) break;
        }
        g.tab_off = tab;
        tab += 2 * (uint32_t)g.slots.size();
    }
    if (!rc) {
        b->ntab = tab;
        if ((rc = balloc(b, (size_t)tab * sizeof(uintptr_t), (void **)&b->d_tab)) == SP_OK) {
            if (hipHostMalloc((void **)&b->h_tab, (size_t)std::max<uint32_t>(tab, 1) * sizeof(uintptr_t)) != hipSuccess)
                rc = SP_ERR_NOMEM;
            else if (hipEventCreateWithFlags(&b->tab_done, hipEventDisableTiming) != hipSuccess)
                rc = SP_ERR_HIP;
        }
    }
    if (!rc && b->groups.size() > 1) {
        // up to kBranches branches (default-priority streams), groups assigned
        // largest first to the branch with the least work so far
        constexpr size_t kBranches = 4;
        const size_t nbr = std::min(kBranches, b->groups.size());
        std::vector<size_t> order(b->groups.size());
        for (size_t gi = 0; gi < order.size(); gi++) order[gi] = gi;
        auto work = [&](size_t gi) { return (uint64_t)b->groups[gi].w.nunits * (uint64_t)b->groups[gi].kc; };
        std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t c) { return work(a) > work(c); });
        std::vector<uint64_t> load(nbr, 0);
        b->group_branch.assign(b->groups.size(), 0);
        for (size_t gi : order) {
            const size_t k2 = (size_t)(std::min_element(load.begin(), load.end()) - load.begin());
            b->group_branch[gi] = (int)k2;
            load[k2] += work(gi);
        }
        b->branch.resize(nbr, nullptr);
        b->join.resize(nbr, nullptr);
        for (size_t i = 0; i < nbr && !rc; i++) {
            if (hipStreamCreateWithFlags(&b->branch[i], hipStreamNonBlocking) != hipSuccess ||
                hipEventCreateWithFlags(&b->join[i], hipEventDisableTiming) != hipSuccess)
                rc = SP_ERR_HIP;
        }
        if (!rc && hipEventCreateWithFlags(&b->fork, hipEventDisableTiming) != hipSuccess) rc = SP_ERR_HIP;
    }
    if (rc) { batch_free(b); return rc; }
    *out = b;
    return SP_OK;
}

extern "C" int sp_batch_destroy(sp_batch *b) {
    if (!b) return SP_ERR_ARG;
    (void)hipSetDevice(b->ctx->device);
    (void)hipStreamSynchronize(b->ctx->stream);
    batch_free(b);
    return SP_OK;
}

// Fill and upload the pointer table (the previous upload must have left the
// pinned buffer before it is rewritten).
static int batch_tables(sp_batch *b, const void *const *X, void *const *Y) {
    SP_HIP_TRY(hipEventSynchronize(b->tab_done));
    const size_t sv = sp_dtype_size(b->dtype);
    for (const Group &g : b->groups) {
        const size_t ns = g.slots.size();
        for (size_t s = 0; s < ns; s++) {
            const Slot &sl = g.slots[s];
            if (!X[sl.prob] || !Y[sl.prob]) return SP_ERR_ARG;
            b->h_tab[g.tab_off + s] = (uintptr_t)X[sl.prob] + (uintptr_t)(sl.xel * (int64_t)sv);
            b->h_tab[g.tab_off + ns + s] = (uintptr_t)Y[sl.prob] + (uintptr_t)(sl.yel * (int64_t)sv);
        }
    }
    if (b->ntab)
        SP_HIP_TRY(hipMemcpyAsync(b->d_tab, b->h_tab, b->ntab * sizeof(uintptr_t), hipMemcpyHostToDevice, b->ctx->stream));
    SP_HIP_TRY(hipEventRecord(b->tab_done, b->ctx->stream));
    return SP_OK;
}

// Every group, on the context stream or forked over the branches and joined
// back into it (also while the context stream is being captured).
static int batch_enqueue(sp_batch *b) {
    sp_ctx *c = b->ctx;
    if (b->branch.empty()) {
        for (const Group &g : b->groups) {
            int rc = run_group(g, c, b->d_tab);
            if (rc) return rc;
        }
        return SP_OK;
    }
    SP_HIP_TRY(hipEventRecord(b->fork, c->stream));
    for (hipStream_t st : b->branch) SP_HIP_TRY(hipStreamWaitEvent(st, b->fork, 0));
    int rc = SP_OK;
    for (size_t gi = 0; gi < b->groups.size() && !rc; gi++) {
        sp_ctx cb = *c;  // the same context, launching on the group's branch
        cb.stream = b->branch[b->group_branch[gi]];
        cb.mark_mid = false;
        rc = run_group(b->groups[gi], &cb, b->d_tab);
    }
    // join every branch, also after a failed launch, so a capture can end
    for (size_t i = 0; i < b->branch.size(); i++) {
        if (hipEventRecord(b->join[i], b->branch[i]) != hipSuccess ||
            hipStreamWaitEvent(c->stream, b->join[i], 0) != hipSuccess)
            rc = rc ? rc : SP_ERR_HIP;
    }
    return rc;
}

static int batch_launch(sp_batch *b) {
    sp_ctx *c = b->ctx;
    if (!(b->flags & SP_BATCH_GRAPH)) return batch_enqueue(b);
    if (!b->gexec || b->captured_on != c->stream) {
        if (b->gexec) { (void)hipGraphExecDestroy(b->gexec); b->gexec = nullptr; }
        if (b->graph) { (void)hipGraphDestroy(b->graph); b->graph = nullptr; }
        SP_HIP_TRY(hipStreamBeginCapture(c->stream, hipStreamCaptureModeRelaxed));
        const int rc = batch_enqueue(b);
        hipGraph_t graph = nullptr;
        const hipError_t e = hipStreamEndCapture(c->stream, &graph);
        if (rc) { if (graph) (void)hipGraphDestroy(graph); return rc; }
        if (e != hipSuccess) return SP_ERR_HIP;
        b->graph = graph;
        SP_HIP_TRY(hipGraphInstantiate(&b->gexec, graph, nullptr, nullptr, 0));
        b->captured_on = c->stream;
    }
    SP_HIP_TRY(hipGraphLaunch(b->gexec, c->stream));
    return SP_OK;
}

extern "C" int sp_batch_run(sp_batch *b, const void *const *X, void *const *Y) {
    if (!b || !X || !Y) return SP_ERR_ARG;
    (void)hipSetDevice(b->ctx->device);
    int rc = batch_tables(b, X, Y);
    if (rc) return rc;
    return batch_launch(b);
}

extern "C" int sp_batch_time(sp_batch *b, const void *const *X, void *const *Y, int iters, uint64_t flush_bytes,
                             float *ms) {
    if (!b || !X || !Y || iters <= 0 || !ms) return SP_ERR_ARG;
    sp_ctx *c = b->ctx;
    (void)hipSetDevice(c->device);
    int rc = batch_tables(b, X, Y);
    if (rc) return rc;
    if (flush_bytes && (c->flush_bytes < flush_bytes)) {
        if (c->flush_buf) (void)hipFree(c->flush_buf);
        c->flush_buf = nullptr;
        c->flush_bytes = 0;
        if (hipMalloc(&c->flush_buf, flush_bytes) != hipSuccess) return SP_ERR_NOMEM;
        c->flush_bytes = flush_bytes;
    }
    for (int it = 0; it < iters; it++) {
        if (flush_bytes && (rc = launch_flush(c->stream, c->flush_buf, flush_bytes, 0x51ED2701u * (uint32_t)(it + 1))))
            return rc;
        SP_HIP_TRY(hipEventRecord(c->ev0, c->stream));
        if ((rc = batch_launch(b))) return rc;
        SP_HIP_TRY(hipEventRecord(c->ev1, c->stream));
        SP_HIP_TRY(hipEventSynchronize(c->ev1));
        SP_HIP_TRY(hipEventElapsedTime(&ms[it], c->ev0, c->ev1));
    }
    return SP_OK;
}

extern "C" int sp_batch_info(const sp_batch *b, uint32_t *ngroups, uint64_t *algo_bytes) {
    if (!b) return SP_ERR_ARG;
    if (ngroups) *ngroups = (uint32_t)b->groups.size();
    if (algo_bytes) *algo_bytes = b->algo_bytes;
    return SP_OK;
}

extern "C" int sp_batch_run_once(sp_ctx *ctx, int nprob, sp_plan *const *plans, const void *const *X,
                                 void *const *Y, const int *k) {
    sp_batch *b = nullptr;
    int rc = sp_batch_create(ctx, nprob, plans, k, 0, &b);
    if (rc) return rc;
    rc = sp_batch_run(b, X, Y);
    const int rc2 = sp_batch_destroy(b);
    return rc ? rc : rc2;
}

// The unbatched counterpart of sp_batch_time: one sp_spmm call per problem,
// back to back on the context stream, timed as a whole with HIP events.
extern "C" int sp_time_plans(sp_ctx *ctx, int nprob, sp_plan *const *plans, const void *const *X, void *const *Y,
                             const int *k, int iters, uint64_t flush_bytes, float *ms) {
    if (!ctx || nprob <= 0 || !plans || !X || !Y || !k || iters <= 0 || !ms) return SP_ERR_ARG;
    for (int i = 0; i < nprob; i++)
        if (!plans[i] || plans[i]->ctx != ctx || k[i] <= 0) return SP_ERR_ARG;
    (void)hipSetDevice(ctx->device);
    if (flush_bytes && ctx->flush_bytes < flush_bytes) {
        if (ctx->flush_buf) (void)hipFree(ctx->flush_buf);
        ctx->flush_buf = nullptr;
        ctx->flush_bytes = 0;
        if (hipMalloc(&ctx->flush_buf, flush_bytes) != hipSuccess) return SP_ERR_NOMEM;
        ctx->flush_bytes = flush_bytes;
    }
    int rc;
    for (int it = 0; it < iters; it++) {
        if (flush_bytes && (rc = launch_flush(ctx->stream, ctx->flush_buf, flush_bytes, 0x51ED2701u * (uint32_t)(it + 1))))
            return rc;
        SP_HIP_TRY(hipEventRecord(ctx->ev0, ctx->stream));
        for (int i = 0; i < nprob; i++)
            if ((rc = sp_spmm(plans[i], k[i], X[i], k[i], Y[i], k[i]))) return rc;
        SP_HIP_TRY(hipEventRecord(ctx->ev1, ctx->stream));
        SP_HIP_TRY(hipEventSynchronize(ctx->ev1));
        SP_HIP_TRY(hipEventElapsedTime(&ms[it], ctx->ev0, ctx->ev1));
    }
    return SP_OK;
}
