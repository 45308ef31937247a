// Learned variant selector: standardised features -> MLP (HIP kernels) ->
// per-kernel-id scores, lowest wins; low-confidence picks fall back to an
// exhaustive search over the best-scored candidates.  No counterpart in the
// reference.
#include <math.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <numeric>
#include <vector>

#include "plan.hpp"

struct sp_selector {
    sp_ctx *ctx = nullptr;
    sp_mlp *mlp = nullptr;
    int sizes[4] = {0};
    float margin = 0.f;
    std::vector<float> mean, stdv, params;
};

static const char kMagic[8] = {'S', 'P', 'S', 'E', 'L', '0', '0', '1'};

extern "C" int sp_selector_create(sp_ctx *ctx, const int *sizes, const float *params, size_t nparams,
                                  const float *feat_mean, const float *feat_std, float margin, sp_selector **out) {
    if (!ctx || !sizes || !params || !feat_mean || !feat_std || !out) return SP_ERR_ARG;
    if (sizes[0] != SP_NFEAT || sizes[3] != SP_K_COUNT || sizes[1] <= 0 || sizes[2] <= 0) return SP_ERR_ARG;
    sp_selector *s = new sp_selector();
    s->ctx = ctx;
    memcpy(s->sizes, sizes, sizeof(s->sizes));
    s->margin = margin;
    s->mean.assign(feat_mean, feat_mean + SP_NFEAT);
    s->stdv.assign(feat_std, feat_std + SP_NFEAT);
    for (float &v : s->stdv)
        if (!(v > 1e-6f)) v = 1e-6f;
    int rc = sp_mlp_create(ctx, sizes[0], sizes[1], sizes[2], sizes[3], 0, &s->mlp);
    if (!rc && sp_mlp_num_params(s->mlp) != nparams) rc = SP_ERR_ARG;
    if (!rc) rc = sp_mlp_set_params(s->mlp, params, nparams);
    if (rc) {
        sp_selector_destroy(s);
        return rc;
    }
    s->params.assign(params, params + nparams);
    *out = s;
    return SP_OK;
}

extern "C" int sp_selector_destroy(sp_selector *s) {
    if (!s) return SP_ERR_ARG;
    if (s->mlp) sp_mlp_destroy(s->mlp);
    delete s;
    return SP_OK;
}

// File: 8-byte magic, int32 sizes[4], float margin, float mean[64],
// float std[64], float params[n] (little-endian, as written by this host).
extern "C" int sp_selector_save(const sp_selector *s, const char *path) {
    if (!s || !path) return SP_ERR_ARG;
    FILE *f = fopen(path, "wb");
    if (!f) return SP_ERR_IO;
    bool ok = fwrite(kMagic, 1, 8, f) == 8 && fwrite(s->sizes, 4, 4, f) == 4 && fwrite(&s->margin, 4, 1, f) == 1 &&
              fwrite(s->mean.data(), 4, SP_NFEAT, f) == SP_NFEAT && fwrite(s->stdv.data(), 4, SP_NFEAT, f) == SP_NFEAT &&
              fwrite(s->params.data(), 4, s->params.size(), f) == s->params.size();
    ok = (fclose(f) == 0) && ok;
    return ok ? SP_OK : SP_ERR_IO;
}

extern "C" int sp_selector_load(sp_ctx *ctx, const char *path, sp_selector **out) {
    if (!ctx || !path || !out) return SP_ERR_ARG;
    FILE *f = fopen(path, "rb");
    if (!f) return SP_ERR_IO;
    char magic[8];
    int sizes[4];
    float margin;
    std::vector<float> mean(SP_NFEAT), stdv(SP_NFEAT);
    bool ok = fread(magic, 1, 8, f) == 8 && !memcmp(magic, kMagic, 8) && fread(sizes, 4, 4, f) == 4 &&
              fread(&margin, 4, 1, f) == 1 && fread(mean.data(), 4, SP_NFEAT, f) == SP_NFEAT &&
              fread(stdv.data(), 4, SP_NFEAT, f) == SP_NFEAT;
    std::vector<float> params;
    if (ok && sizes[0] == SP_NFEAT && sizes[3] == SP_K_COUNT && sizes[1] > 0 && sizes[2] > 0 && sizes[1] <= 4096 &&
        sizes[2] <= 4096) {
        const size_t n = (size_t)sizes[1] * (sizes[0] + 1) + (size_t)sizes[2] * (sizes[1] + 1) +
                         (size_t)sizes[3] * (sizes[2] + 1);
        params.resize(n);
        ok = fread(params.data(), 4, n, f) == n;
    } else {
        ok = false;
    }
    fclose(f);
    if (!ok) return SP_ERR_IO;
    return sp_selector_create(ctx, sizes, params.data(), params.size(), mean.data(), stdv.data(), margin, out);
}

extern "C" int sp_select(sp_selector *s, const sp_coo *A, int fallback_k, int iters, uint64_t flush_bytes,
                         sp_tune_point *choice, float *scores, int *fell_back) {
    if (!s || !A || !choice) return SP_ERR_ARG;
    float feat[SP_NFEAT], pred[SP_K_COUNT];
    int rc = sp_features(A, feat);
    if (rc) return rc;
    for (int i = 0; i < SP_NFEAT; i++) feat[i] = (feat[i] - s->mean[i]) / s->stdv[i];
    if ((rc = sp_mlp_predict(s->mlp, feat, 1, pred))) return rc;
    if (scores) memcpy(scores, pred, sizeof(pred));
    int order[SP_K_COUNT];
    std::iota(order, order + SP_K_COUNT, 0);
    std::stable_sort(order, order + SP_K_COUNT, [&](int a, int b) { return pred[a] < pred[b]; });
    const bool confident = pred[order[1]] - pred[order[0]] >= s->margin;
    memset(choice, 0, sizeof(*choice));
    choice->kernel_id = order[0];
    sp_opts_default(&choice->opts);
    choice->ms = -1.f;
    choice->status = 1;
    if (fell_back) *fell_back = 0;
    if (confident || fallback_k <= 0) return SP_OK;
    // exhaustive search over every point of the fallback_k best-scored ids
    std::vector<sp_tune_point> pts;
    for (int r = 0; r < std::min(fallback_k, (int)SP_K_COUNT); r++) {
        int cnt = 0;
        if ((rc = sp_tune_space(A, order[r], nullptr, 0, &cnt))) return rc;
        const size_t at = pts.size();
        pts.resize(at + cnt);
        if ((rc = sp_tune_space(A, order[r], pts.data() + at, cnt, &cnt))) return rc;
    }
    if ((rc = sp_tune_run(s->ctx, A, pts.data(), (int)pts.size(), iters, flush_bytes))) return rc;
    const int best = sp_tune_best(pts.data(), (int)pts.size(), -1);
    if (best >= 0) *choice = pts[best];
    if (fell_back) *fell_back = 1;
    return SP_OK;
}
