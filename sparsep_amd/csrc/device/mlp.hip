// Variant-selector MLP: in -> h1 -> h2 -> out with ReLU between layers.
// Weights are used in BF16 (round-to-nearest-even copies of FP32 master
// weights), activations and accumulation are FP32.  Training runs on the
// GPU with these kernels: dense forward, masked-MSE loss gradient, ReLU
// backward, weight / input gradients, and Adam on the FP32 masters.  All
// reductions have a fixed order (no atomics), so training is reproducible.
// No counterpart in the reference.
#include <math.h>
#include <string.h>

#include <vector>

#include "plan.hpp"

namespace {

constexpr int TB = 16;  // output tile edge of the dense kernels

__host__ __device__ inline uint16_t f2bf(float f) {
    uint32_t u = __builtin_bit_cast(uint32_t, f);
    if ((u & 0x7F800000u) == 0x7F800000u && (u & 0x007FFFFFu)) return (uint16_t)((u >> 16) | 0x40);  // NaN
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__host__ __device__ inline float bf2f(uint16_t b) { return __builtin_bit_cast(float, (uint32_t)b << 16); }

// out[b][o] = act(sum_i in[b][i] * W[o][i] + bias[o]); pre keeps the sum.
__global__ void dense_fwd(const float *__restrict__ in, const uint16_t *__restrict__ W,
                          const float *__restrict__ bias, float *__restrict__ pre, float *__restrict__ out, int B,
                          int I, int O, int relu) {
    __shared__ float s_in[TB][TB + 1];
    __shared__ float s_w[TB][TB + 1];
    const int tx = threadIdx.x % TB, ty = threadIdx.x / TB;
    const int b = blockIdx.y * TB + ty, o = blockIdx.x * TB + tx;
    float acc = 0.f;
    for (int i0 = 0; i0 < I; i0 += TB) {
        const int bi = blockIdx.y * TB + ty, ii = i0 + tx;
        s_in[ty][tx] = (bi < B && ii < I) ? in[(size_t)bi * I + ii] : 0.f;
        const int wo = blockIdx.x * TB + ty;
        s_w[ty][tx] = (wo < O && ii < I) ? bf2f(W[(size_t)wo * I + ii]) : 0.f;
        __syncthreads();
#pragma unroll
        for (int k = 0; k < TB; k++) acc = __builtin_fmaf(s_in[ty][k], s_w[tx][k], acc);
        __syncthreads();
    }
    if (b < B && o < O) {
        const float v = acc + bias[o];
        if (pre) pre[(size_t)b * O + o] = v;
        out[(size_t)b * O + o] = (relu && v < 0.f) ? 0.f : v;
    }
}

// dY = 2 M^2 (Y - T) / count; per-block partial loss sums (fixed order).
__global__ void mse_grad(const float *__restrict__ Y, const float *__restrict__ T, const float *__restrict__ M,
                         float *__restrict__ dY, float *__restrict__ partial, int n, float inv_count) {
    __shared__ float s[256];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    float l = 0.f;
    if (i < n) {
        const float m = M ? M[i] : 1.f;
        const float d = (Y[i] - T[i]) * m;
        dY[i] = 2.f * d * m * inv_count;
        l = d * d;
    }
    s[threadIdx.x] = l;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = s[0];
}

__global__ void relu_bwd(float *__restrict__ d, const float *__restrict__ pre, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && pre[i] <= 0.f) d[i] = 0.f;
}

// dW[o][i] = sum_b dOut[b][o] in[b][i]; db[o] = sum_b dOut[b][o] (column i == I).
__global__ void dense_bwd_w(const float *__restrict__ dOut, const float *__restrict__ in, float *__restrict__ dW,
                            float *__restrict__ db, int B, int I, int O) {
    const int o = blockIdx.y, i = blockIdx.x * blockDim.x + threadIdx.x;
    if (o >= O || i > I) return;
    float acc = 0.f;
    if (i < I) {
        for (int b = 0; b < B; b++) acc = __builtin_fmaf(dOut[(size_t)b * O + o], in[(size_t)b * I + i], acc);
        dW[(size_t)o * I + i] = acc;
    } else {
        for (int b = 0; b < B; b++) acc += dOut[(size_t)b * O + o];
        db[o] = acc;
    }
}

// dIn[b][i] = sum_o dOut[b][o] W[o][i] (the BF16 weights used forward).
__global__ void dense_bwd_x(const float *__restrict__ dOut, const uint16_t *__restrict__ W, float *__restrict__ dIn,
                            int B, int I, int O) {
    const int b = blockIdx.y, i = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= B || i >= I) return;
    float acc = 0.f;
    for (int o = 0; o < O; o++) acc = __builtin_fmaf(dOut[(size_t)b * O + o], bf2f(W[(size_t)o * I + i]), acc);
    dIn[(size_t)b * I + i] = acc;
}

__global__ void adam_step(float *__restrict__ p, const float *__restrict__ g, float *__restrict__ m,
                          float *__restrict__ v, uint16_t *__restrict__ pb, int n, float lr, float b1, float b2,
                          float eps, float c1, float c2) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float upd = lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    const float np = p[i] - upd;
    p[i] = np;
    if (pb) pb[i] = f2bf(np);
}

__global__ void to_bf16(const float *__restrict__ p, uint16_t *__restrict__ pb, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) pb[i] = f2bf(p[i]);
}

}  // namespace

struct sp_mlp {
    sp_ctx *ctx = nullptr;
    int in = 0, h1 = 0, h2 = 0, out = 0;
    int np = 0;             // parameter count
    int off[6] = {0};       // W1 b1 W2 b2 W3 b3 offsets
    float *p = nullptr, *g = nullptr, *m = nullptr, *v = nullptr;  // FP32 masters, grads, Adam moments
    uint16_t *pb = nullptr;  // BF16 copy (weights used forward)
    int step = 0;
    // activations for a batch
    int cap = 0;
    float *x = nullptr, *a1 = nullptr, *z1 = nullptr, *a2 = nullptr, *z2 = nullptr, *y = nullptr;
    float *t = nullptr, *msk = nullptr, *d3 = nullptr, *d2 = nullptr, *d1 = nullptr, *part = nullptr;
};

static int mlp_alloc(float **p, size_t n) { return hipMalloc((void **)p, (n ? n : 1) * sizeof(float)) == hipSuccess ? SP_OK : SP_ERR_NOMEM; }

static int mlp_reserve(sp_mlp *m, int B) {
    if (B <= m->cap) return SP_OK;
    float **bufs[] = {&m->x, &m->a1, &m->z1, &m->a2, &m->z2, &m->y, &m->t, &m->msk, &m->d3, &m->d2, &m->d1, &m->part};
    for (float **b : bufs)
        if (*b) { (void)hipFree(*b); *b = nullptr; }
    const size_t n[] = {(size_t)B * m->in, (size_t)B * m->h1, (size_t)B * m->h1, (size_t)B * m->h2, (size_t)B * m->h2,
                        (size_t)B * m->out, (size_t)B * m->out, (size_t)B * m->out, (size_t)B * m->out,
                        (size_t)B * m->h2, (size_t)B * m->h1, (size_t)B * m->out / 256 + 2};
    for (int i = 0; i < 12; i++)
        if (mlp_alloc(bufs[i], n[i])) return SP_ERR_NOMEM;
    m->cap = B;
    return SP_OK;
}

extern "C" int sp_mlp_create(sp_ctx *ctx, int in, int h1, int h2, int out, uint64_t seed, sp_mlp **res) {
    if (!ctx || in <= 0 || h1 <= 0 || h2 <= 0 || out <= 0 || !res) return SP_ERR_ARG;
    (void)hipSetDevice(ctx->device);
    sp_mlp *m = new sp_mlp();
    m->ctx = ctx;
    m->in = in; m->h1 = h1; m->h2 = h2; m->out = out;
    const int sizes[6] = {h1 * in, h1, h2 * h1, h2, out * h2, out};
    for (int i = 0; i < 6; i++) { m->off[i] = m->np; m->np += sizes[i]; }
    if (mlp_alloc(&m->p, m->np) || mlp_alloc(&m->g, m->np) || mlp_alloc(&m->m, m->np) || mlp_alloc(&m->v, m->np) ||
        hipMalloc((void **)&m->pb, (size_t)m->np * 2) != hipSuccess) {
        sp_mlp_destroy(m);
        return SP_ERR_NOMEM;
    }
    // He-uniform initialisation from the project PRNG, zero biases
    std::vector<float> h(m->np, 0.f);
    sp_rng r;
    sp_rng_seed(&r, seed);
    const int fan[3] = {in, h1, h2};
    for (int l = 0; l < 3; l++) {
        const float lim = sqrtf(6.f / (float)fan[l]);
        for (int i = 0; i < sizes[2 * l]; i++) h[m->off[2 * l] + i] = (float)((2.0 * sp_rng_uniform(&r) - 1.0) * lim);
    }
    int rc = sp_mlp_set_params(m, h.data(), (size_t)m->np);
    if (!rc) {
        (void)hipMemset(m->m, 0, (size_t)m->np * 4);
        (void)hipMemset(m->v, 0, (size_t)m->np * 4);
    }
    if (rc) { sp_mlp_destroy(m); return rc; }
    *res = m;
    return SP_OK;
}

extern "C" int sp_mlp_destroy(sp_mlp *m) {
    if (!m) return SP_ERR_ARG;
    float *bufs[] = {m->p, m->g, m->m, m->v, m->x, m->a1, m->z1, m->a2, m->z2, m->y, m->t, m->msk, m->d3, m->d2, m->d1, m->part};
    for (float *b : bufs)
        if (b) (void)hipFree(b);
    if (m->pb) (void)hipFree(m->pb);
    delete m;
    return SP_OK;
}

extern "C" size_t sp_mlp_num_params(const sp_mlp *m) { return m ? (size_t)m->np : 0; }

extern "C" int sp_mlp_set_params(sp_mlp *m, const float *host, size_t n) {
    if (!m || !host || n != (size_t)m->np) return SP_ERR_ARG;
    SP_HIP_TRY(hipMemcpy(m->p, host, n * 4, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(to_bf16, dim3((m->np + 255) / 256), dim3(256), 0, m->ctx->stream, m->p, m->pb, m->np);
    SP_HIP_TRY(hipStreamSynchronize(m->ctx->stream));
    m->step = 0;
    return SP_OK;
}

extern "C" int sp_mlp_get_params(sp_mlp *m, float *host, size_t n) {
    if (!m || !host || n != (size_t)m->np) return SP_ERR_ARG;
    SP_HIP_TRY(hipStreamSynchronize(m->ctx->stream));
    SP_HIP_TRY(hipMemcpy(host, m->p, n * 4, hipMemcpyDeviceToHost));
    return SP_OK;
}

static int forward_dev(sp_mlp *m, int B) {
    hipStream_t s = m->ctx->stream;
    const dim3 blk(TB * TB);
    auto grid = [&](int O) { return dim3((O + TB - 1) / TB, (B + TB - 1) / TB); };
    hipLaunchKernelGGL(dense_fwd, grid(m->h1), blk, 0, s, m->x, m->pb + m->off[0], m->p + m->off[1], m->z1, m->a1, B, m->in, m->h1, 1);
    hipLaunchKernelGGL(dense_fwd, grid(m->h2), blk, 0, s, m->a1, m->pb + m->off[2], m->p + m->off[3], m->z2, m->a2, B, m->h1, m->h2, 1);
    hipLaunchKernelGGL(dense_fwd, grid(m->out), blk, 0, s, m->a2, m->pb + m->off[4], m->p + m->off[5], (float *)nullptr, m->y, B, m->h2, m->out, 0);
    return hipGetLastError() == hipSuccess ? SP_OK : SP_ERR_HIP;
}

extern "C" int sp_mlp_predict(sp_mlp *m, const float *X, int B, float *Y) {
    if (!m || !X || !Y || B <= 0) return SP_ERR_ARG;
    int rc = mlp_reserve(m, B);
    if (rc) return rc;
    SP_HIP_TRY(hipMemcpyAsync(m->x, X, (size_t)B * m->in * 4, hipMemcpyHostToDevice, m->ctx->stream));
    if ((rc = forward_dev(m, B))) return rc;
    SP_HIP_TRY(hipMemcpyAsync(Y, m->y, (size_t)B * m->out * 4, hipMemcpyDeviceToHost, m->ctx->stream));
    SP_HIP_TRY(hipStreamSynchronize(m->ctx->stream));
    return SP_OK;
}

// Forward + masked-MSE backward for one batch already in m->x / m->t / m->msk.
static int loss_grad_dev(sp_mlp *m, int B, float *loss) {
    hipStream_t s = m->ctx->stream;
    int rc = forward_dev(m, B);
    if (rc) return rc;
    const int ny = B * m->out, nb = (ny + 255) / 256;
    hipLaunchKernelGGL(mse_grad, dim3(nb), dim3(256), 0, s, m->y, m->t, m->msk, m->d3, m->part, ny, 1.f / (float)ny);
    // layer 3
    hipLaunchKernelGGL(dense_bwd_w, dim3((m->h2 + 1 + 63) / 64, m->out), dim3(64), 0, s, m->d3, m->a2, m->g + m->off[4], m->g + m->off[5], B, m->h2, m->out);
    hipLaunchKernelGGL(dense_bwd_x, dim3((m->h2 + 63) / 64, B), dim3(64), 0, s, m->d3, m->pb + m->off[4], m->d2, B, m->h2, m->out);
    hipLaunchKernelGGL(relu_bwd, dim3((B * m->h2 + 255) / 256), dim3(256), 0, s, m->d2, m->z2, B * m->h2);
    // layer 2
    hipLaunchKernelGGL(dense_bwd_w, dim3((m->h1 + 1 + 63) / 64, m->h2), dim3(64), 0, s, m->d2, m->a1, m->g + m->off[2], m->g + m->off[3], B, m->h1, m->h2);
    hipLaunchKernelGGL(dense_bwd_x, dim3((m->h1 + 63) / 64, B), dim3(64), 0, s, m->d2, m->pb + m->off[2], m->d1, B, m->h1, m->h2);
    hipLaunchKernelGGL(relu_bwd, dim3((B * m->h1 + 255) / 256), dim3(256), 0, s, m->d1, m->z1, B * m->h1);
    // layer 1
    hipLaunchKernelGGL(dense_bwd_w, dim3((m->in + 1 + 63) / 64, m->h1), dim3(64), 0, s, m->d1, m->x, m->g + m->off[0], m->g + m->off[1], B, m->in, m->h1);
    if (hipGetLastError() != hipSuccess) return SP_ERR_HIP;
    if (loss) {
        std::vector<float> part(nb);
        SP_HIP_TRY(hipMemcpyAsync(part.data(), m->part, (size_t)nb * 4, hipMemcpyDeviceToHost, s));
        SP_HIP_TRY(hipStreamSynchronize(s));
        double l = 0;
        for (float v : part) l += v;
        *loss = (float)(l / ny);
    }
    return SP_OK;
}

extern "C" int sp_mlp_loss_grad(sp_mlp *m, const float *X, const float *T, const float *M, int B, float *loss,
                                float *grad) {
    if (!m || !X || !T || B <= 0) return SP_ERR_ARG;
    int rc = mlp_reserve(m, B);
    if (rc) return rc;
    hipStream_t s = m->ctx->stream;
    SP_HIP_TRY(hipMemcpyAsync(m->x, X, (size_t)B * m->in * 4, hipMemcpyHostToDevice, s));
    SP_HIP_TRY(hipMemcpyAsync(m->t, T, (size_t)B * m->out * 4, hipMemcpyHostToDevice, s));
    if (M) SP_HIP_TRY(hipMemcpyAsync(m->msk, M, (size_t)B * m->out * 4, hipMemcpyHostToDevice, s));
    else {
        std::vector<float> ones((size_t)B * m->out, 1.f);
        SP_HIP_TRY(hipMemcpyAsync(m->msk, ones.data(), ones.size() * 4, hipMemcpyHostToDevice, s));
        SP_HIP_TRY(hipStreamSynchronize(s));
    }
    if ((rc = loss_grad_dev(m, B, loss))) return rc;
    if (grad) {
        SP_HIP_TRY(hipMemcpyAsync(grad, m->g, (size_t)m->np * 4, hipMemcpyDeviceToHost, s));
        SP_HIP_TRY(hipStreamSynchronize(s));
    }
    return SP_OK;
}

extern "C" int sp_mlp_adam(sp_mlp *m, float lr, float b1, float b2, float eps) {
    if (!m) return SP_ERR_ARG;
    m->step++;
    const float c1 = 1.f - powf(b1, (float)m->step), c2 = 1.f - powf(b2, (float)m->step);
    hipLaunchKernelGGL(adam_step, dim3((m->np + 255) / 256), dim3(256), 0, m->ctx->stream, m->p, m->g, m->m, m->v, m->pb,
                       m->np, lr, b1, b2, eps, c1, c2);
    return hipGetLastError() == hipSuccess ? SP_OK : SP_ERR_HIP;
}

// Mini-batch Adam over n samples for `epochs`, shuffled with the project
// PRNG; returns the mean loss of the last epoch.
extern "C" int sp_mlp_train(sp_mlp *m, const float *X, const float *T, const float *M, int n, int epochs,
                            int batch, float lr, uint64_t seed, float *final_loss) {
    if (!m || !X || !T || n <= 0 || epochs <= 0 || batch <= 0) return SP_ERR_ARG;
    const int B = batch < n ? batch : n;
    int rc = mlp_reserve(m, B);
    if (rc) return rc;
    std::vector<int> order(n);
    for (int i = 0; i < n; i++) order[i] = i;
    std::vector<float> bx((size_t)B * m->in), bt((size_t)B * m->out), bm((size_t)B * m->out);
    sp_rng r;
    sp_rng_seed(&r, seed);
    hipStream_t s = m->ctx->stream;
    double last = 0;
    for (int e = 0; e < epochs; e++) {
        for (int i = n - 1; i > 0; i--) {
            const int j = (int)(((__uint128_t)sp_rng_next(&r) * (uint64_t)(i + 1)) >> 64);
            std::swap(order[i], order[j]);
        }
        double tot = 0;
        int nbatch = 0;
        for (int s0 = 0; s0 + B <= n; s0 += B) {
            for (int q = 0; q < B; q++) {
                const int src = order[s0 + q];
                memcpy(&bx[(size_t)q * m->in], X + (size_t)src * m->in, (size_t)m->in * 4);
                memcpy(&bt[(size_t)q * m->out], T + (size_t)src * m->out, (size_t)m->out * 4);
                if (M) memcpy(&bm[(size_t)q * m->out], M + (size_t)src * m->out, (size_t)m->out * 4);
                else for (int o = 0; o < m->out; o++) bm[(size_t)q * m->out + o] = 1.f;
            }
            SP_HIP_TRY(hipMemcpyAsync(m->x, bx.data(), bx.size() * 4, hipMemcpyHostToDevice, s));
            SP_HIP_TRY(hipMemcpyAsync(m->t, bt.data(), bt.size() * 4, hipMemcpyHostToDevice, s));
            SP_HIP_TRY(hipMemcpyAsync(m->msk, bm.data(), bm.size() * 4, hipMemcpyHostToDevice, s));
            float l = 0;
            if ((rc = loss_grad_dev(m, B, &l))) return rc;
            if ((rc = sp_mlp_adam(m, lr, 0.9f, 0.999f, 1e-8f))) return rc;
            tot += l;
            nbatch++;
        }
        last = nbatch ? tot / nbatch : 0;
    }
    SP_HIP_TRY(hipStreamSynchronize(s));
    if (final_loss) *final_loss = (float)last;
    return SP_OK;
}
