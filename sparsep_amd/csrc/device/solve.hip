// Conjugate gradient on a plan's matrix: the iterative-solver caller that
// reuses one plan (and its tuned kernel) for every iteration.  Dot products
// reduce in a fixed order (per-block partial sums, then one block), so
// iterates are run-to-run identical.  No counterpart in the reference.
#include <math.h>

#include "plan.hpp"

namespace {

constexpr int RB = 256;      // reduction block
constexpr int RGRID = 1024;  // partial sums per dot product

template <typename T>
__global__ __launch_bounds__(RB) void dot_partial(const T *__restrict__ a, const T *__restrict__ b, uint64_t n,
                                                  double *__restrict__ part) {
    __shared__ double s[RB];
    double acc = 0.0;
    for (uint64_t i = (uint64_t)blockIdx.x * RB + threadIdx.x; i < n; i += (uint64_t)gridDim.x * RB)
        acc += (double)a[i] * (double)b[i];
    s[threadIdx.x] = acc;
    __syncthreads();
    for (int w = RB / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = s[0];
}

__global__ __launch_bounds__(RB) void dot_final(double *__restrict__ part, int np) {
    __shared__ double s[RB];
    double acc = 0.0;
    for (int i = threadIdx.x; i < np; i += RB) acc += part[i];
    s[threadIdx.x] = acc;
    __syncthreads();
    for (int w = RB / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[RGRID] = s[0];
}

// r = b - r (r holds A x on entry); p = r
template <typename T>
__global__ void residual(const T *__restrict__ b, T *__restrict__ r, T *__restrict__ p, uint64_t n) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        const T v = b[i] - r[i];
        r[i] = v;
        p[i] = v;
    }
}

// x += alpha p; r -= alpha q
template <typename T>
__global__ void update_xr(T *__restrict__ x, T *__restrict__ r, const T *__restrict__ p, const T *__restrict__ q,
                          T alpha, uint64_t n) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        x[i] = x[i] + alpha * p[i];
        r[i] = r[i] - alpha * q[i];
    }
}

// p = r + beta p
template <typename T>
__global__ void update_p(T *__restrict__ p, const T *__restrict__ r, T beta, uint64_t n) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = r[i] + beta * p[i];
}

template <typename T>
int cg_t(sp_plan *A, const T *b, T *x, int maxit, double rtol, int *iters, double *relres) {
    sp_ctx *c = A->ctx;
    hipStream_t s = c->stream;
    const uint64_t n = A->nrows;
    T *r = nullptr, *p = nullptr, *q = nullptr;
    double *part = nullptr;
    int rc = SP_OK;
    auto cleanup = [&]() {
        if (r) (void)hipFree(r);
        if (p) (void)hipFree(p);
        if (q) (void)hipFree(q);
        if (part) (void)hipFree(part);
    };
    if (hipMalloc((void **)&r, (n ? n : 1) * sizeof(T)) != hipSuccess ||
        hipMalloc((void **)&p, (n ? n : 1) * sizeof(T)) != hipSuccess ||
        hipMalloc((void **)&q, (n ? n : 1) * sizeof(T)) != hipSuccess ||
        hipMalloc((void **)&part, (RGRID + 1) * sizeof(double)) != hipSuccess) {
        cleanup();
        return SP_ERR_NOMEM;
    }
    const dim3 vg((unsigned)std::min<uint64_t>((n + 255) / 256 + 1, 4096)), vb(256);
    auto dot = [&](const T *u, const T *v, double *out) -> int {
        hipLaunchKernelGGL(dot_partial<T>, dim3(RGRID), dim3(RB), 0, s, u, v, n, part);
        hipLaunchKernelGGL(dot_final, dim3(1), dim3(RB), 0, s, part, RGRID);
        if (hipGetLastError() != hipSuccess) return SP_ERR_HIP;
        if (hipMemcpyAsync(out, part + RGRID, sizeof(double), hipMemcpyDeviceToHost, s) != hipSuccess ||
            hipStreamSynchronize(s) != hipSuccess)
            return SP_ERR_HIP;
        return SP_OK;
    };
    double bb = 0, rr = 0;
    int it = 0;
    do {
        if ((rc = sp_spmv(A, x, r))) break;
        hipLaunchKernelGGL(residual<T>, vg, vb, 0, s, b, r, p, n);
        if ((rc = dot(b, b, &bb)) || (rc = dot(r, r, &rr))) break;
        const double bn = bb > 0 ? sqrt(bb) : 1.0;
        while (it < maxit && sqrt(rr) / bn > rtol) {
            if ((rc = sp_spmv(A, p, q))) break;
            double pq = 0;
            if ((rc = dot(p, q, &pq))) break;
            if (!(pq > 0)) { rc = SP_ERR_STATE; break; }  // not positive definite
            const double alpha = rr / pq;
            hipLaunchKernelGGL(update_xr<T>, vg, vb, 0, s, x, r, p, q, (T)alpha, n);
            double rr1 = 0;
            if ((rc = dot(r, r, &rr1))) break;
            hipLaunchKernelGGL(update_p<T>, vg, vb, 0, s, p, r, (T)(rr1 / rr), n);
            rr = rr1;
            it++;
        }
        if (relres) *relres = sqrt(rr) / bn;
    } while (0);
    if (iters) *iters = it;
    if (!rc && hipStreamSynchronize(s) != hipSuccess) rc = SP_ERR_HIP;
    cleanup();
    return rc;
}

}  // namespace

extern "C" int sp_cg(sp_plan *A, const void *b, void *x, int maxit, double rtol, int *iters, double *relres) {
    if (!A || !b || !x || maxit < 0 || !(rtol >= 0)) return SP_ERR_ARG;
    if (A->nrows != A->ncols) return SP_ERR_ARG;
    (void)hipSetDevice(A->ctx->device);
    switch (A->dtype) {
    case SP_F32: return cg_t<float>(A, (const float *)b, (float *)x, maxit, rtol, iters, relres);
    case SP_F64: return cg_t<double>(A, (const double *)b, (double *)x, maxit, rtol, iters, relres);
    default: return SP_ERR_UNSUPPORTED;
    }
}
