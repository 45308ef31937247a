// Allocation-free CSR SpMV (sp_spmv_csr_noalloc): the entry-balanced
// merge-path SpMV of a CSR-nnz plan, run on CSR arrays the caller already
// holds in device memory, with no plan and no device allocation.  The tile
// coordinates, the problem descriptor and the cross-tile carries live in
// `scratch`, a fixed-size device array of this library; each device holds
// its own copy, whose address is looked up once per device at context
// creation (sp_noalloc_setup) and passed to the kernels.
//
// Kernels, in stream order:
//   coords_kernel      merge-path search of every tile boundary over the
//                      device rowptr (the host-side merge_coords of a plan)
//   csr_merge_kernel   the plan's kernel, row-ordered 2048-item tiles
//   fixup_kernel       the carries of rows crossing tiles, in tile order
// Every scratch word a launch reads is written earlier in the same call,
// so nothing needs clearing.  All calls on one device share the scratch:
// calls that overlap on one GPU (other streams or contexts) race.
#include <mutex>

#include "plan.hpp"

namespace {

// 2048-item tiles (256 threads x 8 items), as a CSR-nnz plan's row-ordered
// merge tiles; SP_NOALLOC_MAX_ITEMS merge items make at most kMaxTiles tiles.
constexpr int kIpt = 8;
constexpr uint32_t kTile = spd::TPB * kIpt;
constexpr uint32_t kMaxTiles = SP_NOALLOC_MAX_ITEMS / kTile;
static_assert(SP_NOALLOC_MAX_ITEMS % kTile == 0, "whole tiles");
constexpr size_t kAlign = 256;
constexpr size_t align_up(size_t v) { return (v + kAlign - 1) / kAlign * kAlign; }

// scratch layout (bytes)
constexpr size_t kOffProb = 0;
constexpr size_t kOffCrow = align_up(kOffProb + sizeof(spd::Prob));
constexpr size_t kOffCnnz = align_up(kOffCrow + (kMaxTiles + 1) * 4);
constexpr size_t kOffPid = align_up(kOffCnnz + (kMaxTiles + 1) * 4);
constexpr size_t kOffCrowCarry = align_up(kOffPid + kMaxTiles * 4);
constexpr size_t kOffCval = align_up(kOffCrowCarry + kMaxTiles * 4);
constexpr size_t kScratchBytes = align_up(kOffCval + kMaxTiles * 8);

static __device__ __align__(256) uint8_t scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
void *g_ptr[kMaxDevices];  // device address of `scratch`, per device

// Thread t <= ntiles: the merge-path coordinate (row, entry) of tile
// boundary t; threads t < ntiles also write the tile's problem id (0).
// Thread 0 stores the problem descriptor the later kernels read.
__global__ void coords_kernel(const uint32_t *__restrict__ rowptr, uint32_t nrows, uint32_t nnz, uint32_t ntiles,
                              spd::Prob P, spd::Prob *__restrict__ pdst, uint32_t *__restrict__ crow,
                              uint32_t *__restrict__ cnnz, uint32_t *__restrict__ pid) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t == 0) *pdst = P;
    if (t > ntiles) return;
    const spd::gptr<const uint32_t> rp = spd::gp(rowptr);
    const uint32_t base = rp[0];
    const uint64_t total = (uint64_t)nrows + nnz;
    const uint32_t d = (uint32_t)min((uint64_t)t * kTile, total);
    const uint32_t i = spd::merge_search(d, nrows, nnz, [&](uint32_t r) { return rp[r + 1] - base; }, 0u);
    crow[t] = i;
    cnnz[t] = d - i + base;
    if (t < ntiles) pid[t] = 0;
}

template <typename T>
int run_t(hipStream_t s, uint8_t *sc, uint32_t nrows, uint32_t ncols, uint32_t nnz, const uint32_t *rowptr,
          const uint32_t *colind, const void *values, const void *x, void *y) {
    using namespace spd;
    using A = acc_t<T>;
    Prob *d_prob = (Prob *)(sc + kOffProb);
    uint32_t *crow = (uint32_t *)(sc + kOffCrow), *cnnz = (uint32_t *)(sc + kOffCnnz);
    uint32_t *pid = (uint32_t *)(sc + kOffPid), *carry_row = (uint32_t *)(sc + kOffCrowCarry);
    A *carry_val = (A *)(sc + kOffCval);
    const uint64_t total = (uint64_t)nrows + nnz;
    const uint32_t ntiles = (uint32_t)((total + kTile - 1) / kTile);

    Prob P{};
    P.ptr = rowptr;
    P.cind = colind;
    P.val = values;
    P.nrows = nrows;
    P.nent = nnz;
    P.nrows_elem = nrows;
    P.ncols_elem = ncols;
    P.ldx = P.ldy = 1;
    Ptrs<T> io{};
    io.xb = (const T *)x;
    io.yb = (T *)y;
    io.ldx = io.ldy = 1;

    hipLaunchKernelGGL(coords_kernel, dim3((ntiles + 1 + TPB - 1) / TPB), dim3(TPB), 0, s, rowptr, nrows, nnz,
                       ntiles, P, d_prob, crow, cnnz, pid);
    if (hipGetLastError() != hipSuccess) return SP_ERR_HIP;
    hipLaunchKernelGGL((csr_merge_kernel<T, kIpt, 1>), dim3(ntiles), dim3(TPB), 0, s, d_prob,
                       (const uint32_t *)nullptr, crow, cnnz, io, carry_row, carry_val, ntiles);
    if (hipGetLastError() != hipSuccess) return SP_ERR_HIP;
    hipLaunchKernelGGL((fixup_kernel<T, 1>), dim3((ntiles + TPB - 1) / TPB), dim3(TPB), 0, s, d_prob, pid,
                       carry_row, carry_val, ntiles, 1u, io, (const uint32_t *)nullptr, (const uint32_t *)nullptr,
                       0u);
    return hipGetLastError() == hipSuccess ? SP_OK : SP_ERR_HIP;
}

}  // namespace

// Looks up this device's copy of `scratch` (the device must be current) and
// keeps its address.  Called by sp_ctx_create, once per device; the entry
// point repeats the lookup on its first call if that one failed.
int sp_noalloc_setup(int dev) {
    if (dev < 0 || dev >= kMaxDevices) return SP_ERR_ARG;
    std::lock_guard<std::mutex> g(g_mu);
    if (g_ptr[dev]) return SP_OK;
    void *p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return SP_ERR_HIP;
    g_ptr[dev] = p;
    return SP_OK;
}

extern "C" int sp_spmv_csr_noalloc(sp_ctx *ctx, int dtype, uint32_t nrows, uint32_t ncols, uint32_t nnz,
                                   const uint32_t *rowptr, const uint32_t *colind, const void *values,
                                   const void *x, void *y) {
    if (!ctx || sp_dtype_size(dtype) == 0) return SP_ERR_ARG;
    if (nrows == 0) return SP_OK;
    if (!rowptr || !y || (nnz && (!colind || !values || !x))) return SP_ERR_ARG;
    if ((uint64_t)nrows + nnz > (uint64_t)SP_NOALLOC_MAX_ITEMS) return SP_ERR_RANGE;
    if (ctx->device < 0 || ctx->device >= kMaxDevices) return SP_ERR_ARG;
    if (hipSetDevice(ctx->device) != hipSuccess) return SP_ERR_HIP;
    if (int rc = sp_noalloc_setup(ctx->device)) return rc;
    void *sc;
    {
        std::lock_guard<std::mutex> g(g_mu);
        sc = g_ptr[ctx->device];
    }
    uint8_t *b = (uint8_t *)sc;
    hipStream_t s = ctx->stream;
    switch (dtype) {
    case SP_I8: return run_t<int8_t>(s, b, nrows, ncols, nnz, rowptr, colind, values, x, y);
    case SP_I16: return run_t<int16_t>(s, b, nrows, ncols, nnz, rowptr, colind, values, x, y);
    case SP_I32: return run_t<int32_t>(s, b, nrows, ncols, nnz, rowptr, colind, values, x, y);
    case SP_I64: return run_t<int64_t>(s, b, nrows, ncols, nnz, rowptr, colind, values, x, y);
    case SP_F32: return run_t<float>(s, b, nrows, ncols, nnz, rowptr, colind, values, x, y);
    case SP_F64: return run_t<double>(s, b, nrows, ncols, nnz, rowptr, colind, values, x, y);
    default: return SP_ERR_ARG;
    }
}
