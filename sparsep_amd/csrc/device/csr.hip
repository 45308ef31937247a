// CSR SpMV on gfx950: the device replacement of SparseP's CSR tasklet loop
// (spmv/1D/CSR-row/dpu/task.c:105-202, identical in CSR-nnz and, with a
// tile-local row pointer, in spmv/2D/{DCSR,RBDCSR,BDCSR}/dpu/task.c).
//
// Two algorithms:
//   VECTOR     row granular.  The part split (partition_by_row or
//              partition_by_nnz) gives each workgroup a contiguous row
//              range; inside it, groups of L lanes (L = 1..64) take rows in
//              interleaved order so a wave's loads stay contiguous, and each
//              group reduces its row with a ds_swizzle / DPP butterfly.
//   SEGMENTED  entry balanced (merge path over row ends and entries,
//              Merrill & Garland 2016).  Every workgroup consumes exactly
//              TPB*IPT merge items (rows + entries), so long rows are split
//              across workgroups.  Within a tile: coalesced loads of
//              colind/vals, products staged in LDS, a per-thread walk, a
//              segmented scan of thread carries; carries that cross tiles
//              are added in tile order by a second small kernel.
// Both are deterministic: every row's sum is formed in a fixed order.
#include "common.hpp"
#include "plan.hpp"

#include <algorithm>
#include <vector>

namespace spd {

constexpr int TPB = 256;

// ---------------------------------------------------------------------------
// VECTOR: one workgroup per part, L lanes per row.
// ---------------------------------------------------------------------------
template <typename T, int L, int K>
__global__ __launch_bounds__(TPB) void csr_vector_kernel(const uint32_t *__restrict__ rowptr,
                                                         const uint32_t *__restrict__ colind,
                                                         const T *__restrict__ vals,
                                                         const T *__restrict__ x, T *__restrict__ y,
                                                         const uint32_t *__restrict__ part, int ldx,
                                                         int ldy) {
    using A = acc_t<T>;
    constexpr int G = TPB / L;
    const uint32_t r0 = part[blockIdx.x], r1 = part[blockIdx.x + 1];
    const int lane = threadIdx.x % L, grp = threadIdx.x / L;
    for (uint32_t r = r0 + grp; r < r1; r += G) {
        const uint32_t b = rowptr[r], e = rowptr[r + 1];
        A acc[K];
#pragma unroll
        for (int q = 0; q < K; q++) acc[q] = A(0);
        uint32_t k = b + lane;
        // two entries per lane per iteration keep more loads in flight
        for (; k + L < e; k += 2 * L) {
            const uint32_t c0 = ldg(colind + k), c1 = ldg(colind + k + L);
            const T v0 = ldg(vals + k), v1 = ldg(vals + k + L);
#pragma unroll
            for (int q = 0; q < K; q++) {
                acc[q] = fma_acc<T>(v0, x[(size_t)c0 * ldx + q], acc[q]);
                acc[q] = fma_acc<T>(v1, x[(size_t)c1 * ldx + q], acc[q]);
            }
        }
        if (k < e) {
            const uint32_t c0 = ldg(colind + k);
            const T v0 = ldg(vals + k);
#pragma unroll
            for (int q = 0; q < K; q++) acc[q] = fma_acc<T>(v0, x[(size_t)c0 * ldx + q], acc[q]);
        }
#pragma unroll
        for (int q = 0; q < K; q++) acc[q] = group_sum<L>(acc[q]);
        if (lane == 0) {
#pragma unroll
            for (int q = 0; q < K; q++) y[(size_t)r * ldy + q] = to_val<T>(acc[q]);
        }
    }
}

// ---------------------------------------------------------------------------
// SEGMENTED (merge path).
// ---------------------------------------------------------------------------

// Merge-path split of diagonal d: the number of row ends consumed when d
// merge items (row ends + entries) have been consumed.  A row end
// rowend[i] is taken before entry j when rowend[i] <= j.
template <typename RE>
__device__ __host__ inline uint32_t merge_search(uint32_t d, uint32_t na, uint32_t nb, RE rowend,
                                                 uint32_t jbase) {
    uint32_t lo = d > nb ? d - nb : 0, hi = d < na ? d : na;
    while (lo < hi) {
        uint32_t mid = (lo + hi) >> 1;
        if (rowend(mid) <= jbase + (d - mid - 1)) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

template <typename A> struct WaveScan {
    // Inclusive segmented scan over a wave; keys are non-decreasing.
    static __device__ __forceinline__ A run(uint32_t key, A val, int lane) {
#pragma unroll
        for (int off = 1; off < WAVE; off <<= 1) {
            uint32_t ok = (uint32_t)__shfl_up((int)key, off, WAVE);
            A ov;
            if constexpr (sizeof(A) == 4) {
                ov = __builtin_bit_cast(A, __shfl_up(__builtin_bit_cast(int, val), off, WAVE));
            } else {
                uint64_t u = __builtin_bit_cast(uint64_t, val);
                int lo = __shfl_up((int)(uint32_t)u, off, WAVE);
                int hi = __shfl_up((int)(uint32_t)(u >> 32), off, WAVE);
                ov = __builtin_bit_cast(A, ((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo);
            }
            if (lane >= off && ok == key) val = ov + val;
        }
        return val;
    }
};

template <typename T, int IPT, int K>
__global__ __launch_bounds__(TPB) void csr_merge_kernel(
    const uint32_t *__restrict__ rowptr, const uint32_t *__restrict__ colind,
    const T *__restrict__ vals, const T *__restrict__ x, T *__restrict__ y,
    const uint32_t *__restrict__ tile_row, const uint32_t *__restrict__ tile_nnz, uint32_t nrows,
    uint32_t *__restrict__ carry_row, acc_t<T> *__restrict__ carry_val, int ldx, int ldy) {
    using A = acc_t<T>;
    constexpr int TILE = TPB * IPT;
    constexpr int NW = TPB / WAVE;
    // One right-hand side: stage the products (the x gather then happens in
    // the coalesced load loop).  Several: stage (column, value) pairs and
    // gather x rows during the walk.
    constexpr bool PROD = (K == 1);
    __shared__ uint32_t s_rowend[TILE + 1];
    __shared__ A s_prod[PROD ? TILE : 1];
    __shared__ uint32_t s_col[PROD ? 1 : TILE];
    __shared__ T s_val[PROD ? 1 : TILE];
    __shared__ uint32_t s_key[TPB];
    __shared__ A s_scan[K][TPB];
    __shared__ A s_wcarry[K][NW];

    const int tid = threadIdx.x, lane = tid % WAVE, wid = tid / WAVE;
    const uint32_t t = blockIdx.x;
    const uint32_t i0 = tile_row[t], j0 = tile_nnz[t];
    const uint32_t na = tile_row[t + 1] - i0;
    const uint32_t nb = tile_nnz[t + 1] - j0;

    // Stage row ends of the rows this tile touches, and the entries.
    const uint32_t nrl = min(na + 1, nrows - i0);
    for (uint32_t k = tid; k < nrl; k += TPB) s_rowend[k] = rowptr[i0 + k + 1];
#pragma unroll 4
    for (uint32_t k = tid; k < nb; k += TPB) {
        const uint32_t c = ldg(colind + j0 + k);
        const T v = ldg(vals + j0 + k);
        if constexpr (PROD) {
            s_prod[k] = mul<T>(v, x[c]);
        } else {
            s_col[k] = c;
            s_val[k] = v;
        }
    }
    __syncthreads();

    // Each thread walks IPT merge items from its own diagonal.
    const uint32_t total = na + nb;
    const uint32_t d0 = min((uint32_t)tid * IPT, total);
    const uint32_t d1 = min(d0 + IPT, total);
    uint32_t li = merge_search(d0, na, nb, [&](uint32_t i) { return s_rowend[i]; }, j0);
    uint32_t lj = d0 - li;
    A acc[K], first[K];
#pragma unroll
    for (int q = 0; q < K; q++) { acc[q] = A(0); first[q] = A(0); }
    bool emitted = false;
    uint32_t first_row = 0;
    for (uint32_t s = d0; s < d1; s++) {
        if (li < na && s_rowend[li] <= j0 + lj) {
            const uint32_t row = i0 + li;
            if (!emitted) {
                emitted = true;
                first_row = row;
#pragma unroll
                for (int q = 0; q < K; q++) first[q] = acc[q];
            } else {
#pragma unroll
                for (int q = 0; q < K; q++) y[(size_t)row * ldy + q] = to_val<T>(acc[q]);
            }
#pragma unroll
            for (int q = 0; q < K; q++) acc[q] = A(0);
            li++;
        } else {
            if constexpr (PROD) {
                acc[0] = acc[0] + s_prod[lj];
            } else {
                const uint32_t c = s_col[lj];
                const T v = s_val[lj];
#pragma unroll
                for (int q = 0; q < K; q++) acc[q] = acc[q] + mul<T>(v, x[(size_t)c * ldx + q]);
            }
            lj++;
        }
    }

    // Segmented scan of thread carries (key = row the carry belongs to).
    const uint32_t key = i0 + li;
    A scan[K];
#pragma unroll
    for (int q = 0; q < K; q++) scan[q] = WaveScan<A>::run(key, acc[q], lane);
    s_key[tid] = key;
#pragma unroll
    for (int q = 0; q < K; q++) s_scan[q][tid] = scan[q];
    __syncthreads();
    if (tid < K) {
        // carry into wave w = segmented total ending at wave w-1's last lane
        const int q = tid;
        A run = A(0);
        uint32_t run_key = 0xFFFFFFFFu;
        for (int w = 0; w < NW; w++) {
            s_wcarry[q][w] = run;
            const uint32_t lk = s_key[w * WAVE + WAVE - 1];
            const A lv = s_scan[q][w * WAVE + WAVE - 1];
            run = (w > 0 && lk == run_key) ? run + lv : lv;
            run_key = lk;
        }
    }
    __syncthreads();
    if (wid > 0 && key == s_key[wid * WAVE - 1]) {
#pragma unroll
        for (int q = 0; q < K; q++) scan[q] = s_wcarry[q][wid] + scan[q];
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < K; q++) s_scan[q][tid] = scan[q];
    __syncthreads();
    if (emitted) {
#pragma unroll
        for (int q = 0; q < K; q++) {
            const A cin = tid > 0 ? s_scan[q][tid - 1] : A(0);
            y[(size_t)first_row * ldy + q] = to_val<T>(cin + first[q]);
        }
    }
    if (tid == TPB - 1) {
        carry_row[t] = key;
#pragma unroll
        for (int q = 0; q < K; q++) carry_val[(size_t)t * K + q] = scan[q];
    }
}

// Adds the carries of row runs that cross tiles, in tile order.
template <typename T, int K>
__global__ void carry_fixup_kernel(const uint32_t *__restrict__ carry_row,
                                   const acc_t<T> *__restrict__ carry_val, uint32_t ntiles,
                                   uint32_t nrows, T *__restrict__ y, int ldy) {
    using A = acc_t<T>;
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ntiles) return;
    const uint32_t r = carry_row[t];
    if (r >= nrows) return;
    if (t > 0 && carry_row[t - 1] == r) return;
    A sum[K];
#pragma unroll
    for (int q = 0; q < K; q++) sum[q] = A(0);
    for (uint32_t u = t; u < ntiles && carry_row[u] == r; u++) {
#pragma unroll
        for (int q = 0; q < K; q++) sum[q] = sum[q] + carry_val[(size_t)u * K + q];
    }
#pragma unroll
    for (int q = 0; q < K; q++) {
        T *dst = y + (size_t)r * ldy + q;
        *dst = to_val<T>((A)(*dst) + sum[q]);
    }
}

// ---------------------------------------------------------------------------
// Launch helpers
// ---------------------------------------------------------------------------

template <typename T, int K>
static int launch_vector(sp_plan *p, const void *x, void *y, int ldx, int ldy) {
    hipStream_t s = p->ctx->stream;
    const T *vx = (const T *)x;
    T *vy = (T *)y;
    const dim3 grid(p->nparts), block(TPB);
#define SPD_VEC(Lv)                                                                              \
    case Lv:                                                                                     \
        hipLaunchKernelGGL((csr_vector_kernel<T, Lv, K>), grid, block, 0, s, p->m.rowptr,        \
                           p->m.colind, (const T *)p->m.vals, vx, vy, p->d_part, ldx, ldy);      \
        break;
    switch (p->lanes) {
        SPD_VEC(1) SPD_VEC(2) SPD_VEC(4) SPD_VEC(8) SPD_VEC(16) SPD_VEC(32) SPD_VEC(64)
    default: return SP_ERR_ARG;
    }
#undef SPD_VEC
    mark_dominant(p->ctx);
    return hipGetLastError() == hipSuccess ? SP_OK : SP_ERR_HIP;
}

template <typename T, int K>
static int launch_merge(sp_plan *p, const void *x, void *y, int ldx, int ldy) {
    hipStream_t s = p->ctx->stream;
    const dim3 grid(p->ntiles), block(TPB);
    using A = acc_t<T>;
#define SPD_MRG(I)                                                                               \
    case I:                                                                                      \
        hipLaunchKernelGGL((csr_merge_kernel<T, I, K>), grid, block, 0, s, p->m.rowptr,          \
                           p->m.colind, (const T *)p->m.vals, (const T *)x, (T *)y, p->d_tile_row, \
                           p->d_tile_pos, p->nrows, p->d_carry_row, (A *)p->d_carry_val, ldx,    \
                           ldy);                                                                 \
        break;
    if (K > 1 && p->ipt > 8) return SP_ERR_UNSUPPORTED;
    switch (p->ipt) {
        SPD_MRG(4) SPD_MRG(8) SPD_MRG(16)
    default: return SP_ERR_ARG;
    }
#undef SPD_MRG
    if (hipGetLastError() != hipSuccess) return SP_ERR_HIP;
    mark_dominant(p->ctx);
    hipLaunchKernelGGL((carry_fixup_kernel<T, K>), dim3((p->ntiles + 255) / 256), dim3(256), 0, s,
                       p->d_carry_row, (const A *)p->d_carry_val, p->ntiles, p->nrows, (T *)y, ldy);
    return hipGetLastError() == hipSuccess ? SP_OK : SP_ERR_HIP;
}

template <typename T>
static int run_typed(sp_plan *p, const void *x, void *y, int k, int ldx, int ldy) {
    // Right-hand sides are processed in chunks of 8/4/2/1 columns.
    int done = 0;
    while (done < k) {
        const int left = k - done;
        const int kc = left >= 8 ? 8 : left >= 4 ? 4 : left >= 2 ? 2 : 1;
        const T *xs = (const T *)x + done;
        T *ys = (T *)y + done;
        int rc;
        if (p->algo == SP_ALGO_VECTOR) {
            rc = kc == 8 ? launch_vector<T, 8>(p, xs, ys, ldx, ldy)
               : kc == 4 ? launch_vector<T, 4>(p, xs, ys, ldx, ldy)
               : kc == 2 ? launch_vector<T, 2>(p, xs, ys, ldx, ldy)
                         : launch_vector<T, 1>(p, xs, ys, ldx, ldy);
        } else {
            if ((uint32_t)kc > p->carry_k) return SP_ERR_ARG;
            rc = kc == 8 ? launch_merge<T, 8>(p, xs, ys, ldx, ldy)
               : kc == 4 ? launch_merge<T, 4>(p, xs, ys, ldx, ldy)
               : kc == 2 ? launch_merge<T, 2>(p, xs, ys, ldx, ldy)
                         : launch_merge<T, 1>(p, xs, ys, ldx, ldy);
        }
        if (rc) return rc;
        done += kc;
    }
    return SP_OK;
}

}  // namespace spd

using namespace spd;

// Tile coordinates of the merge path for tiles of `tile` items.
static void merge_tiles(const sp_csr *A, uint32_t tile, std::vector<uint32_t> &trow,
                        std::vector<uint32_t> &tnnz) {
    const uint64_t total = (uint64_t)A->nrows + A->nnz;
    const uint32_t ntiles = (uint32_t)((total + tile - 1) / tile);
    trow.resize(ntiles + 1);
    tnnz.resize(ntiles + 1);
    const uint32_t *rp = A->rowptr;
    for (uint32_t t = 0; t <= ntiles; t++) {
        const uint64_t d = std::min<uint64_t>((uint64_t)t * tile, total);
        const uint32_t i = merge_search((uint32_t)d, A->nrows, A->nnz,
                                        [&](uint32_t r) { return rp[r + 1]; }, 0);
        trow[t] = i;
        tnnz[t] = (uint32_t)(d - i);
    }
}

int csr_setup(sp_plan *p, const sp_csr *A) {
    const size_t sv = sp_dtype_size(A->dtype);
    // The kernels trust the structure: validate it on the host first.
    if (!A->rowptr || (A->nnz && (!A->colind || !A->values))) return SP_ERR_ARG;
    if (A->rowptr[0] != 0 || A->rowptr[A->nrows] != A->nnz) return SP_ERR_ARG;
    for (uint32_t i = 0; i < A->nrows; i++)
        if (A->rowptr[i + 1] < A->rowptr[i]) return SP_ERR_ARG;
    for (uint32_t k = 0; k < A->nnz; k++)
        if (A->colind[k] >= A->ncols) return SP_ERR_ARG;
    p->m.format = SP_FMT_CSR;
    p->m.nrows = A->nrows;
    p->m.ncols = A->ncols;
    p->m.nnz = A->nnz;
    int rc;
    if ((rc = plan_upload(p, A->rowptr, ((size_t)A->nrows + 1) * 4, (void **)&p->m.rowptr))) return rc;
    if ((rc = plan_upload(p, A->colind, (size_t)A->nnz * 4, (void **)&p->m.colind))) return rc;
    if ((rc = plan_upload(p, A->values, (size_t)A->nnz * sv, &p->m.vals))) return rc;

    if (p->algo == SP_ALGO_VECTOR) {
        uint32_t L = p->opts.lanes_per_row;
        if (L == 0) {
            const double avg = A->nrows ? (double)A->nnz / A->nrows : 1.0;
            L = 1;
            while (L < 64 && L * 2 <= avg * 1.5) L *= 2;
        }
        if (L > 64 || (L & (L - 1))) return SP_ERR_ARG;
        p->lanes = L;
        uint32_t P = p->opts.nparts;
        if (P == 0) {
            // about 4 rows per lane group per workgroup
            const uint64_t rows_per_wg = 4ULL * (TPB / L);
            P = (uint32_t)std::max<uint64_t>(1, (A->nrows + rows_per_wg - 1) / rows_per_wg);
        }
        std::vector<uint32_t> split(P + 1);
        if (p->id == SP_K_CSR_NNZ) rc = sp_partition_by_nnz(A->rowptr, A->nrows, P, split.data());
        else rc = sp_partition_by_row(A->nrows, P, split.data());
        if (rc) return rc;
        // A greedy split always ends at nrows; guard anyway so every row is covered.
        split[P] = A->nrows;
        p->nparts = P;
        p->nsub = TPB / L;
        if ((rc = plan_upload(p, split.data(), split.size() * 4, (void **)&p->d_part))) return rc;
        p->launches = 1;
    } else {
        uint32_t I = p->opts.items_per_lane ? p->opts.items_per_lane : 8;
        if (I != 4 && I != 8 && I != 16) return SP_ERR_ARG;
        p->ipt = I;
        std::vector<uint32_t> trow, tnnz;
        merge_tiles(A, TPB * I, trow, tnnz);
        p->ntiles = (uint32_t)trow.size() - 1;
        p->nparts = p->ntiles;
        p->nsub = TPB;
        if ((rc = plan_upload(p, trow.data(), trow.size() * 4, (void **)&p->d_tile_row))) return rc;
        if ((rc = plan_upload(p, tnnz.data(), tnnz.size() * 4, (void **)&p->d_tile_pos))) return rc;
        p->carry_k = 8;
        const size_t asz = (A->dtype == SP_I64 || A->dtype == SP_F64) ? 8 : 4;
        if ((rc = plan_alloc(p, (size_t)std::max<uint32_t>(p->ntiles, 1) * 4, (void **)&p->d_carry_row))) return rc;
        if ((rc = plan_alloc(p, (size_t)std::max<uint32_t>(p->ntiles, 1) * p->carry_k * asz, &p->d_carry_val))) return rc;
        p->launches = 2;
    }
    return SP_OK;
}

int csr_run(sp_plan *p, const void *x, void *y, int k, int ldx, int ldy) {
    if (p->nrows == 0) return SP_OK;
    switch (p->dtype) {
    case SP_I8: return run_typed<int8_t>(p, x, y, k, ldx, ldy);
    case SP_I16: return run_typed<int16_t>(p, x, y, k, ldx, ldy);
    case SP_I32: return run_typed<int32_t>(p, x, y, k, ldx, ldy);
    case SP_I64: return run_typed<int64_t>(p, x, y, k, ldx, ldy);
    case SP_F32: return run_typed<float>(p, x, y, k, ldx, ldy);
    case SP_F64: return run_typed<double>(p, x, y, k, ldx, ldy);
    default: return SP_ERR_ARG;
    }
}
