// gfx950 SpMV / SpMM kernels over a list of sub-problems.
//
// Every kernel reads a per-workgroup problem id (unit_pid) and a problem
// descriptor (Prob), so one launch serves a 1D plan (one problem), a 2D
// tiling (one problem per tile, writing partial y) or a batch of
// independent SpMVs.  Work coordinates are stored per problem with one
// extra closing entry: unit u of problem p uses coord[u + p] .. coord[u + p + 1].
//
// Device replacements of SparseP's DPU tasklet kernels:
//   csr_merge   entry-balanced CSR (merge path over row ends + entries)
//   csr_vector  row-granular CSR, L lanes per row
//                 (spmv/1D/CSR-row/dpu/task.c:105-202)
//   coo_seg     entry-balanced COO with a segmented reduction; rows that
//               straddle tiles are carried to a fix-up pass
//                 (spmv/1D/COO-nnz/dpu/task.c:21-679, COO-row/dpu/task.c:31-492)
//   bcsr_vector register-blocked BCSR, L lanes per block row
//                 (spmv/1D/BCSR-nnz/dpu/task.c:231-380)
//   bcoo_seg    block-balanced BCOO with per-row segmented carries
//                 (spmv/1D/BCOO-nnz/dpu/task.c:22-894)
//   fixup       adds cross-tile carries in tile order and zeroes rows
//               without entries (replaces the host merges
//               spmv/1D/COO-nnz/host/app.c:275-289 and
//               spmv/1D/BCOO-block/host/app.c:296-313)
// All reductions have a fixed order, so results are run-to-run identical.
#pragma once

#include <type_traits>

#include "common.hpp"

namespace spd {

constexpr int TPB = 256;
constexpr uint32_t NONE = 0xFFFFFFFFu;

struct Prob {
    const uint32_t *ptr;   // CSR rowptr / BCSR browptr
    const uint32_t *rind;  // COO rowind / BCOO browind
    const uint32_t *cind;  // column / block-column index
    const void *val;       // values (blocks row-major r x c)
    uint32_t nrows;        // rows (CSR/COO) or block rows (BCSR/BCOO)
    uint32_t nent;         // entries or blocks
    uint32_t nrows_elem;   // element rows (block formats: rows of the last block row may be cut)
    uint32_t ncols_elem;   // element columns
    int64_t xoff, yoff;    // element offsets into the per-call x / y bases
    int32_t ldx, ldy;      // leading dimensions when x / y come from pointer tables
    uint32_t pbits;        // CSR merge tiles in column order: cind holds col << pbits | position in tile
    uint32_t nhcol;        // hot-x CSR: number of staged ("hot") columns
    const uint32_t *hcol;  // hot-x CSR: the staged columns, slot order
    void *xh;              // hot-x CSR: x at the staged columns, written before each launch
};

template <typename T> struct Ptrs {
    const T *xb;
    T *yb;
    const uintptr_t *xtab;  // per-problem x pointers, or null
    const uintptr_t *ytab;  // per-problem y pointers, or null
    int ldx, ldy;
    int q0;                 // first right-hand side of this launch (tables only)
    __device__ __forceinline__ const T *x(uint32_t pid, const Prob &P) const {
        return xtab ? (const T *)xtab[pid] + q0 : xb + P.xoff * ldx;
    }
    __device__ __forceinline__ T *y(uint32_t pid, const Prob &P) const {
        return ytab ? (T *)ytab[pid] + q0 : yb + P.yoff * ldy;
    }
    __device__ __forceinline__ int lx(const Prob &P) const { return xtab ? P.ldx : ldx; }
    __device__ __forceinline__ int ly(const Prob &P) const { return ytab ? P.ldy : ldy; }
};

// ---------------------------------------------------------------------------
// shared pieces
// ---------------------------------------------------------------------------

template <typename A> __device__ __forceinline__ A shfl_up_any(A v, int off) {
    if constexpr (sizeof(A) == 4) {
        return __builtin_bit_cast(A, __shfl_up(__builtin_bit_cast(int, v), off, WAVE));
    } else {
        uint64_t u = __builtin_bit_cast(uint64_t, v);
        int lo = __shfl_up((int)(uint32_t)u, off, WAVE);
        int hi = __shfl_up((int)(uint32_t)(u >> 32), off, WAVE);
        return __builtin_bit_cast(A, ((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo);
    }
}

// Workgroup-wide inclusive segmented scan of (key, val[K]) over NT threads;
// keys are non-decreasing in thread order.  On return val holds the scan
// and s_scan[q][tid] the same values for neighbours to read.
template <typename A, int K, int NT = TPB>
__device__ __forceinline__ void block_seg_scan(uint32_t key, A (&val)[K], uint32_t *s_key,
                                               A (*s_scan)[NT], A (*s_wc)[NT / WAVE]) {
    constexpr int NW = NT / WAVE;
    const int tid = threadIdx.x, lane = tid % WAVE, wid = tid / WAVE;
#pragma unroll
    for (int off = 1; off < WAVE; off <<= 1) {
        const uint32_t ok = (uint32_t)__shfl_up((int)key, off, WAVE);
#pragma unroll
        for (int q = 0; q < K; q++) {
            const A ov = shfl_up_any(val[q], off);
            if (lane >= off && ok == key) val[q] = ov + val[q];
        }
    }
    s_key[tid] = key;
#pragma unroll
    for (int q = 0; q < K; q++) s_scan[q][tid] = val[q];
    __syncthreads();
    if (tid < K) {
        const int q = tid;
        A run = A(0);
        uint32_t run_key = NONE;
        for (int w = 0; w < NW; w++) {
            s_wc[q][w] = run;
            const uint32_t lk = s_key[w * WAVE + WAVE - 1];
            const A lv = s_scan[q][w * WAVE + WAVE - 1];
            run = (w > 0 && lk == run_key) ? run + lv : lv;
            run_key = lk;
        }
    }
    __syncthreads();
    if (wid > 0 && key == s_key[wid * WAVE - 1]) {
#pragma unroll
        for (int q = 0; q < K; q++) val[q] = s_wc[q][wid] + val[q];
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < K; q++) s_scan[q][tid] = val[q];
    __syncthreads();
}

template <typename RE>
__device__ __host__ inline uint32_t merge_search(uint32_t d, uint32_t na, uint32_t nb, RE rowend,
                                                 uint32_t jbase) {
    uint32_t lo = d > nb ? d - nb : 0, hi = d < na ? d : na;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (rowend(mid) <= jbase + (d - mid - 1)) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// N contiguous values of x at p, which the caller knows to be aligned to the
// load width (16 or 8 bytes): wide loads without a runtime alignment check,
// so that no load sits under a branch.
template <typename T, int N> __device__ __forceinline__ void ld_al(gptr<const T> p, T (&out)[N]) {
    constexpr int B = N * (int)sizeof(T);
    if constexpr (B % 16 == 0) {
        u32x4 w[B / 16];
#pragma unroll
        for (int i = 0; i < B / 16; i++) w[i] = ((gptr<const u32x4>)p)[i];
        __builtin_memcpy(out, w, B);
    } else if constexpr (B == 8) {
        const u32x2 w = *(gptr<const u32x2>)p;
        __builtin_memcpy(out, &w, B);
    } else {
#pragma unroll
        for (int i = 0; i < N; i++) out[i] = p[i];
    }
}

// Four consecutive entries a .. a+3 of a (column index, value) stream (a a
// multiple of 4, both arrays aligned to 4 entries and readable 3 entries
// past their end): one 16-byte index load and one 4 * sizeof(T) value load.
template <typename T>
__device__ __forceinline__ void load4(gptr<const uint32_t> ci, gptr<const T> va, uint32_t a, uint32_t (&c)[4],
                                      T (&v)[4]) {
    const u32x4 w = __builtin_nontemporal_load((gptr<const u32x4>)(ci + a));
    c[0] = w[0];
    c[1] = w[1];
    c[2] = w[2];
    c[3] = w[3];
    ldg_n<T, 4>(va + a, v);
}

// LDS bytes for a staged x slice (2D tiles): three workgroups per CU fit
// next to the tile buffers.
constexpr int SLICE_BYTES = 32768;

// ---------------------------------------------------------------------------
// CSR, entry balanced (merge path).  SLICE (2D tiles, one right-hand side):
// each workgroup takes a contiguous run of units, copies the x slice of the
// current tile (its ncols_elem columns) into LDS when the tile changes, and
// gathers from LDS; the sums are the same as without the slice.
// ---------------------------------------------------------------------------
template <typename T, int IPT, int K, bool SLICE = false, int NT = TPB, bool CH = false>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(CH ? 6 : 1))) void csr_merge_kernel(const Prob *__restrict__ probs,
                                                        const uint32_t *__restrict__ unit_pid,
                                                        const uint32_t *__restrict__ crow,
                                                        const uint32_t *__restrict__ cnnz, Ptrs<T> io,
                                                        uint32_t *__restrict__ carry_row,
                                                        acc_t<T> *__restrict__ carry_val, uint32_t nunits) {
    using A = acc_t<T>;
    constexpr int TILE = NT * IPT;
    // One right-hand side: stage products (the x gather happens in the
    // coalesced load phase).  Several: stage (column, value) pairs and
    // gather x rows during the walk.
    constexpr bool PROD = (K == 1);
    __shared__ uint32_t s_rowend[TILE + 1];
    __shared__ A s_prod[PROD ? TILE : 1];
    __shared__ uint32_t s_col[PROD ? 1 : TILE];
    __shared__ T s_val[PROD ? 1 : TILE];
    __shared__ uint32_t s_key[NT];
    __shared__ A s_scan[K][NT];
    __shared__ A s_wc[K][NT / WAVE];
    __shared__ T s_x[SLICE ? SLICE_BYTES / (int)sizeof(T) : 1];

    const int tid = threadIdx.x;
    // Grid-stride over tiles: one tile per workgroup for a plain launch, a
    // persistent sweep when the grid is capped (batched launches); with a
    // slice, contiguous runs so that consecutive units share the slice.
    uint32_t t0 = blockIdx.x, t1 = nunits, tstep = gridDim.x;
    if constexpr (SLICE) {
        const uint32_t per = (nunits + gridDim.x - 1) / gridDim.x;
        t0 = min(nunits, blockIdx.x * per);
        t1 = min(nunits, t0 + per);
        tstep = 1;
    }
    [[maybe_unused]] uint32_t staged = NONE;
    for (uint32_t t = t0; t < t1; t += tstep) {
    // one-problem launches pass no unit_pid table: the descriptor and tile coordinates
    // are then loaded together instead of after the problem id
    const uint32_t pid = unit_pid ? unit_pid[t] : 0u, ci = t + pid;
    const Prob P = probs[pid];
    const gptr<const T> __restrict__ x = gp(io.x(pid, P));
    const gptr<T> __restrict__ y = gp(io.y(pid, P));
    [[maybe_unused]] const gptr<const uint32_t> __restrict__ rp = gp(P.ptr), rip = gp(P.rind), cip = gp(P.cind);
    const gptr<const T> __restrict__ vals = gp((const T *)P.val);
    const int ldx = io.lx(P), ldy = io.ly(P);
    if constexpr (SLICE) {
        if (pid != staged) {  // the previous unit ended with a barrier, so s_x is free
            for (uint32_t c = tid; c < P.ncols_elem; c += NT) s_x[c] = x[(size_t)c * ldx];
            __syncthreads();
            staged = pid;
        }
    }
    const uint32_t i0 = crow[ci], j0 = cnnz[ci];
    const uint32_t na = crow[ci + 1] - i0, nb = cnnz[ci + 1] - j0;
    // tiles stored in column order carry each entry's row-order position in the low bits
    const uint32_t pb = P.pbits, pmask = (1u << pb) - 1;

    // All loads of the tile are issued before the first is consumed, then
    // all x gathers: two memory round trips per tile.
    const uint32_t nrl = min(na + 1, P.nrows - i0);
    uint32_t re[IPT + 1];
#pragma unroll
    for (int q = 0; q <= IPT; q++) {
        const uint32_t k = tid + q * NT;
        if (k < nrl) re[q] = rp[i0 + k + 1];
    }
    // CH (1- and 2-byte values; the plan's arrays aligned to 4 entries and
    // padded past their end, as every 1D plan's): entries come in 4-entry
    // chunks aligned in the arrays, one 16-byte index load and one
    // 4 * sizeof(T) value load each, so the values stream by bytes.  Chunk c
    // covers array entries a0 + 4c .. a0 + 4c + 3, of which those in
    // [j0, j0 + nb) are the tile's; thread tid takes chunks tid + r NT, and
    // the last thread one more, afterwards, when the tile's entries straddle
    // a further chunk.  Six waves per SIMD are kept (80 registers).  (R-MAT
    // 2^20, kernel time: INT8 78.5 -> 66.7 us, INT16 85.5 -> 72.8 us; for 4-
    // and 8-byte values the lane-strided layout below is faster, 81 against
    // 97 us per FP32 SpMV, because one gather instruction then covers 64
    // consecutive entries of a column-ordered tile instead of every fourth
    // of 256.)  Otherwise: one entry per lane and load, entries tid + q NT.
    // Either way each product goes to its entry's slot.
    constexpr bool chunked = CH;
    auto store = [&](uint32_t e, uint32_t cw, T vw, T xw) {
        const uint32_t pos = pb ? (cw & pmask) : e;
        if constexpr (PROD) {
            s_prod[pos] = mul<T>(vw, xw);
        } else {
            s_col[pos] = cw >> pb;
            s_val[pos] = vw;
        }
    };
    auto gather = [&](uint32_t cw) -> T {
        if constexpr (PROD) return SLICE ? s_x[cw >> pb] : x[(size_t)(cw >> pb) * ldx];
        else return T(0);
    };
    if constexpr (chunked) {
        constexpr int R = IPT / 4;
        static_assert(IPT % 4 == 0, "whole chunks per thread");
        const uint32_t a0 = j0 & ~3u, lead = j0 - a0, nch = (lead + nb + 3) >> 2;
        uint32_t cv[R][4];
        T vv[R][4], xv[R][4];
        // No load under a branch (the compiler would wait for the earlier loads at each
        // branch's end, one gather in flight at a time): chunks past the tile's last load
        // it again, and entries outside the tile gather x[0].  (INT8 R-MAT 2^20 CSR: 68.6
        // against 75-81 us per SpMV; for the lane-strided path below the same change was
        // slower, 86 against 80 us FP32, and is not made there.)
        if (nb) {
#pragma unroll
            for (int r = 0; r < R; r++) load4<T>(cip, vals, a0 + 4 * min((uint32_t)(tid + r * NT), nch - 1), cv[r], vv[r]);
#pragma unroll
            for (int r = 0; r < R; r++)
#pragma unroll
                for (int i = 0; i < 4; i++) {
                    const uint32_t t4 = 4 * (tid + r * NT) + i;  // lead + tile entry
                    xv[r][i] = gather(t4 >= lead && t4 - lead < nb ? cv[r][i] : 0u);
                }
#pragma unroll
            for (int r = 0; r < R; r++)
#pragma unroll
                for (int i = 0; i < 4; i++) {
                    const uint32_t t4 = 4 * (tid + r * NT) + i;
                    if (t4 >= lead && t4 - lead < nb) store(t4 - lead, cv[r][i], vv[r][i], xv[r][i]);
                }
        }
        if (tid == NT - 1 && (uint32_t)(R * NT) < nch) {  // one more chunk: lead + nb > 4 R NT
            uint32_t cw[4];
            T vw[4];
            load4<T>(cip, vals, a0 + 4 * (uint32_t)(R * NT), cw, vw);
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const uint32_t t4 = 4 * (uint32_t)(R * NT) + i;
                if (t4 - lead < nb) store(t4 - lead, cw[i], vw[i], gather(cw[i]));
            }
        }
    } else {
        uint32_t cv[IPT];
        T vv[IPT], xv[IPT];
#pragma unroll
        for (int q = 0; q < IPT; q++) {
            const uint32_t k = tid + q * NT;
            if (k < nb) {
                cv[q] = ldg(cip + j0 + k);
                vv[q] = ldg(vals + j0 + k);
            }
        }
#pragma unroll
        for (int q = 0; q < IPT; q++)
            if (tid + q * NT < nb) xv[q] = gather(cv[q]);
#pragma unroll
        for (int q = 0; q < IPT; q++)
            if (tid + q * NT < nb) store(tid + q * NT, cv[q], vv[q], xv[q]);
    }
#pragma unroll
    for (int q = 0; q <= IPT; q++)
        if (tid + q * NT < nrl) s_rowend[tid + q * NT] = re[q];
    __syncthreads();

    const uint32_t total = na + nb;
    const uint32_t d0 = min((uint32_t)tid * IPT, total), d1 = min(d0 + IPT, total);
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
                T xs[K];
                ld_n<T, K>(x + (size_t)c * ldx, xs);  // the K values of x's row c, wide when aligned
#pragma unroll
                for (int q = 0; q < K; q++) acc[q] = acc[q] + mul<T>(v, xs[q]);
            }
            lj++;
        }
    }
    const uint32_t key = i0 + li;
    block_seg_scan<A, K, NT>(key, acc, s_key, s_scan, s_wc);
    if (emitted) {
#pragma unroll
        for (int q = 0; q < K; q++) {
            const A cin = tid > 0 ? s_scan[q][tid - 1] : A(0);
            y[(size_t)first_row * ldy + q] = to_val<T>(cin + first[q]);
        }
    }
    if (tid == NT - 1) {
        carry_row[t] = key < P.nrows ? key : NONE;
#pragma unroll
        for (int q = 0; q < K; q++) carry_val[(size_t)t * K + q] = acc[q];
    }
    __syncthreads();  // LDS is reused by the next tile
    }
}

// ---------------------------------------------------------------------------
// CSR, merge path with the most used x entries staged in LDS.  The plan picks
// up to Hot<T>::SLOTS "hot" columns (the most frequent ones) and stores every
// merge tile in column order with the hot entries first: entry k of a tile is
// hot iff k < chot[tile], and its index word then holds the LDS slot of its
// column instead of the column.  hot_gather_kernel copies x at the hot
// columns into a contiguous buffer before each launch; each 1024-thread
// workgroup (one per CU) copies that buffer into LDS once and then sweeps
// tiles persistently, so hot gathers are LDS reads and only cold entries
// gather from L2.  Products go to their row-order slots and the segmented
// sums run as in csr_merge_kernel.
// ---------------------------------------------------------------------------
constexpr int HOT_TPB = 1024;
constexpr int HOT_IPT = 4;
template <typename T> struct Hot {
    // LDS left after the tile buffers (row ends, products, scan) of 4 items per lane
    static constexpr int BYTES = sizeof(acc_t<T>) == 8 ? 96 * 1024 : 112 * 1024;
    static constexpr int SLOTS = BYTES / (int)sizeof(T);
};

template <typename T>
__global__ __launch_bounds__(TPB) void hot_gather_kernel(const Prob *__restrict__ probs, Ptrs<T> io) {
    const uint32_t pid = blockIdx.y;
    const Prob P = probs[pid];
    const gptr<const T> __restrict__ x = gp(io.x(pid, P));
    const gptr<const uint32_t> __restrict__ hc = gp(P.hcol);
    const gptr<T> __restrict__ xh = gp((T *)P.xh);
    const int ldx = io.lx(P);
    for (uint32_t s = blockIdx.x * TPB + threadIdx.x; s < P.nhcol; s += gridDim.x * TPB)
        xh[s] = x[(size_t)ldg(hc + s) * ldx];
}

template <typename T>
__global__ __launch_bounds__(HOT_TPB) void csr_hot_kernel(const Prob *__restrict__ probs,
                                                          const uint32_t *__restrict__ unit_pid,
                                                          const uint32_t *__restrict__ crow,
                                                          const uint32_t *__restrict__ cnnz,
                                                          const uint32_t *__restrict__ chot, Ptrs<T> io,
                                                          uint32_t *__restrict__ carry_row,
                                                          acc_t<T> *__restrict__ carry_val, uint32_t nunits) {
    using A = acc_t<T>;
    constexpr int NT = HOT_TPB, IPT = HOT_IPT, TILE = NT * IPT;
    constexpr int H = Hot<T>::SLOTS;
    __shared__ __align__(16) T s_x[H];
    __shared__ uint32_t s_rowend[TILE + 1];
    __shared__ A s_prod[TILE];
    __shared__ uint32_t s_key[NT];
    __shared__ A s_scan[1][NT];
    __shared__ A s_wc[1][NT / WAVE];

    const int tid = threadIdx.x;
    // One workgroup per CU leaves no other workgroup to hide a tile's load
    // latency behind, so the next tile's row ends, indices and values are
    // loaded while the current tile is walked and scanned.
    struct Tile { uint32_t pid, i0, j0, na, nb, nh, nrl; };
    auto fetch = [&](uint32_t tt, Tile &ti, uint32_t (&re)[IPT + 1], uint32_t (&cv)[IPT], T (&vv)[IPT]) {
        ti.pid = unit_pid[tt];
        const uint32_t ci = tt + ti.pid;
        const Prob &Q = probs[ti.pid];
        ti.i0 = crow[ci];
        ti.j0 = cnnz[ci];
        ti.nh = chot[ci];
        ti.na = crow[ci + 1] - ti.i0;
        ti.nb = cnnz[ci + 1] - ti.j0;
        ti.nrl = min(ti.na + 1, Q.nrows - ti.i0);
        const gptr<const uint32_t> __restrict__ rp = gp(Q.ptr), cip = gp(Q.cind);
        const gptr<const T> __restrict__ vals = gp((const T *)Q.val);
#pragma unroll
        for (int q = 0; q <= IPT; q++) {
            const uint32_t k = tid + q * NT;
            if (k < ti.nrl) re[q] = rp[ti.i0 + k + 1];
        }
#pragma unroll
        for (int q = 0; q < IPT; q++) {
            const uint32_t k = tid + q * NT;
            if (k < ti.nb) {
                cv[q] = ldg(cip + ti.j0 + k);
                vv[q] = ldg(vals + ti.j0 + k);
            }
        }
    };
    Tile cur{};
    uint32_t re[IPT + 1], cv[IPT];
    T vv[IPT];
    if (blockIdx.x < nunits) fetch(blockIdx.x, cur, re, cv, vv);
    uint32_t staged = NONE;  // problem whose hot x is in LDS
    for (uint32_t t = blockIdx.x; t < nunits; t += gridDim.x) {
    const uint32_t pid = cur.pid;
    const Prob P = probs[pid];
    if (pid != staged) {
        // the previous tile ended with a barrier, so s_x is free; the copy is
        // 16-byte chunks of the (padded) staged buffer
        const gptr<const u32x4> __restrict__ src = (gptr<const u32x4>)gp((const T *)P.xh);
        const uint32_t nchunk = (P.nhcol * (uint32_t)sizeof(T) + 15) / 16;
        for (uint32_t c = tid; c < nchunk; c += NT) ((u32x4 *)s_x)[c] = src[c];
        staged = pid;
    }
    const gptr<const T> __restrict__ x = gp(io.x(pid, P));
    const gptr<T> __restrict__ y = gp(io.y(pid, P));
    const int ldx = io.lx(P), ldy = io.ly(P);
    const uint32_t i0 = cur.i0, j0 = cur.j0, na = cur.na, nb = cur.nb, nh = cur.nh, nrl = cur.nrl;
    const uint32_t pb = P.pbits, pmask = (1u << pb) - 1;

    // cold gathers first (global), then the staging copy must be visible for the LDS reads
    T xv[IPT];
#pragma unroll
    for (int q = 0; q < IPT; q++) {
        const uint32_t k = tid + q * NT;
        if (k < nb && k >= nh) xv[q] = x[(size_t)(cv[q] >> pb) * ldx];
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < IPT; q++) {
        const uint32_t k = tid + q * NT;
        if (k < nh) xv[q] = s_x[cv[q] >> pb];
    }
#pragma unroll
    for (int q = 0; q < IPT; q++)
        if (tid + q * NT < nb) s_prod[cv[q] & pmask] = mul<T>(vv[q], xv[q]);
#pragma unroll
    for (int q = 0; q <= IPT; q++)
        if (tid + q * NT < nrl) s_rowend[tid + q * NT] = re[q];
    __syncthreads();

    // the next tile's loads are in flight during the walk and the scan
    Tile nxt{};
    uint32_t re2[IPT + 1], cv2[IPT];
    T vv2[IPT];
    if (t + gridDim.x < nunits) fetch(t + gridDim.x, nxt, re2, cv2, vv2);

    const uint32_t total = na + nb;
    const uint32_t d0 = min((uint32_t)tid * IPT, total), d1 = min(d0 + IPT, total);
    uint32_t li = merge_search(d0, na, nb, [&](uint32_t i) { return s_rowend[i]; }, j0);
    uint32_t lj = d0 - li;
    A acc[1] = {A(0)}, first = A(0);
    bool emitted = false;
    uint32_t first_row = 0;
    for (uint32_t s = d0; s < d1; s++) {
        if (li < na && s_rowend[li] <= j0 + lj) {
            const uint32_t row = i0 + li;
            if (!emitted) {
                emitted = true;
                first_row = row;
                first = acc[0];
            } else {
                y[(size_t)row * ldy] = to_val<T>(acc[0]);
            }
            acc[0] = A(0);
            li++;
        } else {
            acc[0] = acc[0] + s_prod[lj];
            lj++;
        }
    }
    const uint32_t key = i0 + li;
    block_seg_scan<A, 1, NT>(key, acc, s_key, s_scan, s_wc);
    if (emitted) {
        const A cin = tid > 0 ? s_scan[0][tid - 1] : A(0);
        y[(size_t)first_row * ldy] = to_val<T>(cin + first);
    }
    if (tid == NT - 1) {
        carry_row[t] = key < P.nrows ? key : NONE;
        carry_val[t] = acc[0];
    }
    __syncthreads();  // LDS tile buffers are reused by the next tile
    cur = nxt;
#pragma unroll
    for (int q = 0; q <= IPT; q++) re[q] = re2[q];
#pragma unroll
    for (int q = 0; q < IPT; q++) {
        cv[q] = cv2[q];
        vv[q] = vv2[q];
    }
    }
}

// ---------------------------------------------------------------------------
// CSR, row granular: L lanes per row, rows of the unit interleaved over
// the TPB/L lane groups.
// ---------------------------------------------------------------------------
template <typename T, int L, int K>
__global__ __launch_bounds__(TPB) void csr_vector_kernel(const Prob *__restrict__ probs,
                                                         const uint32_t *__restrict__ unit_pid,
                                                         const uint32_t *__restrict__ urow, Ptrs<T> io,
                                                         uint32_t nunits) {
    using A = acc_t<T>;
    constexpr int G = TPB / L;
    for (uint32_t u = blockIdx.x; u < nunits; u += gridDim.x) {
    const uint32_t pid = unit_pid[u];
    const Prob P = probs[pid];
    const gptr<const T> __restrict__ x = gp(io.x(pid, P));
    const gptr<T> __restrict__ y = gp(io.y(pid, P));
    [[maybe_unused]] const gptr<const uint32_t> __restrict__ rp = gp(P.ptr), rip = gp(P.rind), cip = gp(P.cind);
    const gptr<const T> __restrict__ vals = gp((const T *)P.val);
    const gptr<const uint32_t> __restrict__ colind = cip;
    const int ldx = io.lx(P), ldy = io.ly(P);
    const uint32_t r0 = urow[u + pid], r1 = urow[u + pid + 1];
    const int lane = threadIdx.x % L, grp = threadIdx.x / L;
    for (uint32_t r = r0 + grp; r < r1; r += G) {
        const uint32_t b = rp[r], e = rp[r + 1];
        A acc[K];
#pragma unroll
        for (int q = 0; q < K; q++) acc[q] = A(0);
        uint32_t k = b + lane;
        // four entries per lane in flight (index and value loads, then the four gathers)
        // before any is consumed: long rows are otherwise bound by load latency
        if constexpr (K == 1) {
            for (; k + 3 * L < e; k += 4 * L) {
                uint32_t c[4];
                T v[4], xv[4];
#pragma unroll
                for (int j = 0; j < 4; j++) { c[j] = ldg(colind + k + j * L); v[j] = ldg(vals + k + j * L); }
#pragma unroll
                for (int j = 0; j < 4; j++) xv[j] = x[(size_t)c[j] * ldx];
#pragma unroll
                for (int j = 0; j < 4; j++) acc[0] = fma_acc<T>(v[j], xv[j], acc[0]);
            }
        }
        for (; k + L < e; k += 2 * L) {
            const uint32_t c0 = ldg(colind + k), c1 = ldg(colind + k + L);
            const T v0 = ldg(vals + k), v1 = ldg(vals + k + L);
            T x0[K], x1[K];  // rows c0 and c1 of x, wide loads when aligned
            ld_n<T, K>(x + (size_t)c0 * ldx, x0);
            ld_n<T, K>(x + (size_t)c1 * ldx, x1);
#pragma unroll
            for (int q = 0; q < K; q++) {
                acc[q] = fma_acc<T>(v0, x0[q], acc[q]);
                acc[q] = fma_acc<T>(v1, x1[q], acc[q]);
            }
        }
        if (k < e) {
            const uint32_t c0 = ldg(colind + k);
            const T v0 = ldg(vals + k);
            T x0[K];
            ld_n<T, K>(x + (size_t)c0 * ldx, x0);
#pragma unroll
            for (int q = 0; q < K; q++) acc[q] = fma_acc<T>(v0, x0[q], acc[q]);
        }
#pragma unroll
        for (int q = 0; q < K; q++) acc[q] = group_sum<L>(acc[q]);
        if (lane == 0) {
#pragma unroll
            for (int q = 0; q < K; q++) y[(size_t)r * ldy + q] = to_val<T>(acc[q]);
        }
    }
    }
}

// ---------------------------------------------------------------------------
// CSR, row granular, one right-hand side, 4-entry chunks: L lanes per row,
// each lane taking whole 4-entry chunks aligned in the arrays (one 16-byte
// index load and one 4 * sizeof(T) value load per chunk; the plan's arrays
// are aligned to 4 entries and readable 3 entries past their end).  A row's
// chunks start at its first entry rounded down to a multiple of 4; entries
// of a chunk outside the row add nothing and gather x[0].  Two chunks per
// lane are in flight in the main loop; no load sits under a branch.  The
// row loop of spmv/1D/CSR-row/dpu/task.c:105-202 with four entries per lane
// and load.  (27-point stencil 64^3 x 4, FP32: 245 us with 32 lanes against
// 301-363 us for one entry per lane and 303 us for the merge-path kernel.)
// Q (quads): the plan found every row's length a multiple of 4 and every
// chunk's columns c, c+1, c+2, c+3 with c a multiple of 4 (rows made of
// aligned runs of four columns, as in a stencil with 4 unknowns per point);
// a chunk then reads its four x values in one load instead of four gathers
// when x is contiguous and aligned to four values.
// ---------------------------------------------------------------------------
template <typename T, int L, bool Q = false>
__global__ __launch_bounds__(TPB) void csr_chunk_kernel(const Prob *__restrict__ probs,
                                                        const uint32_t *__restrict__ unit_pid,
                                                        const uint32_t *__restrict__ urow, Ptrs<T> io,
                                                        uint32_t nunits) {
    using A = acc_t<T>;
    constexpr int G = TPB / L;
    for (uint32_t u = blockIdx.x; u < nunits; u += gridDim.x) {
    const uint32_t pid = unit_pid[u];
    const Prob P = probs[pid];
    const gptr<const T> __restrict__ x = gp(io.x(pid, P));
    const gptr<T> __restrict__ y = gp(io.y(pid, P));
    const gptr<const uint32_t> __restrict__ rp = gp(P.ptr), cip = gp(P.cind);
    const gptr<const T> __restrict__ vals = gp((const T *)P.val);
    const int ldx = io.lx(P), ldy = io.ly(P);
    const uint32_t r0 = urow[u + pid], r1 = urow[u + pid + 1];
    const int lane = threadIdx.x % L, grp = threadIdx.x / L;
    if constexpr (Q) {
        typedef T T4 __attribute__((ext_vector_type(4)));
        if (ldx == 1 && (uintptr_t)io.x(pid, P) % sizeof(T4) == 0) {
            // a chunk lies inside its row or outside it; one outside reads x[0 .. 3] and adds 0
            auto quad = [&](uint32_t k, uint32_t b, uint32_t e, A acc) {
                uint32_t c[4];
                T v[4];
                load4<T>(cip, vals, k, c, v);
                const bool ok = k >= b && k < e;
                const T4 xv = *(gptr<const T4>)(x + (ok ? c[0] : 0u));
#pragma unroll
                for (int i = 0; i < 4; i++) acc = fma_acc<T>(ok ? v[i] : T(0), xv[i], acc);
                return acc;
            };
            for (uint32_t r = r0 + grp; r < r1; r += G) {
                const uint32_t b = rp[r], e = rp[r + 1];
                A acc = A(0);
                uint32_t k = b + 4 * (uint32_t)lane;
                for (; k + 4 * L < e; k += 8 * L) {  // two chunks per lane in flight, both inside the row
                    uint32_t c[2][4];
                    T v[2][4];
                    load4<T>(cip, vals, k, c[0], v[0]);
                    load4<T>(cip, vals, k + 4 * L, c[1], v[1]);
                    const T4 x0 = *(gptr<const T4>)(x + c[0][0]), x1 = *(gptr<const T4>)(x + c[1][0]);
#pragma unroll
                    for (int i = 0; i < 4; i++) acc = fma_acc<T>(v[0][i], x0[i], acc);
#pragma unroll
                    for (int i = 0; i < 4; i++) acc = fma_acc<T>(v[1][i], x1[i], acc);
                }
                if (k < e) acc = quad(k, b, e, acc);
                acc = group_sum<L>(acc);
                if (lane == 0) y[(size_t)r * ldy] = to_val<T>(acc);
            }
            continue;
        }
    }
    // chunk at k: its entries' products, those outside [b, e) zero
    auto chunk = [&](uint32_t k, uint32_t b, uint32_t e, A acc) {
        uint32_t c[4];
        T v[4], xv[4];
        load4<T>(cip, vals, k, c, v);
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const bool ok = k + i >= b && k + i < e;
            xv[i] = x[(size_t)(ok ? c[i] : 0u) * ldx];
            if (!ok) v[i] = T(0);
        }
#pragma unroll
        for (int i = 0; i < 4; i++) acc = fma_acc<T>(v[i], xv[i], acc);
        return acc;
    };
    for (uint32_t r = r0 + grp; r < r1; r += G) {
        const uint32_t b = rp[r], e = rp[r + 1];
        A acc = A(0);
        if (e > b) {
            uint32_t k = (b & ~3u) + 4 * (uint32_t)lane;
            for (; k + 4 * L < e; k += 8 * L) {  // two chunks per lane in flight
                uint32_t c[2][4];
                T v[2][4], xv[2][4];
                load4<T>(cip, vals, k, c[0], v[0]);
                load4<T>(cip, vals, k + 4 * L, c[1], v[1]);
#pragma unroll
                for (int h = 0; h < 2; h++)
#pragma unroll
                    for (int i = 0; i < 4; i++) {
                        const uint32_t kk = k + h * 4 * L + i;
                        const bool ok = kk >= b && kk < e;
                        xv[h][i] = x[(size_t)(ok ? c[h][i] : 0u) * ldx];
                        if (!ok) v[h][i] = T(0);
                    }
#pragma unroll
                for (int h = 0; h < 2; h++)
#pragma unroll
                    for (int i = 0; i < 4; i++) acc = fma_acc<T>(v[h][i], xv[h][i], acc);
            }
            if (k < e) acc = chunk(k, b, e, acc);
        }
        acc = group_sum<L>(acc);
        if (lane == 0) y[(size_t)r * ldy] = to_val<T>(acc);
    }
    }
}

// ---------------------------------------------------------------------------
// CSR, row granular, K >= 8 right-hand sides: lanes across the right-hand
// sides.  A wave takes a row; lane l = GS*g + r works on columns V r ..
// V r + V - 1 (one 16-byte piece of an x row, V = 16 / sizeof(T), GS = K / V
// lanes per entry) of entry slot g (G = 64 / GS entries of the row per
// step).  The row's column indices and values are read 64 at a time, one
// entry per lane, and handed out by shuffles, so each step is one coalesced
// read of G rows of x and V multiply-adds per lane; the G partial sums are
// added across lanes at the end of the row.  The row loop of the reference
// tasklet (spmv/1D/CSR-row/dpu/task.c:105-202) with the right-hand sides
// spread over lanes instead of entries, so no per-column reduction over many
// lanes is needed.  (One column per lane, 4-byte x loads, took twice as long:
// 27-point stencil, FP32, 32 columns: 1648 against 832 us.)  Used where a row
// of x is 64 bytes or more (K * sizeof(T) >= 64); narrower rows keep
// csr_vector_kernel, whose lanes read whole x rows in 16-byte pieces (FP32,
// 8 columns on the 27-point stencil: 532 against 667 us).
// ---------------------------------------------------------------------------
template <typename T, int K>
__global__ __launch_bounds__(TPB) void csr_rhs_kernel(const Prob *__restrict__ probs,
                                                      const uint32_t *__restrict__ unit_pid,
                                                      const uint32_t *__restrict__ urow, Ptrs<T> io,
                                                      uint32_t nunits) {
    static_assert(K == 8 || K == 16 || K == 32, "8, 16 or 32 right-hand sides per pass");
    using A = acc_t<T>;
    constexpr int V = 16 / (int)sizeof(T);  // columns per lane: one 16-byte piece of an x row
    static_assert(K % V == 0, "whole 16-byte pieces per x row");
    constexpr int GS = K / V;               // lanes per entry
    constexpr int NW = TPB / WAVE, G = WAVE / GS, U = 4;
    const int lane = threadIdx.x % WAVE, wid = threadIdx.x / WAVE;
    const int g = lane / GS, r4 = lane % GS;
    for (uint32_t u = blockIdx.x; u < nunits; u += gridDim.x) {
    const uint32_t pid = unit_pid[u];
    const Prob P = probs[pid];
    const gptr<const T> __restrict__ x = gp(io.x(pid, P));
    const gptr<T> __restrict__ y = gp(io.y(pid, P));
    const gptr<const uint32_t> __restrict__ rp = gp(P.ptr), colind = gp(P.cind);
    const gptr<const T> __restrict__ vals = gp((const T *)P.val);
    const int ldx = io.lx(P), ldy = io.ly(P);
    const uint32_t r0 = urow[u + pid], r1 = urow[u + pid + 1];
    // whole 16-byte pieces of x and y rows (uniform over the launch); the x case is a
    // template argument so that no load sits under a branch
    const bool xw = ((uintptr_t)io.x(pid, P) % 16 == 0) && (ldx % V == 0);
    const bool yw = ((uintptr_t)io.y(pid, P) % 16 == 0) && (ldy % V == 0);
    auto rows = [&](auto xw_c) {
    constexpr bool XW = decltype(xw_c)::value;
    // This wave's rows r0 + wid + NW t: their entry ranges in one load (64 rows a
    // round), and the next row's first 64 entries loaded while a row runs, so a
    // row starts with its indices and values at hand.
    for (uint32_t R0 = r0 + wid; R0 < r1; R0 += NW * WAVE) {
    const uint32_t nrw = min((uint32_t)WAVE, (r1 - R0 + NW - 1) / NW);
    const uint32_t Il = R0 + NW * (uint32_t)lane;
    const uint32_t bl = (uint32_t)lane < nrw ? rp[Il] : 0u, el = (uint32_t)lane < nrw ? rp[Il + 1] : 0u;
    auto first = [&](uint32_t t, uint32_t &cn, T &vn) {  // row t's first batch, one entry per lane
        const uint32_t bb = (uint32_t)__builtin_amdgcn_readlane((int)bl, (int)t);
        const uint32_t ee = (uint32_t)__builtin_amdgcn_readlane((int)el, (int)t);
        const uint32_t nn = min((uint32_t)WAVE, ee - bb);
        cn = (uint32_t)lane < nn ? ldg(colind + bb + lane) : 0u;
        vn = (uint32_t)lane < nn ? ldg(vals + bb + lane) : T(0);
    };
    uint32_t cn;
    T vn;
    first(0, cn, vn);
    for (uint32_t t = 0; t < nrw; t++) {
        const uint32_t r = R0 + NW * t;
        const uint32_t b = (uint32_t)__builtin_amdgcn_readlane((int)bl, (int)t);
        const uint32_t e = (uint32_t)__builtin_amdgcn_readlane((int)el, (int)t);
        uint32_t cl = cn;
        T vl = vn;
        if (t + 1 < nrw) first(t + 1, cn, vn);
        A acc[V];
#pragma unroll
        for (int j = 0; j < V; j++) acc[j] = A(0);
        for (uint32_t base = b; base < e; base += WAVE) {
            const uint32_t nb = min((uint32_t)WAVE, e - base);
            if (base != b) {  // rows of more than 64 entries
                cl = (uint32_t)lane < nb ? ldg(colind + base + lane) : 0u;
                vl = (uint32_t)lane < nb ? ldg(vals + base + lane) : T(0);
            }
            for (uint32_t t0 = 0; t0 < nb; t0 += G * U) {
                T v[U], xv[U][V];
#pragma unroll
                for (int w = 0; w < U; w++) {
                    const uint32_t s = t0 + w * G + g;
                    const bool act = s < nb;
                    // every lane takes part in the shuffles (a lane reads no value from an
                    // inactive one); an idle slot reads entry 0's x row and adds 0 times it
                    const int src = act ? (int)s : 0;
                    const uint32_t c = (uint32_t)__shfl((int)cl, src, WAVE);
                    const T vs = (T)__shfl(vl, src, WAVE);
                    v[w] = act ? vs : T(0);
                    const gptr<const T> px = x + (size_t)c * ldx + V * r4;
                    if constexpr (XW) {
                        const u32x4 p4 = *(gptr<const u32x4>)px;
                        __builtin_memcpy(xv[w], &p4, 16);
                    } else {
#pragma unroll
                        for (int j = 0; j < V; j++) xv[w][j] = px[j];
                    }
                }
#pragma unroll
                for (int w = 0; w < U; w++)
#pragma unroll
                    for (int j = 0; j < V; j++) acc[j] = fma_acc<T>(v[w], xv[w][j], acc[j]);
            }
        }
#pragma unroll
        for (int j = 0; j < V; j++) acc[j] = xor_sum<GS, WAVE>(acc[j]);
        if (g == 0) {
            const gptr<T> py = y + (size_t)r * ldy + V * r4;
            T o[V];
#pragma unroll
            for (int j = 0; j < V; j++) o[j] = to_val<T>(acc[j]);
            if (yw) {
                u32x4 p4;
                __builtin_memcpy(&p4, o, 16);
                *(gptr<u32x4>)py = p4;
            } else {
#pragma unroll
                for (int j = 0; j < V; j++) py[j] = o[j];
            }
        }
    }
    }
    };
    if (xw) rows(std::true_type{});
    else rows(std::false_type{});
    }
}

// ---------------------------------------------------------------------------
// COO, entry balanced with a segmented reduction.  Entries are sorted by
// row.  A thread owns IPT consecutive entries; a row is written by the
// thread holding its last entry, with the partial sums of earlier threads
// of the tile added in; partials of earlier tiles arrive in the fix-up.
// QD (with CH, row-ordered tiles): the plan found every group of four
// entries in the arrays to be one row's columns c .. c+3 with c a multiple
// of 4; a chunk then reads its four x values in one load when x is
// contiguous and aligned to four values.
// ---------------------------------------------------------------------------
template <typename T, int IPT, int K, bool SLICE = false, int NT = TPB, bool ATOM = false, bool CH = false,
          bool QD = false>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(CH ? 6 : 1))) void coo_seg_kernel(const Prob *__restrict__ probs,
                                                      const uint32_t *__restrict__ unit_pid,
                                                      const uint32_t *__restrict__ cent, Ptrs<T> io,
                                                      uint32_t *__restrict__ carry_row,
                                                      acc_t<T> *__restrict__ carry_val, uint32_t nunits) {
    using A = acc_t<T>;
    constexpr int TILE = NT * IPT;
    constexpr bool PROD = (K == 1);
    __shared__ uint32_t s_row[TILE];
    __shared__ A s_prod[PROD ? TILE : 1];
    __shared__ uint32_t s_col[PROD ? 1 : TILE];
    __shared__ T s_val[PROD ? 1 : TILE];
    __shared__ uint32_t s_key[NT];
    __shared__ A s_scan[K][NT];
    __shared__ A s_wc[K][NT / WAVE];
    __shared__ T s_x[SLICE ? SLICE_BYTES / (int)sizeof(T) : 1];

    const int tid = threadIdx.x;
    // as csr_merge_kernel: grid-stride, or contiguous runs sharing a staged x slice
    uint32_t t0 = blockIdx.x, t1 = nunits, tstep = gridDim.x;
    if constexpr (SLICE) {
        const uint32_t per = (nunits + gridDim.x - 1) / gridDim.x;
        t0 = min(nunits, blockIdx.x * per);
        t1 = min(nunits, t0 + per);
        tstep = 1;
    }
    [[maybe_unused]] uint32_t staged = NONE;
    for (uint32_t t = t0; t < t1; t += tstep) {
    const uint32_t pid = unit_pid[t], ci = t + pid;
    const Prob P = probs[pid];
    const gptr<const T> __restrict__ x = gp(io.x(pid, P));
    const gptr<T> __restrict__ y = gp(io.y(pid, P));
    [[maybe_unused]] const gptr<const uint32_t> __restrict__ rp = gp(P.ptr), rip = gp(P.rind), cip = gp(P.cind);
    const gptr<const T> __restrict__ vals = gp((const T *)P.val);
    const int ldx = io.lx(P), ldy = io.ly(P);
    if constexpr (SLICE) {
        if (pid != staged) {  // the previous unit ended with a barrier, so s_x is free
            for (uint32_t c = tid; c < P.ncols_elem; c += NT) s_x[c] = x[(size_t)c * ldx];
            __syncthreads();
            staged = pid;
        }
    }
    const uint32_t e0 = cent[ci], n = cent[ci + 1] - e0;
    const uint32_t next_row = (e0 + n < P.nent) ? rip[e0 + n] : NONE;
    // tiles stored in column order carry each entry's row-order position in the low bits
    const uint32_t pb = P.pbits, pmask = (1u << pb) - 1;

    // tile entry e: its row to s_row[e], its product (or column and value)
    // to its row-order slot
    auto store = [&](uint32_t e, uint32_t rw, uint32_t cw, T vw, T xw) {
        const uint32_t pos = pb ? (cw & pmask) : e;
        if constexpr (PROD) {
            s_prod[pos] = mul<T>(vw, xw);
        } else {
            s_col[pos] = cw >> pb;
            s_val[pos] = vw;
        }
        s_row[e] = rw;
    };
    auto gather = [&](uint32_t cw) -> T {
        if constexpr (PROD) return SLICE ? s_x[cw >> pb] : x[(size_t)(cw >> pb) * ldx];
        else return T(0);
    };
    if constexpr (CH) {
        // the plan's arrays aligned to 4 entries and padded past their end: 4-entry
        // chunks aligned in the arrays (16-byte row and column loads, one
        // 4 * sizeof(T) value load), as in csr_merge_kernel; used for 1- and
        // 2-byte values (R-MAT 2^20: INT8 83.6 -> 68.8 us, INT16 90.3 -> 79.0 us)
        constexpr int R = IPT / 4;
        static_assert(IPT % 4 == 0, "whole chunks per thread");
        const uint32_t a0 = e0 & ~3u, lead = e0 - a0, nch = (lead + n + 3) >> 2;
        uint32_t rv[R][4], cv[R][4];
        T vv[R][4], xv[R][4];
        if (n) {  // no load under a branch, as in csr_merge_kernel's chunked path
#pragma unroll
            for (int r = 0; r < R; r++) {
                const uint32_t c = min((uint32_t)(tid + r * NT), nch - 1);
                const u32x4 w = __builtin_nontemporal_load((gptr<const u32x4>)(rip + a0 + 4 * c));
                rv[r][0] = w[0];
                rv[r][1] = w[1];
                rv[r][2] = w[2];
                rv[r][3] = w[3];
                load4<T>(cip, vals, a0 + 4 * c, cv[r], vv[r]);
            }
            typedef T T4 __attribute__((ext_vector_type(4)));
            if (QD && ldx == 1 && (uintptr_t)io.x(pid, P) % sizeof(T4) == 0) {
                // every chunk of the arrays is an aligned run of four columns of one row
#pragma unroll
                for (int r = 0; r < R; r++) {
                    const T4 q = *(gptr<const T4>)(x + cv[r][0]);
#pragma unroll
                    for (int i = 0; i < 4; i++) xv[r][i] = q[i];
                }
            } else {
#pragma unroll
                for (int r = 0; r < R; r++)
#pragma unroll
                    for (int i = 0; i < 4; i++) {
                        const uint32_t t4 = 4 * (tid + r * NT) + i;  // lead + tile entry
                        xv[r][i] = gather(t4 >= lead && t4 - lead < n ? cv[r][i] : 0u);
                    }
            }
#pragma unroll
            for (int r = 0; r < R; r++)
#pragma unroll
                for (int i = 0; i < 4; i++) {
                    const uint32_t t4 = 4 * (tid + r * NT) + i;
                    if (t4 >= lead && t4 - lead < n) store(t4 - lead, rv[r][i], cv[r][i], vv[r][i], xv[r][i]);
                }
        }
        if (tid == NT - 1 && (uint32_t)(R * NT) < nch) {  // one more chunk: lead + n > 4 R NT
            const uint32_t a = a0 + 4 * (uint32_t)(R * NT);
            const u32x4 w = __builtin_nontemporal_load((gptr<const u32x4>)(rip + a));
            uint32_t cw[4];
            T vw[4];
            load4<T>(cip, vals, a, cw, vw);
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const uint32_t t4 = 4 * (uint32_t)(R * NT) + i;
                if (t4 - lead < n) store(t4 - lead, w[i], cw[i], vw[i], gather(cw[i]));
            }
        }
    } else {
        uint32_t rv[IPT], cv[IPT];
        T vv[IPT], xv[IPT];
#pragma unroll
        for (int q = 0; q < IPT; q++) {
            const uint32_t k = tid + q * NT;
            if (k < n) {
                rv[q] = ldg(rip + e0 + k);
                cv[q] = ldg(cip + e0 + k);
                vv[q] = ldg(vals + e0 + k);
            }
        }
#pragma unroll
        for (int q = 0; q < IPT; q++)
            if (tid + q * NT < n) xv[q] = gather(cv[q]);
#pragma unroll
        for (int q = 0; q < IPT; q++)
            if (tid + q * NT < n) store(tid + q * NT, rv[q], cv[q], vv[q], xv[q]);
    }
    __syncthreads();

    const uint32_t a = min((uint32_t)tid * IPT, n), b = min(a + IPT, n);
    A acc[K], first[K];
#pragma unroll
    for (int q = 0; q < K; q++) { acc[q] = A(0); first[q] = A(0); }
    bool emitted = false, final_emitted = false;
    uint32_t first_row = 0, cur = a < b ? s_row[a] : NONE;
    for (uint32_t e = a; e < b; e++) {
        const uint32_t r = s_row[e];
        if (r != cur) {
            if (!emitted) {
                emitted = true;
                first_row = cur;
#pragma unroll
                for (int q = 0; q < K; q++) first[q] = acc[q];
            } else {
#pragma unroll
                for (int q = 0; q < K; q++) y[(size_t)cur * ldy + q] = to_val<T>(acc[q]);
            }
#pragma unroll
            for (int q = 0; q < K; q++) acc[q] = A(0);
            cur = r;
        }
        if constexpr (PROD) {
            acc[0] = acc[0] + s_prod[e];
        } else {
            const uint32_t c = s_col[e];
            const T v = s_val[e];
            T xs[K];
            ld_n<T, K>(x + (size_t)c * ldx, xs);
#pragma unroll
            for (int q = 0; q < K; q++) acc[q] = acc[q] + mul<T>(v, xs[q]);
        }
    }
    if (a < b) {
        const uint32_t nxt = b < n ? s_row[b] : next_row;
        if (nxt != cur) {  // this thread holds the row's last entry
            final_emitted = true;
            if (!emitted) {
                emitted = true;
                first_row = cur;
#pragma unroll
                for (int q = 0; q < K; q++) first[q] = acc[q];
            } else {
#pragma unroll
                for (int q = 0; q < K; q++) y[(size_t)cur * ldy + q] = to_val<T>(acc[q]);
            }
#pragma unroll
            for (int q = 0; q < K; q++) acc[q] = A(0);
        }
    }
    const uint32_t key = cur;
    block_seg_scan<A, K, NT>(key, acc, s_key, s_scan, s_wc);
    // ATOM (the reference's lock-based merges): the tile's first row, which may
    // have begun in an earlier tile, and the row it carries out are added to y
    // with atomics (both rows are zeroed before the launch); no carries.
    if (emitted) {
        const bool cont = tid > 0 && s_key[tid - 1] == first_row;
        const bool shared = ATOM && first_row == s_row[0];
        for (int q = 0; q < K; q++) {
            const A cin = cont ? s_scan[q][tid - 1] : A(0);
            if (shared) atomic_add_y<T>(y + (size_t)first_row * ldy + q, cin + first[q]);
            else y[(size_t)first_row * ldy + q] = to_val<T>(cin + first[q]);
        }
    }
    const uint32_t last = n ? (n - 1) / IPT : 0;
    if ((uint32_t)tid == last) {
        if constexpr (ATOM) {
            if (n && !final_emitted) {
                for (int q = 0; q < K; q++) atomic_add_y<T>(y + (size_t)key * ldy + q, acc[q]);
            }
        } else {
            carry_row[t] = (n == 0 || final_emitted) ? NONE : key;
#pragma unroll
            for (int q = 0; q < K; q++) carry_val[(size_t)t * K + q] = acc[q];
        }
    }
    __syncthreads();  // LDS is reused by the next tile
    }
}

// ---------------------------------------------------------------------------
// BCSR, register blocked: L lanes per block row; a lane takes whole R x C
// blocks (contiguous values) and keeps R partial sums per right-hand side.
// ---------------------------------------------------------------------------
template <typename T, int R, int C, int L, int K>
__global__ __launch_bounds__(TPB) void bcsr_vector_kernel(const Prob *__restrict__ probs,
                                                          const uint32_t *__restrict__ unit_pid,
                                                          const uint32_t *__restrict__ urow, Ptrs<T> io,
                                                         uint32_t nunits) {
    using A = acc_t<T>;
    constexpr int G = TPB / L;
    for (uint32_t u = blockIdx.x; u < nunits; u += gridDim.x) {
    const uint32_t pid = unit_pid[u];
    const Prob P = probs[pid];
    const gptr<const T> __restrict__ x = gp(io.x(pid, P));
    const gptr<T> __restrict__ y = gp(io.y(pid, P));
    [[maybe_unused]] const gptr<const uint32_t> __restrict__ rp = gp(P.ptr), rip = gp(P.rind), cip = gp(P.cind);
    const gptr<const T> __restrict__ bv = gp((const T *)P.val);
    const int ldx = io.lx(P), ldy = io.ly(P);
    const uint32_t r0 = urow[u + pid], r1 = urow[u + pid + 1];
    const int lane = threadIdx.x % L, grp = threadIdx.x / L;
    const uint32_t ncols = P.ncols_elem;
    for (uint32_t I = r0 + grp; I < r1; I += G) {
        const uint32_t b = rp[I], e = rp[I + 1];
        A acc[R][K];
#pragma unroll
        for (int i = 0; i < R; i++)
#pragma unroll
            for (int q = 0; q < K; q++) acc[i][q] = A(0);
        for (uint32_t k = b + lane; k < e; k += L) {
            const uint32_t J = ldg(cip + k);
            T v[R * C];
            ldg_n<T, R * C>(bv + (size_t)k * (R * C), v);
            if (K == 1 && ldx == 1 && J * C + C <= ncols) {
                // whole block column inside the matrix: the C x values in one load when aligned
                T xs[C];
                ld_n<T, C>(x + (size_t)J * C, xs);
#pragma unroll
                for (int c = 0; c < C; c++)
#pragma unroll
                    for (int i = 0; i < R; i++) acc[i][0] = fma_acc<T>(v[i * C + c], xs[c], acc[i][0]);
                continue;
            }
#pragma unroll
            for (int c = 0; c < C; c++) {
                const uint32_t col = J * C + c;
                if (col < ncols) {
                    T xs[K];  // row col of x, wide loads when aligned
                    ld_n<T, K>(x + (size_t)col * ldx, xs);
#pragma unroll
                    for (int q = 0; q < K; q++) {
#pragma unroll
                        for (int i = 0; i < R; i++) acc[i][q] = fma_acc<T>(v[i * C + c], xs[q], acc[i][q]);
                    }
                }
            }
        }
#pragma unroll
        for (int i = 0; i < R; i++)
#pragma unroll
            for (int q = 0; q < K; q++) acc[i][q] = group_sum<L>(acc[i][q]);
        if (lane == 0) {
#pragma unroll
            for (int i = 0; i < R; i++) {
                const uint32_t row = I * R + i;
                if (row < P.nrows_elem) {
#pragma unroll
                    for (int q = 0; q < K; q++) y[(size_t)row * ldy + q] = to_val<T>(acc[i][q]);
                }
            }
        }
    }
    }
}

// ---------------------------------------------------------------------------
// BCSR, coalesced (one right-hand side): a wave takes one block row at a time
// and its lanes split every block into 16-byte chunks, so one load
// instruction reads 64 consecutive chunks (1 KB of consecutive blocks) where
// the register-blocked kernel touches 64 separate blocks.  A chunk lies in
// one block row (C * sizeof(T) is a multiple of 16); its lane multiplies it
// with the matching 16 bytes of x, and a butterfly over the lanes that share
// a row forms the row sums.
// ---------------------------------------------------------------------------
template <typename T, int R, int C> struct Coal {
    static constexpr int CH = 16 / (int)sizeof(T);  // elements per chunk
    static constexpr int CPR = C / CH;              // chunks per block row
    static constexpr int G = R * CPR;               // chunks per block
    static constexpr int PB = WAVE / G;             // blocks per wave step
    static constexpr bool ok = (C % CH == 0) && G <= WAVE && (WAVE % G == 0);
};

template <typename T, int R, int C>
__global__ __launch_bounds__(TPB) void bcsr_coal_kernel(const Prob *__restrict__ probs,
                                                        const uint32_t *__restrict__ unit_pid,
                                                        const uint32_t *__restrict__ urow, Ptrs<T> io,
                                                        uint32_t nunits) {
    using A = acc_t<T>;
    using CO = Coal<T, R, C>;
    static_assert(CO::ok, "block shape not eligible");
    constexpr int NW = TPB / WAVE;
    const int lane = threadIdx.x % WAVE, wid = threadIdx.x / WAVE;
    const int blk_off = lane / CO::G, g = lane % CO::G, i = g / CO::CPR, c0 = (g % CO::CPR) * CO::CH;
    for (uint32_t u = blockIdx.x; u < nunits; u += gridDim.x) {
    const uint32_t pid = unit_pid[u];
    const Prob P = probs[pid];
    const gptr<const T> __restrict__ x = gp(io.x(pid, P));
    const gptr<T> __restrict__ y = gp(io.y(pid, P));
    const gptr<const uint32_t> __restrict__ rp = gp(P.ptr), cip = gp(P.cind);
    const gptr<const T> __restrict__ bv = gp((const T *)P.val);
    const int ldx = io.lx(P), ldy = io.ly(P);
    const uint32_t r0 = urow[u + pid], r1 = urow[u + pid + 1];
    const uint32_t ncols = P.ncols_elem;
    // x read in whole pieces without bound checks when every block column is whole and x is
    // contiguous and aligned (uniform over the unit; a template argument, so that no load
    // sits under a branch)
    const bool fast = ldx == 1 && ncols % C == 0 &&
                      ((uintptr_t)io.x(pid, P) % (CO::CH * sizeof(T) >= 16 ? 16 : CO::CH * sizeof(T))) == 0;
    auto xpiece = [&](auto fast_c, uint32_t col, T (&xs)[CO::CH]) {
        if constexpr (decltype(fast_c)::value) {
            ld_al<T, CO::CH>(x + col, xs);
        } else if (ldx == 1 && col + CO::CH <= ncols) {
            ld_n<T, CO::CH>(x + col, xs);
        } else {
#pragma unroll
            for (int j = 0; j < CO::CH; j++) xs[j] = col + j < ncols ? x[(size_t)(col + j) * ldx] : T(0);
        }
    };
    auto rows = [&](auto fast_c) {
    for (uint32_t I = r0 + wid; I < r1; I += NW) {
        const uint32_t b = rp[I], e = rp[I + 1];
        A acc = A(0);
        // long rows (power-law hubs): 4 wave steps per iteration, loads first,
        // while the whole window lies in the row (the same test on every lane)
        uint32_t base = b;
        for (; base + 4 * CO::PB <= e; base += 4 * CO::PB) {
            const uint32_t k = base + blk_off;
            uint32_t J[4];
            T v[4][CO::CH], xs[4][CO::CH];
#pragma unroll
            for (int u = 0; u < 4; u++) J[u] = ldg(cip + k + u * CO::PB);
#pragma unroll
            for (int u = 0; u < 4; u++) ldg_n<T, CO::CH>(bv + (size_t)(k + u * CO::PB) * (R * C) + g * CO::CH, v[u]);
#pragma unroll
            for (int u = 0; u < 4; u++) xpiece(fast_c, J[u] * C + c0, xs[u]);
#pragma unroll
            for (int u = 0; u < 4; u++)
#pragma unroll
                for (int j = 0; j < CO::CH; j++) acc = fma_acc<T>(v[u][j], xs[u][j], acc);
        }
        // the rest: U wave steps (about 32 blocks) per iteration with all loads
        // issued first; lanes past the row's end reload its last block and drop the product
        constexpr int U = CO::PB >= 32 ? 1 : 32 / CO::PB > 8 ? 8 : 32 / CO::PB;
        for (uint32_t k = base + blk_off; k < e; k += U * CO::PB) {
            uint32_t J[U];
            T v[U][CO::CH], xs[U][CO::CH];
#pragma unroll
            for (int u = 0; u < U; u++) J[u] = ldg(cip + min(k + u * CO::PB, e - 1));
#pragma unroll
            for (int u = 0; u < U; u++) ldg_n<T, CO::CH>(bv + (size_t)min(k + u * CO::PB, e - 1) * (R * C) + g * CO::CH, v[u]);
#pragma unroll
            for (int u = 0; u < U; u++) xpiece(fast_c, J[u] * C + c0, xs[u]);
#pragma unroll
            for (int u = 0; u < U; u++) {
                const bool in = k + u * CO::PB < e;  // a select, not a branch
#pragma unroll
                for (int j = 0; j < CO::CH; j++) acc = fma_acc<T>(in ? v[u][j] : T(0), xs[u][j], acc);
            }
        }
        // lanes of row i: all chunks of the row (low bits) and all block slots (high bits)
        acc = xor_sum<1, CO::CPR>(acc);
        acc = xor_sum<CO::G, WAVE>(acc);
        if (blk_off == 0 && (g % CO::CPR) == 0) {
            const uint32_t row = I * R + i;
            if (row < P.nrows_elem) y[(size_t)row * ldy] = to_val<T>(acc);
        }
    }
    };
    if (fast) rows(std::true_type{});
    else rows(std::false_type{});
    }
}

// ---------------------------------------------------------------------------
// Several right-hand sides (K = 8, 16, 32), CSR and BCSR with R*C <= 16: a
// wave takes one (block) row at a time, its lanes split into 64/K groups of
// K, lane q of a group holds right-hand side q.  A group takes every
// (64/K)-th entry or block of the row: the K lanes read the same values (one
// transaction) and K consecutive x values of each column (coalesced), so
// each lane keeps only R sums, and one butterfly over the groups ends the
// row.  The matrix is read once for all K columns.
// ---------------------------------------------------------------------------
template <typename T, int R, int C, int K>
__global__ __launch_bounds__(TPB) void rows_rhs_kernel(const Prob *__restrict__ probs,
                                                       const uint32_t *__restrict__ unit_pid,
                                                       const uint32_t *__restrict__ urow, Ptrs<T> io,
                                                       uint32_t nunits) {
    using A = acc_t<T>;
    static_assert(K == 8 || K == 16 || K == 32, "lanes across 8, 16 or 32 right-hand sides");
    constexpr int NW = TPB / WAVE, GB = WAVE / K, U = 2;
    const int lane = threadIdx.x % WAVE, wid = threadIdx.x / WAVE;
    const int q = lane % K, gb = lane / K;
    for (uint32_t u = blockIdx.x; u < nunits; u += gridDim.x) {
    const uint32_t pid = unit_pid[u];
    const Prob P = probs[pid];
    const gptr<const T> __restrict__ x = gp(io.x(pid, P));
    const gptr<T> __restrict__ y = gp(io.y(pid, P));
    const gptr<const uint32_t> __restrict__ rp = gp(P.ptr), cip = gp(P.cind);
    const gptr<const T> __restrict__ bv = gp((const T *)P.val);
    const int ldx = io.lx(P), ldy = io.ly(P);
    const uint32_t r0 = urow[u + pid], r1 = urow[u + pid + 1];
    const uint32_t ncols = P.ncols_elem;
    for (uint32_t I = r0 + wid; I < r1; I += NW) {
        const uint32_t b = rp[I], e = rp[I + 1];
        A acc[R];
#pragma unroll
        for (int i = 0; i < R; i++) acc[i] = A(0);
        // U group steps per iteration, loads first; lanes past the row's end
        // reload its last entry and drop the product
        for (uint32_t k = b + gb; k < e; k += U * GB) {
            uint32_t J[U];
            T v[U][R * C];
#pragma unroll
            for (int w = 0; w < U; w++) {
                const uint32_t kk = min(k + w * GB, e - 1);
                J[w] = ldg(cip + kk);
                ldg_n<T, R * C>(bv + (size_t)kk * (R * C), v[w]);
            }
            T xs[U][C];
#pragma unroll
            for (int w = 0; w < U; w++)
#pragma unroll
                for (int c = 0; c < C; c++) {
                    const uint32_t col = J[w] * C + c;
                    xs[w][c] = col < ncols ? x[(size_t)col * ldx + q] : T(0);
                }
#pragma unroll
            for (int w = 0; w < U; w++)
                if (k + w * GB < e) {
#pragma unroll
                    for (int c = 0; c < C; c++)
#pragma unroll
                        for (int i = 0; i < R; i++) acc[i] = fma_acc<T>(v[w][i * C + c], xs[w][c], acc[i]);
                }
        }
#pragma unroll
        for (int i = 0; i < R; i++) acc[i] = xor_sum<K, WAVE>(acc[i]);
        if (gb == 0) {
#pragma unroll
            for (int i = 0; i < R; i++) {
                const uint32_t row = I * R + i;
                if (row < P.nrows_elem) y[(size_t)row * ldy + q] = to_val<T>(acc[i]);
            }
        }
    }
    }
}

// ---------------------------------------------------------------------------
// BCSR 4 x 4 blocks times 8 or 16 right-hand sides on the matrix cores: the
// block-row product of the reference (spmv/1D/BCSR-nnz/dpu/task.c:337-370)
// as a dense 4 x 4 x K contraction per block.  A wave takes a block row; its
// block columns are read 64 at a time, one per lane, and handed out by
// shuffles.  (32 right-hand sides go to bcsr4_valu_kernel, which is faster on
// the 27-point stencil: FP32 331-345 against 514 us, FP64 633-637 against
// 727 us.)
//
// FP32, v_mfma_f32_4x4x1_16b_f32: 16 independent 4x4x1 products per
// instruction; product m = lane / 4 gets A[i] from lane 4m + i and B[j] from
// lane 4m + j, and lane 4m + j holds D[v][j] in its register v.  Lane
// l = K*g + 4*b + j takes matrix block g of the step (G = 64 / K blocks per
// step), right-hand side column 4b + j, and matrix row j: A = block row j,
// column c; B = x row (4 J + c), column 4b + j.  Four instructions (c = 0..3)
// form the block's 4 x K contribution; the G partial sums are added across
// lanes at the end of the row.  A lane reads its block row (A values) as
// one 16-byte piece and its four B values element by element: the 4K-byte
// rows of x come in whole, coalesced, with K lanes per row.  (Staging the
// block values in LDS and reading x in 16-byte pieces with a DPP transpose
// across each quad measured slower on the 27-point stencil.)
//
// FP64, v_mfma_f64_4x4x4_4b_f64: 4 independent 4x4x4 products per
// instruction; product m takes A[i][k] from lane 16k + 4m + i and B[k][j]
// from lane 16k + 4m + j, and lane 16i + 4m + j holds D[i][j].  Lane
// l = 16h + 4m + o supplies block[o][h] (A, from the block values staged in
// LDS 32 blocks at a time in 16-byte chunks) and x row (4 J + h) (B); the
// columns that B's j = o stands for are a relabelling of the right-hand
// sides: 16: one instruction per block, columns 4m + o; 8: two blocks per
// instruction (products 0-1 and 2-3), columns 4(m%2) + o.
// ---------------------------------------------------------------------------
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef double f64x2 __attribute__((ext_vector_type(2)));
typedef float f32x2 __attribute__((ext_vector_type(2)));

template <typename T, int K>
__global__ __launch_bounds__(TPB) void bcsr_mfma_kernel(const Prob *__restrict__ probs,
                                                        const uint32_t *__restrict__ unit_pid,
                                                        const uint32_t *__restrict__ urow, Ptrs<T> io,
                                                        uint32_t nunits) {
    static_assert(K == 8 || K == 16, "8 or 16 right-hand sides per pass");
    static_assert(sizeof(T) == 4 || sizeof(T) == 8, "FP32 or FP64");
    constexpr int NW = TPB / WAVE;
    constexpr int NB64 = 32;  // FP64: blocks per LDS batch (4 KB of values per wave)
    __shared__ __align__(16) double s_val[sizeof(T) == 8 ? NW : 1][sizeof(T) == 8 ? NB64 * 16 : 1];
    const int lane = threadIdx.x % WAVE, wid = threadIdx.x / WAVE;
    for (uint32_t u = blockIdx.x; u < nunits; u += gridDim.x) {
    const uint32_t pid = unit_pid[u];
    const Prob P = probs[pid];
    const gptr<const T> __restrict__ x = gp(io.x(pid, P));
    const gptr<T> __restrict__ y = gp(io.y(pid, P));
    const gptr<const uint32_t> __restrict__ rp = gp(P.ptr), cip = gp(P.cind);
    const gptr<const T> __restrict__ bv = gp((const T *)P.val);
    const int ldx = io.lx(P), ldy = io.ly(P);
    const uint32_t r0 = urow[u + pid], r1 = urow[u + pid + 1];
    const uint32_t ncols = P.ncols_elem;
    // This wave's block rows r0 + wid + NW t: their block ranges in one load (64 rows
    // a round), and the next row's first block columns loaded while a row runs.
    constexpr uint32_t NBATCH = sizeof(T) == 4 ? WAVE : NB64;  // blocks per batch
    for (uint32_t R0 = r0 + wid; R0 < r1; R0 += NW * WAVE) {
    const uint32_t nrw = min((uint32_t)WAVE, (r1 - R0 + NW - 1) / NW);
    const uint32_t Il = R0 + NW * (uint32_t)lane;
    const uint32_t bl = (uint32_t)lane < nrw ? rp[Il] : 0u, el = (uint32_t)lane < nrw ? rp[Il + 1] : 0u;
    auto first_cols = [&](uint32_t t) {  // row t's first batch of block columns, one per lane
        const uint32_t bb = (uint32_t)__builtin_amdgcn_readlane((int)bl, (int)t);
        const uint32_t ee = (uint32_t)__builtin_amdgcn_readlane((int)el, (int)t);
        return (uint32_t)lane < min(NBATCH, ee - bb) ? ldg(cip + bb + lane) : 0u;
    };
    uint32_t Jn = first_cols(0);
    for (uint32_t t = 0; t < nrw; t++) {
        const uint32_t I = R0 + NW * t;
        const uint32_t b = (uint32_t)__builtin_amdgcn_readlane((int)bl, (int)t);
        const uint32_t e = (uint32_t)__builtin_amdgcn_readlane((int)el, (int)t);
        const uint32_t Jfirst = Jn;
        if (t + 1 < nrw) Jn = first_cols(t + 1);
        if constexpr (sizeof(T) == 4) {
            constexpr int G = WAVE / K;  // matrix blocks per step
            constexpr int U = 8 / G;     // steps whose loads go out together
            const int g = lane / K, q = lane % K, j = lane % 4;
            f32x4 acc = {0.f, 0.f, 0.f, 0.f};
            // blocks in batches of 64: one coalesced load of their block columns,
            // handed to the lanes that need them by a shuffle
            for (uint32_t base = b; base < e; base += WAVE) {
                const uint32_t nb = min((uint32_t)WAVE, e - base);
                const uint32_t Jl = base == b ? Jfirst : ((uint32_t)lane < nb ? ldg(cip + base + lane) : 0u);
                for (uint32_t t0 = 0; t0 < nb; t0 += G * U) {
                    T a[U][4], xb[U][4];
                    bool act[U];
                    uint32_t J[U];
#pragma unroll
                    for (int w = 0; w < U; w++) {
                        const uint32_t sl = t0 + w * G + g;
                        act[w] = sl < nb;
                        const uint32_t sc = act[w] ? sl : 0u;  // a valid block; its product is dropped
                        J[w] = (uint32_t)__shfl((int)Jl, (int)sc, WAVE);
                        ldg_n<T, 4>(bv + (size_t)(base + sc) * 16 + j * 4, a[w]);  // matrix row j of the block
                    }
#pragma unroll
                    for (int w = 0; w < U; w++)
#pragma unroll
                        for (int c = 0; c < 4; c++) {
                            const uint32_t col = J[w] * 4 + c;
                            xb[w][c] = (act[w] && col < ncols) ? x[(size_t)col * ldx + q] : T(0);
                        }
#pragma unroll
                    for (int w = 0; w < U; w++)
#pragma unroll
                        for (int c = 0; c < 4; c++)
                            acc = __builtin_amdgcn_mfma_f32_4x4x1f32(act[w] ? a[w][c] : T(0), xb[w][c], acc, 0, 0, 0);
                }
            }
#pragma unroll
            for (int v = 0; v < 4; v++) acc[v] = xor_sum<K, WAVE>(acc[v]);
            if (g == 0) {
#pragma unroll
                for (int v = 0; v < 4; v++) {
                    const uint32_t row = I * 4 + v;
                    if (row < P.nrows_elem) y[(size_t)row * ldy + q] = acc[v];
                }
            }
        } else {
            constexpr int BPS = K == 8 ? 2 : 1;  // matrix blocks per instruction
            constexpr int U = 4 / BPS;           // steps whose loads go out together
            double *const sv = s_val[wid];
            const int h = lane / 16, m = (lane / 4) % 4, o = lane % 4;
            const int gblk = BPS == 2 ? m / 2 : 0;  // block of this product within the step
            // right-hand side column of this lane (a relabelling, see above)
            const int colq = K == 16 ? 4 * m + o : 4 * (m % 2) + o;
            double acc = 0.0;
            for (uint32_t base = b; base < e; base += NB64) {
                const uint32_t nb = min((uint32_t)NB64, e - base);
                // the batch's values to LDS in 16-byte chunks, its block columns to lanes
                __builtin_amdgcn_wave_barrier();
                for (uint32_t c = lane; c < nb * 8; c += WAVE)
                    *(u32x4 *)(sv + c * 2) = __builtin_nontemporal_load((gptr<const u32x4>)(bv + (size_t)base * 16) + c);
                const uint32_t Jl = base == b ? Jfirst : ((uint32_t)lane < nb ? ldg(cip + base + lane) : 0u);
                __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
                __builtin_amdgcn_wave_barrier();
                for (uint32_t t0 = 0; t0 < nb; t0 += BPS * U) {
                    double xv[U];
                    bool act[U];
                    uint32_t sidx[U];
#pragma unroll
                    for (int w = 0; w < U; w++) {
                        const uint32_t sl = t0 + w * BPS + gblk;
                        act[w] = sl < nb;
                        sidx[w] = act[w] ? sl : 0u;
                        const uint32_t row = (uint32_t)__shfl((int)Jl, (int)sidx[w], WAVE) * 4 + h;
                        const bool ok = act[w] && row < ncols;
                        xv[w] = ok ? (double)x[(size_t)row * ldx + colq] : 0.0;
                    }
#pragma unroll
                    for (int w = 0; w < U; w++) {
                        const double av = act[w] ? sv[sidx[w] * 16 + o * 4 + h] : 0.0;
                        acc = __builtin_amdgcn_mfma_f64_4x4x4f64(av, xv[w], acc, 0, 0, 0);
                    }
                }
                __builtin_amdgcn_wave_barrier();  // the next batch overwrites the staged values
            }
            const uint32_t row = I * 4 + h;
            if constexpr (BPS == 2) acc += lane_xor<8>(acc);  // the other block of the step: m ^ 2
            if ((BPS == 1 || gblk == 0) && row < P.nrows_elem) y[(size_t)row * ldy + colq] = (T)acc;
        }
    }
    }
    }
}

// ---------------------------------------------------------------------------
// BCSR 4 x 4 FP32 or FP64 blocks times 32 right-hand sides on the vector
// ALUs (the block-row product of spmv/1D/BCSR-nnz/dpu/task.c:337-370 over a
// 4 x 32 panel of x per block).  A wave takes a block row; it stages the
// values of its blocks in LDS 4 KB at a time (coalesced 16-byte loads) and
// hands their block columns to the lanes by shuffles.  With V values per
// 16-byte piece (4 FP32, 2 FP64), lane l = (32 / V) g + r takes block g of the
// step (8 FP32 or 4 FP64 blocks per step) and right-hand sides V r .. V r +
// V - 1: it reads the block's 16 values from LDS (the lanes of a block read
// the same words) and the four x rows' 16-byte pieces it needs, and keeps a
// 4 x V tile of sums.  At the end of the row the block slots' tiles are
// added by a scattered butterfly (each exchange hands over half of the sums a
// lane keeps: FP32 14 lane exchanges instead of 48), after which every lane
// stores two values of y.  When every block column is whole the x rows load unchecked
// and an empty slot of the last step skips its products; otherwise empty
// slots read a block of zeros and x rows past the matrix read as zeros.
// The FP32 kernel issues its vector ALU instructions most of the time
// (counters on the 27-point stencil 64^3 x 4: 179M instructions, 87% of the
// SIMD cycles, before the two changes above): they took it from 409-417 us to
// 345 us.  The FP32 matrix-core kernel takes 511-515 us: its 4-byte x loads (one column per
// lane, as the 4 x 4 x 1 instruction wants its B operand) issue four times as
// many load instructions; at 8 columns the matrix-core kernel stays faster
// (207-213 against 250-264 us).
// ---------------------------------------------------------------------------
template <typename T, int K>
__global__ __launch_bounds__(TPB) void bcsr4_valu_kernel(const Prob *__restrict__ probs,
                                                         const uint32_t *__restrict__ unit_pid,
                                                         const uint32_t *__restrict__ urow, Ptrs<T> io,
                                                         uint32_t nunits) {
    static_assert(K == 32, "32 right-hand sides per pass");
    static_assert(__is_same(T, float) || __is_same(T, double), "FP32 or FP64 blocks");
    constexpr int NW = TPB / WAVE;
    constexpr int V = 16 / (int)sizeof(T);  // values per 16-byte piece
    constexpr int GS = K / V;               // lanes per block
    constexpr int G = WAVE / GS;            // blocks per step (8 FP32, 4 FP64)
    constexpr int NB = 4096 / (16 * (int)sizeof(T));  // blocks per staged batch (4 KB)
    constexpr int PPB = (int)sizeof(T);               // 16-byte pieces per block
    static_assert(NB * PPB == 4 * WAVE, "four pieces per lane and batch");
    typedef T VT __attribute__((ext_vector_type(V)));
    // a wave's batch of blocks, then a block of zeros that the empty slots of a step read
    __shared__ __align__(16) T s_val[NW][(NB + 1) * 16];
    const int lane = threadIdx.x % WAVE, wid = threadIdx.x / WAVE;
    const int g = lane / GS, r = lane % GS;
    T *const sv = s_val[wid];
    if (lane < 16) sv[NB * 16 + lane] = T(0);
    for (uint32_t u = blockIdx.x; u < nunits; u += gridDim.x) {
    const uint32_t pid = unit_pid[u];
    const Prob P = probs[pid];
    const gptr<const T> __restrict__ x = gp(io.x(pid, P));
    const gptr<T> __restrict__ y = gp(io.y(pid, P));
    const gptr<const uint32_t> __restrict__ rp = gp(P.ptr), cip = gp(P.cind);
    const gptr<const T> __restrict__ bv = gp((const T *)P.val);
    const int ldx = io.lx(P), ldy = io.ly(P);
    const uint32_t r0 = urow[u + pid], r1 = urow[u + pid + 1];
    const uint32_t ncols = P.ncols_elem;
    // whole 16-byte pieces of x and y rows (uniform over the launch); the x case is a
    // template argument so that no load sits under a branch
    const bool xw = ((uintptr_t)io.x(pid, P) % 16 == 0) && (ldx % V == 0);
    const bool yw = ((uintptr_t)io.y(pid, P) % 16 == 0) && (ldy % V == 0);
    auto rows = [&](auto xw_c, auto wh_c) {
        constexpr bool XW = decltype(xw_c)::value;
        constexpr bool WH = decltype(wh_c)::value;  // whole 16-byte x pieces and whole block columns
        for (uint32_t I = r0 + wid; I < r1; I += NW) {
            const uint32_t b = rp[I], e = rp[I + 1];
            T acc[4][V];
#pragma unroll
            for (int i = 0; i < 4; i++)
#pragma unroll
                for (int q = 0; q < V; q++) acc[i][q] = T(0);
            for (uint32_t base = b; base < e; base += NB) {
                const uint32_t nb = min((uint32_t)NB, e - base);
                // the batch's values to LDS in 16-byte pieces, its block columns to lanes
                __builtin_amdgcn_wave_barrier();
#pragma unroll
                for (int c0 = 0; c0 < 4; c0++) {
                    const uint32_t c = (uint32_t)lane + (uint32_t)c0 * WAVE;  // piece c of the batch
                    if (c < nb * PPB)
                        *(u32x4 *)(sv + c * V) = __builtin_nontemporal_load((gptr<const u32x4>)(bv + (size_t)base * 16) + c);
                }
                const uint32_t Jl = (uint32_t)lane < nb ? ldg(cip + base + lane) : 0u;
                __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
                __builtin_amdgcn_wave_barrier();
                if constexpr (WH) {
                    // whole block columns: the block's four x rows load unchecked and an empty
                    // slot skips its products (FP32: two blocks per lane and step, all eight x
                    // rows in flight, measured slower: 122 VGPRs leave four waves per SIMD
                    // instead of five)
                    for (uint32_t t0 = 0; t0 < nb; t0 += G) {
                        const uint32_t sl = t0 + g;
                        const bool act = sl < nb;
                        const uint32_t J = (uint32_t)__shfl((int)Jl, (int)(act ? sl : 0u), WAVE);
                        const gptr<const T> px = x + (size_t)J * 4 * ldx + V * r;
                        VT xv[4], av[16 / V];
#pragma unroll
                        for (int c = 0; c < 4; c++) xv[c] = *(gptr<const VT>)(px + (size_t)c * ldx);
#pragma unroll
                        for (int p = 0; p < 16 / V; p++) av[p] = *(const VT *)(sv + (act ? sl : 0u) * 16 + p * V);
                        if (act) {
#pragma unroll
                            for (int i = 0; i < 4; i++)
#pragma unroll
                                for (int c = 0; c < 4; c++)
#pragma unroll
                                    for (int q = 0; q < V; q++)
                                        acc[i][q] = fma_acc<T>(av[(i * 4 + c) / V][(i * 4 + c) % V], xv[c][q], acc[i][q]);
                        }
                    }
                } else {
                    for (uint32_t t0 = 0; t0 < nb; t0 += G) {
                        const uint32_t sl = t0 + g;
                        const bool act = sl < nb;
                        const uint32_t J = (uint32_t)__shfl((int)Jl, (int)(act ? sl : 0u), WAVE);
                        // an empty slot multiplies the zero block by zeros
                        const T *const pa = sv + (act ? sl : (uint32_t)NB) * 16;
                        T a[16], xr[4][V];
#pragma unroll
                        for (int p = 0; p < 16 / V; p++) {
                            const VT v = *(const VT *)(pa + p * V);
#pragma unroll
                            for (int q = 0; q < V; q++) a[p * V + q] = v[q];
                        }
#pragma unroll
                        for (int c = 0; c < 4; c++) {
                            const uint32_t col = J * 4 + c;
                            const bool ok = act && col < ncols;
                            const gptr<const T> px = x + (size_t)(ok ? col : 0u) * ldx + V * r;
                            if constexpr (XW) {
                                const VT v = *(gptr<const VT>)px;
#pragma unroll
                                for (int q = 0; q < V; q++) xr[c][q] = ok ? v[q] : T(0);
                            } else {
#pragma unroll
                                for (int q = 0; q < V; q++) xr[c][q] = ok ? px[q] : T(0);
                            }
                        }
#pragma unroll
                        for (int i = 0; i < 4; i++)
#pragma unroll
                            for (int c = 0; c < 4; c++)
#pragma unroll
                                for (int q = 0; q < V; q++) acc[i][q] = fma_acc<T>(a[i * 4 + c], xr[c][q], acc[i][q]);
                    }
                }
                __builtin_amdgcn_wave_barrier();  // the next batch overwrites the staged values
            }
            // sums over the G block slots, scattered: each exchange hands over half of the sums
            // a lane keeps (block rows by lane bits 5 and 4; FP32 then column pairs by lane bit
            // 3), so every lane ends with two sums of block row (lane >> 4) & 3, at columns
            // V r and V r + 1 (FP32: 4 r + 2 and 4 r + 3 when lane bit 3 is set)
            const bool hb = lane & 32, mb = lane & 16, lb = lane & 8;
            T h[2][V], c[V], d[2];
#pragma unroll
            for (int i = 0; i < 2; i++)
#pragma unroll
                for (int q = 0; q < V; q++) {
                    const T keep = hb ? acc[i + 2][q] : acc[i][q], send = hb ? acc[i][q] : acc[i + 2][q];
                    h[i][q] = keep + lane_xor<32>(send);
                }
#pragma unroll
            for (int q = 0; q < V; q++) {
                const T keep = mb ? h[1][q] : h[0][q], send = mb ? h[0][q] : h[1][q];
                c[q] = keep + lane_xor<16>(send);
            }
            uint32_t col0 = V * r;
            if constexpr (V == 4) {
#pragma unroll
                for (int q = 0; q < 2; q++) {
                    const T keep = lb ? c[q + 2] : c[q], send = lb ? c[q] : c[q + 2];
                    d[q] = keep + lane_xor<8>(send);
                }
                col0 += lb ? 2 : 0;
            } else {
                d[0] = c[0];
                d[1] = c[1];
            }
            const uint32_t row = I * 4 + (mb ? 1 : 0) + (hb ? 2 : 0);
            if (row < P.nrows_elem) {
                const gptr<T> py = y + (size_t)row * ldy + col0;
                if (yw) {
                    typedef T T2 __attribute__((ext_vector_type(2)));
                    *(gptr<T2>)py = T2{d[0], d[1]};
                } else {
                    py[0] = d[0];
                    py[1] = d[1];
                }
            }
        }
    };
    if (xw && ncols % 4 == 0) rows(std::true_type{}, std::true_type{});
    else if (xw) rows(std::true_type{}, std::false_type{});
    else rows(std::false_type{}, std::false_type{});
    }
}

// ---------------------------------------------------------------------------
// BCOO, block balanced: phase 1 forms the R row sums of every block of the
// tile with lane groups of C over the block's elements (coalesced value
// loads); phase 2 is the COO segmented walk over blocks with R-wide sums.
// ---------------------------------------------------------------------------
//
// PTR (BCSR, entry balanced): the same kernel over merge-path tiles of a BCSR
// matrix (block-row ends + blocks, coordinates crow = cent and cnnz); a
// block's row comes from a binary search over the tile's block-row ends in
// LDS instead of a stored row index, and a tile's last row is complete when
// its end lies in the tile.  Block rows without blocks are zeroed by the
// fix-up.
template <typename T, int R, int C, int IPT, bool PTR = false, bool ATOM = false>
__global__ __launch_bounds__(TPB) void bcoo_seg_kernel(const Prob *__restrict__ probs,
                                                       const uint32_t *__restrict__ unit_pid,
                                                       const uint32_t *__restrict__ cent,
                                                       const uint32_t *__restrict__ cnnz, Ptrs<T> io,
                                                       uint32_t *__restrict__ carry_row,
                                                       acc_t<T> *__restrict__ carry_val,
                                                       uint32_t nunits) {
    using A = acc_t<T>;
    constexpr int TILE = TPB * IPT;  // blocks (BCOO) or block-row ends + blocks (PTR) per tile
    static_assert(TPB % (R * C) == 0, "block must tile the workgroup");
    __shared__ uint32_t s_brow[TILE];
    __shared__ uint32_t s_bcol[TILE];
    __shared__ uint32_t s_rowend[PTR ? TILE + 1 : 1];
    __shared__ A s_part[TILE * R];
    __shared__ uint32_t s_key[TPB];
    __shared__ A s_scan[R][TPB];
    __shared__ A s_wc[R][TPB / WAVE];

    const int tid = threadIdx.x;
    for (uint32_t t = blockIdx.x; t < nunits; t += gridDim.x) {
    const uint32_t pid = unit_pid[t], ci = t + pid;
    const Prob P = probs[pid];
    const gptr<const T> __restrict__ x = gp(io.x(pid, P));
    const gptr<T> __restrict__ y = gp(io.y(pid, P));
    [[maybe_unused]] const gptr<const uint32_t> __restrict__ rp = gp(P.ptr), rip = gp(P.rind), cip = gp(P.cind);
    const gptr<const T> __restrict__ bv = gp((const T *)P.val);
    const int ldx = io.lx(P), ldy = io.ly(P);
    uint32_t b0, n, next_brow, i0 = 0, na = 0;
    if constexpr (PTR) {
        i0 = cent[ci];
        na = cent[ci + 1] - i0;
        b0 = cnnz[ci];
        n = cnnz[ci + 1] - b0;
        next_brow = NONE;  // unused: completeness comes from the row ends
        const uint32_t nrl = min(na + 1, P.nrows - i0);  // the ends of rows i0 .. i0 + na
        // all loads first (lanes past an end load its last element again), then the
        // LDS stores: a load under a branch would wait for the loads before it
        uint32_t re[IPT + 1], bc[IPT];
#pragma unroll
        for (int q = 0; q <= IPT; q++) re[q] = nrl ? rp[i0 + min((uint32_t)(tid + q * TPB), nrl - 1) + 1] : 0u;
#pragma unroll
        for (int q = 0; q < IPT; q++) bc[q] = n ? ldg(cip + b0 + min((uint32_t)(tid + q * TPB), n - 1)) : 0u;
#pragma unroll
        for (int q = 0; q <= IPT; q++) {
            const uint32_t k = tid + q * TPB;
            if (k < nrl) s_rowend[k] = re[q];
            if (q < IPT && k < n) s_bcol[k] = bc[q];
        }
        __syncthreads();
        // row of block b0 + k: i0 + the number of the tile's row ends at or before it
#pragma unroll
        for (int q = 0; q < IPT; q++) {
            const uint32_t k = tid + q * TPB;
            if (k < n) {
                uint32_t lo = 0, hi = na;
                while (lo < hi) {
                    const uint32_t mid = (lo + hi) >> 1;
                    if (s_rowend[mid] <= b0 + k) lo = mid + 1;
                    else hi = mid;
                }
                s_brow[k] = i0 + lo;
            }
        }
    } else {
        b0 = cent[ci];
        n = cent[ci + 1] - b0;
        next_brow = (b0 + n < P.nent) ? rip[b0 + n] : NONE;
        if (n) {  // loads first, then the LDS stores, as above
            uint32_t br[IPT], bc[IPT];
#pragma unroll
            for (int q = 0; q < IPT; q++) {
                const uint32_t k = min((uint32_t)(tid + q * TPB), n - 1);
                br[q] = ldg(rip + b0 + k);
                bc[q] = ldg(cip + b0 + k);
            }
#pragma unroll
            for (int q = 0; q < IPT; q++) {
                const uint32_t k = tid + q * TPB;
                if (k < n) {
                    s_brow[k] = br[q];
                    s_bcol[k] = bc[q];
                }
            }
        }
    }
    __syncthreads();
    using CO = Coal<T, R, C>;
    // x read in whole pieces without bound checks when every block column is whole and x is
    // contiguous and aligned (uniform over the tile; a template argument, so that no load
    // sits under a branch)
    constexpr int XW = (CO::ok && CO::G > 1) ? CO::CH : C;  // x values per piece
    const bool fast = ldx == 1 && P.ncols_elem % C == 0 &&
                      ((uintptr_t)io.x(pid, P) % (XW * sizeof(T) >= 16 ? 16 : XW * sizeof(T))) == 0;
    auto phase1 = [&](auto fast_c) {
    constexpr bool FAST = decltype(fast_c)::value;
    if constexpr (CO::ok && CO::G > 1) {
        // phase 1, blocks wider than 16 bytes: lanes take 16-byte chunks of
        // consecutive blocks (one load instruction covers 1 KB of values), and
        // the lanes holding one block row's chunks sum them
        if (n) {
            const uint32_t nl = n - 1;
            const int g = tid % CO::G;  // TPB is a multiple of G: a lane keeps its chunk slot
            const int c0 = (g % CO::CPR) * CO::CH;
            // every lane stores its sum, past the tile's end too (those slots of s_part are
            // never read): a store under a branch between two iterations would hold the
            // next iteration's loads back
#pragma unroll
            for (int q = 0; q < IPT * CO::G; q++) {
                const uint32_t blk = (uint32_t)(tid + q * TPB) / CO::G;
                const uint32_t bc = min(blk, nl);  // lanes past the tile reload its last block
                const uint32_t col = s_bcol[bc] * C + c0;
                T v[CO::CH], xs[CO::CH];
                ldg_n<T, CO::CH>(bv + (size_t)(b0 + bc) * (R * C) + g * CO::CH, v);
                if constexpr (FAST) {
                    ld_al<T, CO::CH>(x + col, xs);
                } else if (ldx == 1 && col + CO::CH <= P.ncols_elem) {
                    ld_n<T, CO::CH>(x + col, xs);
                } else {
#pragma unroll
                    for (int j = 0; j < CO::CH; j++)
                        xs[j] = col + j < P.ncols_elem ? x[(size_t)(col + j) * ldx] : T(0);
                }
                A p = A(0);
#pragma unroll
                for (int j = 0; j < CO::CH; j++) p = p + mul<T>(v[j], xs[j]);
                p = xor_sum<1, CO::CPR>(p);
                if (g % CO::CPR == 0) s_part[blk * R + g / CO::CPR] = p;
            }
        }
    } else {
    // phase 1: a thread takes whole blocks (R*C contiguous values in wide loads,
    // the C x values in one load when aligned) and forms their R row sums
    if (n) {
#pragma unroll
    for (int q = 0; q < IPT; q++) {
        const uint32_t blk = tid + q * TPB;
        const uint32_t bc = min(blk, n - 1);  // lanes past the tile reload its last block
        const uint32_t J = s_bcol[bc];
        T v[R * C];
        ldg_n<T, R * C>(bv + (size_t)(b0 + bc) * (R * C), v);
        T xs[C];
        if constexpr (FAST) {
            ld_al<T, C>(x + (size_t)J * C, xs);
        } else if (ldx == 1 && J * C + C <= P.ncols_elem) {
            ld_n<T, C>(x + (size_t)J * C, xs);
        } else {
#pragma unroll
            for (int c = 0; c < C; c++) {
                const uint32_t col = J * C + c;
                xs[c] = col < P.ncols_elem ? x[(size_t)col * ldx] : T(0);
            }
        }
#pragma unroll
        for (int i = 0; i < R; i++) {  // past the tile's end too, as above
            A p = A(0);
#pragma unroll
            for (int c = 0; c < C; c++) p = p + mul<T>(v[i * C + c], xs[c]);
            s_part[blk * R + i] = p;
        }
    }
    }
    }
    };
    if (fast) phase1(std::true_type{});
    else phase1(std::false_type{});
    __syncthreads();

    const uint32_t a = min((uint32_t)tid * IPT, n), bb = min(a + IPT, n);
    A acc[R], first[R];
#pragma unroll
    for (int i = 0; i < R; i++) { acc[i] = A(0); first[i] = A(0); }
    bool emitted = false, final_emitted = false;
    uint32_t first_row = 0, cur = a < bb ? s_brow[a] : NONE;
    auto emit_block_row = [&](uint32_t brow, const A (&v)[R]) {
#pragma unroll
        for (int i = 0; i < R; i++) {
            const uint32_t row = brow * R + i;
            if (row < P.nrows_elem) y[(size_t)row * ldy] = to_val<T>(v[i]);
        }
    };
    for (uint32_t e = a; e < bb; e++) {
        const uint32_t r = s_brow[e];
        if (r != cur) {
            if (!emitted) {
                emitted = true;
                first_row = cur;
#pragma unroll
                for (int i = 0; i < R; i++) first[i] = acc[i];
            } else {
                emit_block_row(cur, acc);
            }
#pragma unroll
            for (int i = 0; i < R; i++) acc[i] = A(0);
            cur = r;
        }
#pragma unroll
        for (int i = 0; i < R; i++) acc[i] = acc[i] + s_part[e * R + i];
    }
    if (a < bb) {
        // the thread's last row ends here if the next block is in another row (BCOO), or
        // (last block of a PTR tile) if the row's end lies in the tile
        const bool ends = bb < n ? s_brow[bb] != cur : PTR ? s_rowend[cur - i0] <= b0 + n : next_brow != cur;
        if (ends) {
            final_emitted = true;
            if (!emitted) {
                emitted = true;
                first_row = cur;
#pragma unroll
                for (int i = 0; i < R; i++) first[i] = acc[i];
            } else {
                emit_block_row(cur, acc);
            }
#pragma unroll
            for (int i = 0; i < R; i++) acc[i] = A(0);
        }
    }
    const uint32_t key = cur;
    block_seg_scan<A, R>(key, acc, s_key, s_scan, s_wc);
    // ATOM (BCOO only, the reference's lock-based merges): the tile's first
    // block row and the block row it carries out are added to y with atomics
    // (their rows are zeroed before the launch); no carries.
    auto add_block_row = [&](uint32_t brow, const A (&v)[R]) {
#pragma unroll
        for (int i = 0; i < R; i++) {
            const uint32_t row = brow * R + i;
            if (row < P.nrows_elem) atomic_add_y<T>(y + (size_t)row * ldy, v[i]);
        }
    };
    if (emitted) {
        const bool cont = tid > 0 && s_key[tid - 1] == first_row;
        A v[R];
#pragma unroll
        for (int i = 0; i < R; i++) v[i] = (cont ? s_scan[i][tid - 1] : A(0)) + first[i];
        if (ATOM && first_row == s_brow[0]) add_block_row(first_row, v);
        else emit_block_row(first_row, v);
    }
    const uint32_t last = n ? (n - 1) / IPT : 0;
    if ((uint32_t)tid == last) {
        if constexpr (ATOM) {
            if (n && !final_emitted) add_block_row(key, acc);
        } else {
            const uint32_t kr = (n == 0 || final_emitted) ? NONE : key;
#pragma unroll
            for (int i = 0; i < R; i++) {
                carry_row[(size_t)i * nunits + t] = kr;
                carry_val[(size_t)i * nunits + t] = acc[i];
            }
        }
    }
    __syncthreads();  // LDS is reused by the next tile
    }
}

// ---------------------------------------------------------------------------
// Fix-up: threads [0, S*nunits) add carried partial sums of row runs that
// cross tiles (slot s = row inside a block, S = 1 for element formats), in
// tile order; threads beyond zero the listed rows that have no entries.
// ---------------------------------------------------------------------------
template <typename T, int K>
__global__ void fixup_kernel(const Prob *__restrict__ probs, const uint32_t *__restrict__ unit_pid,
                             const uint32_t *__restrict__ carry_row, const acc_t<T> *__restrict__ carry_val,
                             uint32_t nunits, uint32_t S, Ptrs<T> io, const uint32_t *__restrict__ zpid,
                             const uint32_t *__restrict__ zrow, uint32_t nzero) {
    using A = acc_t<T>;
    const uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t ncarry = (uint64_t)S * nunits;
    if (g < ncarry) {
        const uint32_t s = (uint32_t)(g / nunits), t = (uint32_t)(g % nunits);
        const uint32_t *cr = carry_row + (size_t)s * nunits;
        const A *cv = carry_val + (size_t)s * nunits * K;
        // Loads that depend only on t go out together (the kernel is a chain of
        // dependent loads, so its time is their latency): this unit's carry,
        // its neighbour's, and the carried value (every unit writes one).
        const uint32_t r = cr[t];
        const uint32_t pid = unit_pid[t];
        const uint32_t rprev = t > 0 ? cr[t - 1] : NONE;
        const uint32_t pprev = t > 0 ? unit_pid[t - 1] : NONE;
        A sum[K];
#pragma unroll
        for (int q = 0; q < K; q++) sum[q] = A(0) + cv[(size_t)t * K + q];
        const Prob P = probs[pid];
        if (r == NONE) return;
        if (pprev == pid && rprev == r) return;  // not the first unit of this run
        for (uint32_t u = t + 1; u < nunits && unit_pid[u] == pid && cr[u] == r; u++) {
#pragma unroll
            for (int q = 0; q < K; q++) sum[q] = sum[q] + cv[(size_t)u * K + q];
        }
        const uint32_t row = r * S + s;
        if (row >= P.nrows_elem) return;
        const gptr<T> yp = gp(io.y(pid, P)) + (size_t)row * io.ly(P);
#pragma unroll
        for (int q = 0; q < K; q++) yp[q] = to_val<T>((A)yp[q] + sum[q]);
        return;
    }
    const uint64_t z = g - ncarry;
    if (z >= nzero) return;
    const uint32_t pid = zpid[z];
    const Prob P = probs[pid];
    const gptr<T> yp = gp(io.y(pid, P)) + (size_t)zrow[z] * io.ly(P);
#pragma unroll
    for (int q = 0; q < K; q++) yp[q] = T(0);
}

// ---------------------------------------------------------------------------
// CSR merge fix-up from a list of row runs.  Which rows cross tiles depends
// only on the tiling (tile t carries the row at its end coordinate), so the
// plan lists the runs once: one thread per run adds the run's carries in
// tile order to the row's y, as fixup_kernel does.  Every load a thread needs
// goes out after its one run record arrives (table launches add ytab[pid]).
// ---------------------------------------------------------------------------
struct Run {
    uint32_t pid, t0, n, row;  // problem, first unit, units, row carried
    int64_t yel;               // P.yoff + row: y element of the row (no pointer tables)
    int32_t ldy;               // P.ldy (pointer tables)
    uint32_t pad;
};

template <typename T, int K>
__global__ void run_fixup_kernel(const Run *__restrict__ runs, uint32_t nruns, const acc_t<T> *__restrict__ carry_val,
                                 Ptrs<T> io) {
    using A = acc_t<T>;
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= nruns) return;
    const Run R = runs[j];
    const gptr<T> yp = io.ytab ? gp((T *)io.ytab[R.pid] + io.q0) + (size_t)R.row * R.ldy : gp(io.yb) + R.yel * io.ldy;
    A sum[K];
#pragma unroll
    for (int q = 0; q < K; q++) sum[q] = A(0) + carry_val[(size_t)R.t0 * K + q];
    for (uint32_t u = 1; u < R.n; u++) {
#pragma unroll
        for (int q = 0; q < K; q++) sum[q] = sum[q] + carry_val[(size_t)(R.t0 + u) * K + q];
    }
#pragma unroll
    for (int q = 0; q < K; q++) yp[q] = to_val<T>((A)yp[q] + sum[q]);
}

}  // namespace spd
