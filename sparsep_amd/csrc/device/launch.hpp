// Typed runner for a Work: picks the kernel instantiation and launches the
// main kernel, then the fix-up (carries + rows without entries) when the
// kind needs one.  Included only by the per-dtype instantiation files.
#pragma once

#include <algorithm>

#include "plan.hpp"
#include "work.hpp"

namespace spd {

template <typename T, int K>
static int launch_main(const Work &w, hipStream_t s, Ptrs<T> io) {
    using A = acc_t<T>;
    const dim3 grid(w.grid_cap && w.grid_cap < w.nunits ? w.grid_cap : w.nunits), block(TPB);
    A *cv = (A *)w.d_carry_val;
    if constexpr (K > 8) {
        // wide chunks: row-granular kernels only (no carries, no LDS staging)
        if (w.kind == WK_CSR_VEC) {
            if constexpr (K * sizeof(T) >= 64) {  // x rows of 64 bytes or more: lanes across the right-hand sides
                hipLaunchKernelGGL((csr_rhs_kernel<T, K>), grid, block, 0, s, w.d_prob, w.d_unit_pid, w.d_c0, io, w.nunits);
            } else {
#define SPD_VECW(Lv) case Lv: hipLaunchKernelGGL((csr_vector_kernel<T, Lv, K>), grid, block, 0, s, w.d_prob, w.d_unit_pid, w.d_c0, io, w.nunits); break;
                switch (w.lanes) {
                    SPD_VECW(1) SPD_VECW(2) SPD_VECW(4) SPD_VECW(8) SPD_VECW(16) SPD_VECW(32) SPD_VECW(64)
                default: return SP_ERR_ARG;
                }
#undef SPD_VECW
            }
            return hipGetLastError() == hipSuccess ? SP_OK : SP_ERR_HIP;
        }
        if (w.kind == WK_BCSR_VEC) {
            if constexpr (__is_same(T, float) || __is_same(T, double)) {
                if (w.R == 4 && w.C == 4) {  // 4 x 4 blocks: the matrix-core kernel (32 columns: vector ALUs)
                    if constexpr (K == 32)
                        hipLaunchKernelGGL((bcsr4_valu_kernel<T, 32>), grid, block, 0, s, w.d_prob, w.d_unit_pid, w.d_c0, io, w.nunits);
                    else
                        hipLaunchKernelGGL((bcsr_mfma_kernel<T, K>), grid, block, 0, s, w.d_prob, w.d_unit_pid, w.d_c0, io, w.nunits);
                    return hipGetLastError() == hipSuccess ? SP_OK : SP_ERR_HIP;
                }
            }
            if constexpr (sizeof(T) == 8) {
                // 8-byte values: lanes across the right-hand sides (R sums per lane) beat R*K
                // register-blocked sums (stencil FP64 4x4, 32 columns: 1.57 vs 2.20 ms)
#define SPD_RR(Rv, Cv) \
    if (w.R == Rv && w.C == Cv) { hipLaunchKernelGGL((rows_rhs_kernel<T, Rv, Cv, K>), grid, block, 0, s, w.d_prob, w.d_unit_pid, w.d_c0, io, w.nunits); return hipGetLastError() == hipSuccess ? SP_OK : SP_ERR_HIP; }
                SPD_RR(1, 1) SPD_RR(2, 2) SPD_RR(2, 4) SPD_RR(4, 2) SPD_RR(4, 4)
#undef SPD_RR
                return SP_ERR_UNSUPPORTED;
            } else {
                const uint32_t lanes = w.lanes ? w.lanes : 16;  // coalesced plans: 16 lanes for several RHS
#define SPD_BLW(Rv, Cv, Lv) \
    if (w.R == Rv && w.C == Cv && lanes == Lv) { hipLaunchKernelGGL((bcsr_vector_kernel<T, Rv, Cv, Lv, K>), grid, block, 0, s, w.d_prob, w.d_unit_pid, w.d_c0, io, w.nunits); return hipGetLastError() == hipSuccess ? SP_OK : SP_ERR_HIP; }
#define SPD_BRCW(Rv, Cv) SPD_BLW(Rv, Cv, 1) SPD_BLW(Rv, Cv, 4) SPD_BLW(Rv, Cv, 16) SPD_BLW(Rv, Cv, 64)
                if constexpr (K <= 32) { SPD_BRCW(1, 1) SPD_BRCW(2, 2) SPD_BRCW(2, 4) }
                if constexpr (K <= 16) { SPD_BRCW(4, 4) SPD_BRCW(4, 2) }
#undef SPD_BRCW
#undef SPD_BLW
                return SP_ERR_UNSUPPORTED;
            }
        }
        return SP_ERR_UNSUPPORTED;
    } else {
    if (w.kind == WK_CSR_HOT) {
        if constexpr (K == 1) {
            // x at the hot columns for every problem, then one workgroup per CU (grid_cap)
            const uint32_t gx = std::min<uint32_t>(std::max<uint32_t>((w.hcol_max + TPB - 1) / TPB, 1), 256);
            hipLaunchKernelGGL((hot_gather_kernel<T>), dim3(gx, w.nprob), dim3(TPB), 0, s, w.d_prob, io);
            hipLaunchKernelGGL((csr_hot_kernel<T>), grid, dim3(HOT_TPB), 0, s, w.d_prob, w.d_unit_pid, w.d_c0,
                               w.d_c1, w.d_c2, io, w.d_carry_row, cv, w.nunits);
            return hipGetLastError() == hipSuccess ? SP_OK : SP_ERR_HIP;
        } else {
            return SP_ERR_UNSUPPORTED;  // one right-hand side per pass (max_rhs_chunk)
        }
    }
    switch (w.kind) {
    case WK_CSR_MERGE: {
        const uint32_t *upid = w.nprob == 1 ? nullptr : w.d_unit_pid;  // one problem: pid 0
        if constexpr (K == 1) {
            if (w.slice) {
                switch (w.ipt) {
                case 4: hipLaunchKernelGGL((csr_merge_kernel<T, 4, 1, true>), grid, block, 0, s, w.d_prob, upid, w.d_c0, w.d_c1, io, w.d_carry_row, cv, w.nunits); break;
                case 8: hipLaunchKernelGGL((csr_merge_kernel<T, 8, 1, true>), grid, block, 0, s, w.d_prob, upid, w.d_c0, w.d_c1, io, w.d_carry_row, cv, w.nunits); break;
                default: return SP_ERR_ARG;
                }
                break;
            }
            // 1- and 2-byte values, arrays aligned to 4 entries and padded: 4-entry chunk loads
            // (4- and 8-byte values keep lane-strided loads, whose gathers share more lines)
            if constexpr (sizeof(T) <= 2) {
            if (w.chunk) {
                switch (w.ipt) {
                case 4: hipLaunchKernelGGL((csr_merge_kernel<T, 4, 1, false, TPB, true>), grid, block, 0, s, w.d_prob, upid, w.d_c0, w.d_c1, io, w.d_carry_row, cv, w.nunits); break;
                case 8: hipLaunchKernelGGL((csr_merge_kernel<T, 8, 1, false, TPB, true>), grid, block, 0, s, w.d_prob, upid, w.d_c0, w.d_c1, io, w.d_carry_row, cv, w.nunits); break;
                case 16: hipLaunchKernelGGL((csr_merge_kernel<T, 8, 1, false, 2 * TPB, true>), grid, dim3(2 * TPB), 0, s, w.d_prob, upid, w.d_c0, w.d_c1, io, w.d_carry_row, cv, w.nunits); break;
                default: return SP_ERR_ARG;
                }
                break;
            }
            }
        }
        switch (w.ipt) {
        case 4: hipLaunchKernelGGL((csr_merge_kernel<T, 4, K>), grid, block, 0, s, w.d_prob, upid, w.d_c0, w.d_c1, io, w.d_carry_row, cv, w.nunits); break;
        case 8: hipLaunchKernelGGL((csr_merge_kernel<T, 8, K>), grid, block, 0, s, w.d_prob, upid, w.d_c0, w.d_c1, io, w.d_carry_row, cv, w.nunits); break;
        case 16:
            if constexpr (K == 1) {
                // 4096-item tiles as 512 threads x 8 items: the registers (and occupancy) of 8 items per lane
                hipLaunchKernelGGL((csr_merge_kernel<T, 8, 1, false, 2 * TPB>), grid, dim3(2 * TPB), 0, s, w.d_prob, upid, w.d_c0, w.d_c1, io, w.d_carry_row, cv, w.nunits);
            } else {
                hipLaunchKernelGGL((csr_merge_kernel<T, 16, K>), grid, block, 0, s, w.d_prob, upid, w.d_c0, w.d_c1, io, w.d_carry_row, cv, w.nunits);
            }
            break;
        default: return SP_ERR_ARG;
        }
        break;
    }
    case WK_CSR_VEC:
        if constexpr (K * sizeof(T) >= 64) {  // x rows of 64 bytes or more: lanes across the right-hand sides
            hipLaunchKernelGGL((csr_rhs_kernel<T, 8>), grid, block, 0, s, w.d_prob, w.d_unit_pid, w.d_c0, io, w.nunits);
            break;
        }
        if constexpr (K == 1) {
            if (w.ipt == 4 && w.chunk) {  // 4-entry chunks per lane (the plan's arrays: aligned and padded)
#define SPD_CHQ(Lv) \
    case Lv: \
        if (w.quad) hipLaunchKernelGGL((csr_chunk_kernel<T, Lv, true>), grid, block, 0, s, w.d_prob, w.d_unit_pid, w.d_c0, io, w.nunits); \
        else hipLaunchKernelGGL((csr_chunk_kernel<T, Lv>), grid, block, 0, s, w.d_prob, w.d_unit_pid, w.d_c0, io, w.nunits); \
        break;
                switch (w.lanes) {
                    SPD_CHQ(4) SPD_CHQ(8) SPD_CHQ(16) SPD_CHQ(32) SPD_CHQ(64)
                default: return SP_ERR_ARG;
                }
#undef SPD_CHQ
                break;
            }
        }
#define SPD_VEC(Lv) case Lv: hipLaunchKernelGGL((csr_vector_kernel<T, Lv, K>), grid, block, 0, s, w.d_prob, w.d_unit_pid, w.d_c0, io, w.nunits); break;
        switch (w.lanes) {
            SPD_VEC(1) SPD_VEC(2) SPD_VEC(4) SPD_VEC(8) SPD_VEC(16) SPD_VEC(32) SPD_VEC(64)
        default: return SP_ERR_ARG;
        }
#undef SPD_VEC
        break;
    case WK_COO_SEG:
        if (w.atomic) {  // 1D plans only: no slices
            switch (w.ipt) {
            case 4: hipLaunchKernelGGL((coo_seg_kernel<T, 4, K, false, TPB, true>), grid, block, 0, s, w.d_prob, w.d_unit_pid, w.d_c0, io, w.d_carry_row, cv, w.nunits); break;
            case 8: hipLaunchKernelGGL((coo_seg_kernel<T, 8, K, false, TPB, true>), grid, block, 0, s, w.d_prob, w.d_unit_pid, w.d_c0, io, w.d_carry_row, cv, w.nunits); break;
            case 16:
                if constexpr (K == 1)
                    hipLaunchKernelGGL((coo_seg_kernel<T, 8, 1, false, 2 * TPB, true>), grid, dim3(2 * TPB), 0, s, w.d_prob, w.d_unit_pid, w.d_c0, io, w.d_carry_row, cv, w.nunits);
                else
                    hipLaunchKernelGGL((coo_seg_kernel<T, 16, K, false, TPB, true>), grid, block, 0, s, w.d_prob, w.d_unit_pid, w.d_c0, io, w.d_carry_row, cv, w.nunits);
                break;
            default: return SP_ERR_ARG;
            }
            break;
        }
        if constexpr (K == 1) {
            if (w.slice) {
                switch (w.ipt) {
                case 4: hipLaunchKernelGGL((coo_seg_kernel<T, 4, 1, true>), grid, block, 0, s, w.d_prob, w.d_unit_pid, w.d_c0, io, w.d_carry_row, cv, w.nunits); break;
                case 8: hipLaunchKernelGGL((coo_seg_kernel<T, 8, 1, true>), grid, block, 0, s, w.d_prob, w.d_unit_pid, w.d_c0, io, w.d_carry_row, cv, w.nunits); break;
                default: return SP_ERR_ARG;
                }
                break;
            }
            if (w.quad) {  // row-ordered chunk tiles reading x four values per load
                switch (w.ipt) {
                case 4: hipLaunchKernelGGL((coo_seg_kernel<T, 4, 1, false, TPB, false, true, true>), grid, block, 0, s, w.d_prob, w.d_unit_pid, w.d_c0, io, w.d_carry_row, cv, w.nunits); break;
                case 8: hipLaunchKernelGGL((coo_seg_kernel<T, 8, 1, false, TPB, false, true, true>), grid, block, 0, s, w.d_prob, w.d_unit_pid, w.d_c0, io, w.d_carry_row, cv, w.nunits); break;
                case 16: hipLaunchKernelGGL((coo_seg_kernel<T, 8, 1, false, 2 * TPB, false, true, true>), grid, dim3(2 * TPB), 0, s, w.d_prob, w.d_unit_pid, w.d_c0, io, w.d_carry_row, cv, w.nunits); break;
                default: return SP_ERR_ARG;
                }
                break;
            }
            if constexpr (sizeof(T) <= 2) {  // as for CSR above
            if (w.chunk) {
                switch (w.ipt) {
                case 4: hipLaunchKernelGGL((coo_seg_kernel<T, 4, 1, false, TPB, false, true>), grid, block, 0, s, w.d_prob, w.d_unit_pid, w.d_c0, io, w.d_carry_row, cv, w.nunits); break;
                case 8: hipLaunchKernelGGL((coo_seg_kernel<T, 8, 1, false, TPB, false, true>), grid, block, 0, s, w.d_prob, w.d_unit_pid, w.d_c0, io, w.d_carry_row, cv, w.nunits); break;
                case 16: hipLaunchKernelGGL((coo_seg_kernel<T, 8, 1, false, 2 * TPB, false, true>), grid, dim3(2 * TPB), 0, s, w.d_prob, w.d_unit_pid, w.d_c0, io, w.d_carry_row, cv, w.nunits); break;
                default: return SP_ERR_ARG;
                }
                break;
            }
            }
        }
        switch (w.ipt) {
        case 4: hipLaunchKernelGGL((coo_seg_kernel<T, 4, K>), grid, block, 0, s, w.d_prob, w.d_unit_pid, w.d_c0, io, w.d_carry_row, cv, w.nunits); break;
        case 8: hipLaunchKernelGGL((coo_seg_kernel<T, 8, K>), grid, block, 0, s, w.d_prob, w.d_unit_pid, w.d_c0, io, w.d_carry_row, cv, w.nunits); break;
        case 16:
            if constexpr (K == 1) {
                // 4096-entry tiles as 512 threads x 8 items (the registers of 8 items per lane)
                hipLaunchKernelGGL((coo_seg_kernel<T, 8, 1, false, 2 * TPB>), grid, dim3(2 * TPB), 0, s, w.d_prob, w.d_unit_pid, w.d_c0, io, w.d_carry_row, cv, w.nunits);
            } else {
                hipLaunchKernelGGL((coo_seg_kernel<T, 16, K>), grid, block, 0, s, w.d_prob, w.d_unit_pid, w.d_c0, io, w.d_carry_row, cv, w.nunits);
            }
            break;
        default: return SP_ERR_ARG;
        }
        break;
    case WK_BCSR_VEC: {
        if constexpr (K == 8 && (__is_same(T, float) || __is_same(T, double))) {
            if (w.R == 4 && w.C == 4) {  // 4 x 4 blocks: the matrix-core kernel
                hipLaunchKernelGGL((bcsr_mfma_kernel<T, 8>), grid, block, 0, s, w.d_prob, w.d_unit_pid, w.d_c0, io, w.nunits);
                break;
            }
        }
        if (w.lanes == 0 && K == 1) {
            // coalesced kernel (the plan checked the shape against the value size)
#define SPD_CO(Rv, Cv) \
    if constexpr (Coal<T, Rv, Cv>::ok) \
        if (w.R == Rv && w.C == Cv) { hipLaunchKernelGGL((bcsr_coal_kernel<T, Rv, Cv>), grid, block, 0, s, w.d_prob, w.d_unit_pid, w.d_c0, io, w.nunits); break; }
            do {
                SPD_CO(2, 2) SPD_CO(4, 4) SPD_CO(8, 8) SPD_CO(2, 4) SPD_CO(4, 2)
                return SP_ERR_UNSUPPORTED;
            } while (0);
#undef SPD_CO
            break;
        }
        const uint32_t lanes = w.lanes ? w.lanes : 16;
#define SPD_BL(Rv, Cv, Lv) \
    if (w.R == Rv && w.C == Cv && lanes == Lv) { hipLaunchKernelGGL((bcsr_vector_kernel<T, Rv, Cv, Lv, K>), grid, block, 0, s, w.d_prob, w.d_unit_pid, w.d_c0, io, w.nunits); break; }
#define SPD_BRC(Rv, Cv) SPD_BL(Rv, Cv, 1) SPD_BL(Rv, Cv, 4) SPD_BL(Rv, Cv, 16) SPD_BL(Rv, Cv, 64)
        do {
            SPD_BRC(1, 1) SPD_BRC(2, 2) SPD_BRC(4, 4) SPD_BRC(8, 8) SPD_BRC(2, 4) SPD_BRC(4, 2)
            return SP_ERR_UNSUPPORTED;
        } while (0);
#undef SPD_BRC
#undef SPD_BL
        break;
    }
    case WK_BCSR_SEG: {
        if (K != 1) return SP_ERR_UNSUPPORTED;
#define SPD_BS(Rv, Cv) \
    if (w.R == Rv && w.C == Cv) { \
        if (w.ipt == 2) hipLaunchKernelGGL((bcoo_seg_kernel<T, Rv, Cv, 2, true>), grid, block, 0, s, w.d_prob, w.d_unit_pid, w.d_c0, w.d_c1, io, w.d_carry_row, cv, w.nunits); \
        else hipLaunchKernelGGL((bcoo_seg_kernel<T, Rv, Cv, 4, true>), grid, block, 0, s, w.d_prob, w.d_unit_pid, w.d_c0, w.d_c1, io, w.d_carry_row, cv, w.nunits); \
        break; }
        do {
            SPD_BS(1, 1) SPD_BS(2, 2) SPD_BS(4, 4) SPD_BS(8, 8) SPD_BS(2, 4) SPD_BS(4, 2)
            return SP_ERR_UNSUPPORTED;
        } while (0);
#undef SPD_BS
        break;
    }
    case WK_BCOO_SEG: {
        if (K != 1) return SP_ERR_UNSUPPORTED;
#define SPD_BO(Rv, Cv) \
    if (w.R == Rv && w.C == Cv) { \
        if (w.atomic) { \
            if (w.ipt == 2) hipLaunchKernelGGL((bcoo_seg_kernel<T, Rv, Cv, 2, false, true>), grid, block, 0, s, w.d_prob, w.d_unit_pid, w.d_c0, nullptr, io, w.d_carry_row, cv, w.nunits); \
            else hipLaunchKernelGGL((bcoo_seg_kernel<T, Rv, Cv, 4, false, true>), grid, block, 0, s, w.d_prob, w.d_unit_pid, w.d_c0, nullptr, io, w.d_carry_row, cv, w.nunits); \
            break; } \
        if (w.ipt == 2) hipLaunchKernelGGL((bcoo_seg_kernel<T, Rv, Cv, 2>), grid, block, 0, s, w.d_prob, w.d_unit_pid, w.d_c0, nullptr, io, w.d_carry_row, cv, w.nunits); \
        else hipLaunchKernelGGL((bcoo_seg_kernel<T, Rv, Cv, 4>), grid, block, 0, s, w.d_prob, w.d_unit_pid, w.d_c0, nullptr, io, w.d_carry_row, cv, w.nunits); \
        break; }
        do {
            SPD_BO(1, 1) SPD_BO(2, 2) SPD_BO(4, 4) SPD_BO(8, 8) SPD_BO(2, 4) SPD_BO(4, 2)
            return SP_ERR_UNSUPPORTED;
        } while (0);
#undef SPD_BO
        break;
    }
    default: return SP_ERR_ARG;
    }
    return hipGetLastError() == hipSuccess ? SP_OK : SP_ERR_HIP;
    }
}

template <typename T, int K>
static int launch_chunk(const Work &w, sp_ctx *ctx, Ptrs<T> io) {
    hipStream_t s = ctx->stream;
    if (w.atomic) {
        // rows shared between tiles (and rows without entries) start at zero; the main
        // kernel adds the shared rows' partial sums atomically, so nothing follows it
        using A = acc_t<T>;
        if (w.nprez) {
            hipLaunchKernelGGL((fixup_kernel<T, K>), dim3((w.nprez + 255) / 256), dim3(256), 0, s, w.d_prob,
                               w.d_unit_pid, w.d_carry_row, (const A *)w.d_carry_val, 0u, 1u, io, w.d_prez_pid,
                               w.d_prez_row, w.nprez);
            if (hipGetLastError() != hipSuccess) return SP_ERR_HIP;
        }
        if (w.nunits) {
            const int rc = launch_main<T, K>(w, s, io);
            if (rc) return rc;
        }
        mark_dominant(ctx);
        return SP_OK;
    }
    if (w.nunits == 0 && w.nzero == 0) return SP_OK;
    if (w.nunits) {
        int rc = launch_main<T, K>(w, s, io);
        if (rc) return rc;
    }
    mark_dominant(ctx);
    if (w.d_runs) {  // CSR merge with its run list
        if (w.nruns) {
            using A = acc_t<T>;
            hipLaunchKernelGGL((run_fixup_kernel<T, K>), dim3((w.nruns + 255) / 256), dim3(256), 0, s, w.d_runs, w.nruns,
                               (const A *)w.d_carry_val, io);
            if (hipGetLastError() != hipSuccess) return SP_ERR_HIP;
        }
        return SP_OK;
    }
    const bool carries = w.kind == WK_CSR_MERGE || w.kind == WK_COO_SEG || w.kind == WK_BCOO_SEG || w.kind == WK_CSR_HOT ||
                         w.kind == WK_BCSR_SEG;
    const uint32_t S = (w.kind == WK_BCOO_SEG || w.kind == WK_BCSR_SEG) ? w.R : 1;
    const uint64_t nthreads = (carries ? (uint64_t)S * w.nunits : 0) + w.nzero;
    if (nthreads) {
        using A = acc_t<T>;
        const uint32_t nun = carries ? w.nunits : 0;
        hipLaunchKernelGGL((fixup_kernel<T, K>), dim3((unsigned)((nthreads + 255) / 256)), dim3(256), 0, s,
                           w.d_prob, w.d_unit_pid, w.d_carry_row, (const A *)w.d_carry_val, nun, S, io,
                           w.d_zpid, w.d_zrow, w.nzero);
        if (hipGetLastError() != hipSuccess) return SP_ERR_HIP;
    }
    return SP_OK;
}

}  // namespace spd

// One launch group whose problems take x / y from device pointer tables and
// their leading dimensions from the descriptors; kc right-hand sides.
template <typename T>
int run_tables_t(const Work &w, sp_ctx *ctx, const uintptr_t *xtab, const uintptr_t *ytab, int kc) {
    using namespace spd;
    Ptrs<T> io;
    io.xb = nullptr;
    io.yb = nullptr;
    io.xtab = xtab;
    io.ytab = ytab;
    io.ldx = io.ldy = 0;
    io.q0 = 0;
    switch (kc) {
    case 32: return launch_chunk<T, 32>(w, ctx, io);
    case 16: return launch_chunk<T, 16>(w, ctx, io);
    case 8: return launch_chunk<T, 8>(w, ctx, io);
    case 4: return launch_chunk<T, 4>(w, ctx, io);
    case 2: return launch_chunk<T, 2>(w, ctx, io);
    case 1: return launch_chunk<T, 1>(w, ctx, io);
    default: return SP_ERR_ARG;
    }
}

template <typename T>
int run_work_t(const Work &w, sp_ctx *ctx, const void *xb, void *yb, const uintptr_t *xtab,
               const uintptr_t *ytab, int k, int ldx, int ldy) {
    using namespace spd;
    int done = 0;
    const int kmax = max_rhs_chunk(w);
    if (w.zero_all && w.zero_rows && yb) {
        // Partial buffers of tiled plans: rows a tile has no entries for stay 0.
        int rc = launch_fill_zero(ctx->stream, yb, (uint64_t)w.zero_rows * ldy * sizeof(T));
        if (rc) return rc;
    }
    while (done < k) {
        const int left = k - done;
        int kc = left >= 32 ? 32 : left >= 16 ? 16 : left >= 8 ? 8 : left >= 4 ? 4 : left >= 2 ? 2 : 1;
        while (kc > kmax) kc >>= 1;
        if ((w.kind == WK_CSR_MERGE || w.kind == WK_COO_SEG) && (uint32_t)kc > w.carry_k) kc = (int)w.carry_k;
        Ptrs<T> io;
        io.xb = xb ? (const T *)xb + done : nullptr;
        io.yb = yb ? (T *)yb + done : nullptr;
        io.xtab = xtab;
        io.ytab = ytab;
        io.ldx = ldx;
        io.ldy = ldy;
        io.q0 = done;
        int rc = kc == 32 ? launch_chunk<T, 32>(w, ctx, io)
               : kc == 16 ? launch_chunk<T, 16>(w, ctx, io)
               : kc == 8 ? launch_chunk<T, 8>(w, ctx, io)
               : kc == 4 ? launch_chunk<T, 4>(w, ctx, io)
               : kc == 2 ? launch_chunk<T, 2>(w, ctx, io)
                         : launch_chunk<T, 1>(w, ctx, io);
        if (rc) return rc;
        done += kc;
    }
    return SP_OK;
}
