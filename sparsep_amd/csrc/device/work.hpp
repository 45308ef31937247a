// Host-side description of one launch group over a list of sub-problems,
// and the typed runners that launch it.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

#include "kern.hpp"

struct sp_ctx;

enum WorkKind { WK_CSR_MERGE = 0, WK_CSR_VEC = 1, WK_COO_SEG = 2, WK_BCSR_VEC = 3, WK_BCOO_SEG = 4, WK_CSR_HOT = 5,
               WK_BCSR_SEG = 6 };

struct Work {
    int kind = WK_CSR_MERGE;
    int dtype = 0;
    uint32_t nprob = 0, nunits = 0;
    spd::Prob *d_prob = nullptr;
    uint32_t *d_unit_pid = nullptr;
    uint32_t *d_c0 = nullptr, *d_c1 = nullptr;  // coordinates, index unit + pid
    uint32_t *d_c2 = nullptr;                   // hot-x CSR: staged entries per tile, index unit + pid
    uint32_t hcol_max = 0;                      // hot-x CSR: largest staged-column count of the problems
    uint32_t ipt = 8, lanes = 1, R = 1, C = 1;
    uint32_t *d_carry_row = nullptr;
    void *d_carry_val = nullptr;
    uint32_t carry_k = 0;  // right-hand sides the carry buffer holds
    uint32_t nzero = 0;
    uint32_t *d_zpid = nullptr, *d_zrow = nullptr;
    bool zero_all = false;  // zero every output row first (2D partial buffers)
    uint64_t zero_rows = 0; // rows to zero when zero_all (times ldy elements)
    uint32_t grid_cap = 0;  // persistent launch: at most this many workgroups (0 = one per unit)
    bool slice = false;     // 2D CSR/COO tiles: one right-hand side gathers from an LDS copy of the tile's x slice
    spd::Run *d_runs = nullptr;  // CSR merge: row runs that cross tiles (fix-up list), nruns of them
    uint32_t nruns = 0;
    bool atomic = false;         // COO/BCOO: rows crossing tiles merged by atomics (no carries)
    bool chunk = false;          // CSR merge / COO: every problem's arrays aligned to 4 entries and padded
                                 // past their end (library allocations), so tiles load 4-entry chunks
    bool quad = false;           // chunked CSR rows / row-ordered COO chunk tiles: every chunk is four
                                 // consecutive columns from a multiple of 4 inside one row (x read four
                                 // values at a time)
    uint32_t nprez = 0;          // atomic: rows zeroed before the main launch
    uint32_t *d_prez_pid = nullptr, *d_prez_row = nullptr;
};

// Widest right-hand-side chunk a kernel kind takes in one pass: the
// row-granular CSR kernel keeps K accumulators per lane; FP32/FP64 4 x 4
// BCSR blocks run on the matrix cores (8 or 16 per pass) or the vector ALUs
// (32 per pass); the register-blocked
// BCSR kernel R*K (R*K <= 64); for 8-byte values and R*C <= 16, lanes run
// across up to 32 right-hand sides with R sums each; the kernels with carries
// keep K partial sums per carry slot.
inline int max_rhs_chunk(const Work &w) {
    const bool wide8 = (w.dtype == SP_I64 || w.dtype == SP_F64) && w.R * w.C <= 16;
    switch (w.kind) {
    case WK_CSR_VEC: return 32;
    case WK_BCSR_VEC:
        if ((w.dtype == SP_F32 || w.dtype == SP_F64) && w.R == 4 && w.C == 4) return 32;  // matrix-core / vector-ALU kernels
        return wide8 ? 32 : w.R <= 2 ? 32 : w.R <= 4 ? 16 : 8;
    case WK_BCOO_SEG: return 1;
    case WK_BCSR_SEG: return 1;
    case WK_CSR_HOT: return 1;
    default: return 8;
    }
}

// Launches `w` for k right-hand sides (processed in chunks of 32/16/8/4/2/1).
// x/y: per-call bases (element offsets from the descriptors apply), or
// xtab/ytab: per-problem device pointer tables.
int run_work(const Work &w, sp_ctx *ctx, const void *xb, void *yb, const uintptr_t *xtab,
             const uintptr_t *ytab, int k, int ldx, int ldy);

template <typename T>
int run_work_t(const Work &w, sp_ctx *ctx, const void *xb, void *yb, const uintptr_t *xtab,
               const uintptr_t *ytab, int k, int ldx, int ldy);

extern template int run_work_t<int8_t>(const Work &, sp_ctx *, const void *, void *, const uintptr_t *,
                                       const uintptr_t *, int, int, int);
extern template int run_work_t<int16_t>(const Work &, sp_ctx *, const void *, void *, const uintptr_t *,
                                        const uintptr_t *, int, int, int);
extern template int run_work_t<int32_t>(const Work &, sp_ctx *, const void *, void *, const uintptr_t *,
                                        const uintptr_t *, int, int, int);
extern template int run_work_t<int64_t>(const Work &, sp_ctx *, const void *, void *, const uintptr_t *,
                                        const uintptr_t *, int, int, int);
extern template int run_work_t<float>(const Work &, sp_ctx *, const void *, void *, const uintptr_t *,
                                      const uintptr_t *, int, int, int);
extern template int run_work_t<double>(const Work &, sp_ctx *, const void *, void *, const uintptr_t *,
                                       const uintptr_t *, int, int, int);

// Host view of one sub-problem while a Work is being built.
struct HostProb {
    spd::Prob dev{};               // device pointers + sizes
    const uint32_t *h_ptr = nullptr;   // CSR rowptr / BCSR browptr (host)
    const uint32_t *h_rind = nullptr;  // COO rowind / BCOO browind (host)
    std::vector<uint32_t> unit_split;  // coordinates of this problem's units (size units+1)
};
