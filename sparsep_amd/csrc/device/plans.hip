// Plan builders: turn a host matrix plus the reference partitioners' splits
// into the device Work of one variant.  SparseP's two levels map onto the
// GPU as: part (ex DPU, `nparts`) -> a contiguous range of work units;
// sub-part (ex tasklet, `nsub`) -> one work unit (a workgroup, or a run of
// fixed-size tiles for the entry-balanced kernels).  The split arrays come
// from the same partitioners the host surface exports.
#include <string.h>

#include <algorithm>
#include <thread>
#include <vector>

#include "plan.hpp"

using spd::merge_search;
using spd::Prob;
using spd::TPB;

namespace {

size_t acc_size(int dt) { return (dt == SP_I64 || dt == SP_F64) ? 8 : 4; }

struct Builder {
    sp_plan *p;
    std::vector<Prob> probs;
    std::vector<uint32_t> unit_pid, c0, c1, c2;
    std::vector<uint32_t> zpid, zrow;

    explicit Builder(sp_plan *pl) : p(pl) {}

    void add(const Prob &P, const std::vector<uint32_t> &k0, const std::vector<uint32_t> *k1,
             const std::vector<uint32_t> *k2 = nullptr) {
        if (k2) {
            if (k2->empty()) c2.push_back(0);
            else c2.insert(c2.end(), k2->begin(), k2->end());
        }
        const uint32_t pid = (uint32_t)probs.size();
        probs.push_back(P);
        const size_t nu = k0.empty() ? 0 : k0.size() - 1;
        for (size_t u = 0; u < nu; u++) unit_pid.push_back(pid);
        if (k0.empty()) c0.push_back(0);
        else c0.insert(c0.end(), k0.begin(), k0.end());
        if (k1) {
            if (k1->empty()) c1.push_back(0);
            else c1.insert(c1.end(), k1->begin(), k1->end());
        }
    }

    int finish(int kind, bool carries, uint32_t slots) {
        Work &w = p->work;
        w.kind = kind;
        w.dtype = p->dtype;
        w.nprob = (uint32_t)probs.size();
        w.nunits = (uint32_t)unit_pid.size();
        int rc;
        if ((rc = plan_upload(p, probs.data(), probs.size() * sizeof(Prob), (void **)&w.d_prob))) return rc;
        if ((rc = plan_upload(p, unit_pid.data(), unit_pid.size() * 4, (void **)&w.d_unit_pid))) return rc;
        if ((rc = plan_upload(p, c0.data(), c0.size() * 4, (void **)&w.d_c0))) return rc;
        if (!c1.empty() && (rc = plan_upload(p, c1.data(), c1.size() * 4, (void **)&w.d_c1))) return rc;
        if (!c2.empty() && (rc = plan_upload(p, c2.data(), c2.size() * 4, (void **)&w.d_c2))) return rc;
        w.nzero = (uint32_t)zrow.size();
        if (w.nzero) {
            if ((rc = plan_upload(p, zpid.data(), zpid.size() * 4, (void **)&w.d_zpid))) return rc;
            if ((rc = plan_upload(p, zrow.data(), zrow.size() * 4, (void **)&w.d_zrow))) return rc;
        }
        if (carries) {
            w.carry_k = (kind == WK_BCOO_SEG || kind == WK_BCSR_SEG) ? 1 : 8;
            // atomic merges write no carries (batches allocate their own)
            if (p->algo != SP_ALGO_SEGMENTED_ATOMIC) {
                const size_t n = (size_t)std::max<uint32_t>(w.nunits, 1) * slots;
                if ((rc = plan_alloc(p, n * 4, (void **)&w.d_carry_row))) return rc;
                if ((rc = plan_alloc(p, n * w.carry_k * acc_size(p->dtype), &w.d_carry_val))) return rc;
            }
        }
        if (kind == WK_CSR_MERGE && carries && !w.nzero) {
            // tile t carries row c0[t + pid + 1] when that row exists; consecutive
            // tiles carrying the same row form one run
            std::vector<spd::Run> runs;
            for (uint32_t t = 0; t < w.nunits; t++) {
                const uint32_t pid = unit_pid[t], r = c0[(size_t)t + pid + 1];
                const Prob &P = probs[pid];
                if (r >= P.nrows || r >= P.nrows_elem) continue;
                if (!runs.empty()) {
                    spd::Run &b = runs.back();
                    if (b.pid == pid && b.row == r && b.t0 + b.n == t) { b.n++; continue; }
                }
                runs.push_back(spd::Run{pid, t, 1, r, P.yoff + (int64_t)r, P.ldy, 0});
            }
            w.nruns = (uint32_t)runs.size();
            if ((rc = plan_upload(p, runs.data(), runs.size() * sizeof(spd::Run), (void **)&w.d_runs))) return rc;
            p->launches = 1 + (w.nruns ? 1 : 0);
        } else {
            p->launches = 1 + ((carries || w.nzero) ? 1 : 0);
        }
        p->h_probs = probs;
        p->h_unit_pid = unit_pid;
        p->h_c0 = c0;
        p->h_c1 = c1;
        p->h_c2 = c2;
        p->h_zpid = zpid;
        p->h_zrow = zrow;
        return SP_OK;
    }
};

// Contiguous, strictly increasing boundaries from a list of cut points.
void push_cut(std::vector<uint32_t> &v, uint32_t c) {
    if (v.empty() || c > v.back()) v.push_back(c);
}

// Merge-path tile coordinates of a CSR problem (tile = TPB*ipt items).
void merge_coords(const uint32_t *rp, uint32_t nrows, uint32_t nnz, uint32_t tile, std::vector<uint32_t> &crow,
                  std::vector<uint32_t> &cnnz) {
    const uint64_t total = (uint64_t)nrows + nnz;
    const uint32_t nt = (uint32_t)((total + tile - 1) / tile);
    crow.resize(nt + 1);
    cnnz.resize(nt + 1);
    const uint32_t base = rp[0];
    for (uint32_t t = 0; t <= nt; t++) {
        const uint64_t d = std::min<uint64_t>((uint64_t)t * tile, total);
        const uint32_t i = merge_search((uint32_t)d, nrows, nnz, [&](uint32_t r) { return rp[r + 1] - base; }, 0);
        crow[t] = i;
        cnnz[t] = (uint32_t)(d - i) + base;
    }
}

// Row units of one row range [r0, r1) of a CSR (or BCSR block-row) pointer
// array: the tasklet-level split of the reference (by rows or by entries,
// with its row granularity), one unit per tasklet.  A unit longer than
// max_rows rows is cut into even pieces: the nnz-balanced split puts every
// trailing empty row into its last tasklet (a 2D stripe of a banded matrix
// can leave half the rows empty), and one workgroup walking those rows
// would serialise the launch.
int row_units(const uint32_t *rp, uint32_t r0, uint32_t r1, uint32_t ntasklets, int by_nnz, uint32_t gran,
              uint32_t max_rows, std::vector<uint32_t> &out) {
    const uint32_t rows = r1 - r0;
    if (rows == 0) return SP_OK;
    std::vector<uint32_t> split(ntasklets + 1);
    int rc = by_nnz ? sp_partition_tsklt_by_nnz(rp, r0, rows, rp[r1] - rp[r0], ntasklets, gran, split.data())
                    : sp_partition_tsklt_by_row(rows, ntasklets, gran, split.data());
    if (rc) return rc;
    split[ntasklets] = rows;  // the reference can leave trailing rows uncovered; cover them
    uint32_t prev = 0;
    for (uint32_t i = 0; i <= ntasklets; i++) {
        const uint32_t cut = std::max(prev, std::min(split[i], rows));
        if (max_rows && cut - prev > max_rows)
            for (uint32_t s = prev + max_rows; s < cut; s += max_rows) push_cut(out, r0 + s);
        push_cut(out, r0 + cut);
        prev = cut;
    }
    return SP_OK;
}

uint32_t auto_lanes(double avg) {
    uint32_t L = 1;
    while (L < 64 && L * 2 <= avg * 1.5) L *= 2;
    return L;
}

// Lanes per row of the 1D row-granular CSR kernel: the smallest power of two
// covering the average row (the exhaustive autotuner's timings over 500
// synthetic matrices favour it over auto_lanes by ~4% geometric mean; the
// extra lanes buy latency hiding on matrices of a few million entries).
uint32_t auto_lanes_1d(double avg) {
    uint32_t L = 1;
    while (L < 64 && L < avg) L *= 2;
    return L;
}

// Lanes per row of the chunked row kernel (4 entries per lane and load): the
// smallest power of two covering the average row's chunks, 4 to 64.
uint32_t chunk_lanes(double avg) {
    uint32_t L = 4;
    while (L < 64 && 4.0 * L < avg) L *= 2;
    return L;
}

// Rows made of aligned runs of four columns (csr_chunk_kernel's quads): every
// row's length a multiple of 4 and every group of four entries (which then
// never crosses a row) the columns c, c+1, c+2, c+3 with c a multiple of 4,
// in a matrix whose column count is a multiple of 4.
static bool quad_rows(const uint32_t *rowptr, uint32_t nrows, const uint32_t *c, uint32_t nnz, uint32_t ncols) {
    if (nnz == 0 || ncols % 4) return false;
    for (uint32_t i = 0; i < nrows; i++)
        if ((rowptr[i + 1] - rowptr[i]) % 4) return false;
    for (uint32_t j = 0; j < nnz; j += 4)
        if (c[j] % 4 || c[j + 1] != c[j] + 1 || c[j + 2] != c[j] + 2 || c[j + 3] != c[j] + 3) return false;
    return true;
}

// Shapes the coalesced BCSR kernel takes: a block row holds whole 16-byte
// chunks and a block at most 64 of them (see spd::Coal).  It gives a whole
// wave to each block row, so it is chosen only when the average block row
// fills at least half a wave with chunks.
bool coal_ok(int dtype, uint32_t R, uint32_t C, double avg_blocks) {
    const uint32_t ch = 16 / (uint32_t)sp_dtype_size(dtype);
    if (C % ch) return false;
    const uint32_t g = R * (C / ch);
    const bool shape = (R == 2 && C == 2) || (R == 4 && C == 4) || (R == 8 && C == 8) || (R == 2 && C == 4) ||
                       (R == 4 && C == 2);
    return shape && g <= 64 && 64 % g == 0 && avg_blocks * g >= 32.0;
}

uint32_t pick_ipt(const sp_opts &o, uint32_t def) {
    const uint32_t I = o.items_per_lane ? o.items_per_lane : def;
    return (I == 4 || I == 8 || I == 16) ? I : 0;
}

// Tile an entry range [e0, e1) into fixed tiles of at most `tile` entries.
void tile_range(uint32_t e0, uint32_t e1, uint32_t tile, std::vector<uint32_t> &out) {
    push_cut(out, e0);
    for (uint64_t e = (uint64_t)e0 + tile; e < e1; e += tile) push_cut(out, (uint32_t)e);
    push_cut(out, e1);
}

// Column-ordered merge tiles: entry k of tile [j0, j1) becomes the key
// col << pb | (k - j0); each tile's keys are sorted (column order, ties in row
// order) and its values permuted to match.  Tiles are independent, so they
// are spread over host threads.
void column_ordered_tiles(const uint32_t *colind, const void *values, uint32_t nnz,
                          const std::vector<uint32_t> &cnnz, uint32_t pb, size_t sv, std::vector<uint32_t> &key,
                          std::vector<char> &val) {
    key.resize(nnz);
    val.resize((size_t)nnz * sv);
    const size_t nt = cnnz.empty() ? 0 : cnnz.size() - 1;
    const char *src = (const char *)values;
    auto work = [&](size_t t0, size_t step) {
        for (size_t t = t0; t < nt; t += step) {
            const uint32_t j0 = cnnz[t], j1 = cnnz[t + 1];
            for (uint32_t k = j0; k < j1; k++) key[k] = (colind[k] << pb) | (k - j0);
            std::sort(key.begin() + j0, key.begin() + j1);
            const uint32_t mask = (1u << pb) - 1;
            for (uint32_t k = j0; k < j1; k++)
                memcpy(&val[(size_t)k * sv], src + (size_t)(j0 + (key[k] & mask)) * sv, sv);
        }
    };
    const size_t nth = std::min<size_t>({16, std::max(1u, std::thread::hardware_concurrency()), nt / 64 + 1});
    std::vector<std::thread> th;
    for (size_t i = 1; i < nth; i++) th.emplace_back(work, i, nth);
    work(0, nth);
    for (auto &t : th) t.join();
}

// Atomic merges (SP_ALGO_SEGMENTED_ATOMIC): the rows a tile may share with
// its neighbours, those of its first and last entry (R element rows each for
// blocks), are zeroed before the launch, together with the listed rows that
// have no entries; the kernel then adds those rows' partial sums atomically.
int atomic_prezero(sp_plan *p, const uint32_t *rind, const std::vector<uint32_t> &units, uint32_t R,
                   uint32_t nrows_elem, const std::vector<uint32_t> &zrow) {
    std::vector<uint32_t> rows(zrow);
    for (size_t u = 0; u + 1 < units.size(); u++) {
        if (units[u + 1] == units[u]) continue;
        for (const uint32_t r : {rind[units[u]], rind[units[u + 1] - 1]})
            for (uint32_t i = 0; i < R; i++)
                if ((uint64_t)r * R + i < nrows_elem) rows.push_back(r * R + i);
    }
    std::sort(rows.begin(), rows.end());
    rows.erase(std::unique(rows.begin(), rows.end()), rows.end());
    Work &w = p->work;
    w.atomic = true;
    w.nprez = (uint32_t)rows.size();
    const std::vector<uint32_t> pid(rows.size(), 0);
    int rc;
    if ((rc = plan_upload(p, pid.data(), pid.size() * 4, (void **)&w.d_prez_pid))) return rc;
    if ((rc = plan_upload(p, rows.data(), rows.size() * 4, (void **)&w.d_prez_row))) return rc;
    p->launches = 1 + (w.nprez ? 1 : 0);
    return SP_OK;
}

// Blocks per lane of the BCOO kernel: its LDS holds R sums per block of the
// tile, so tall blocks with 8-byte sums take half the tile (FP64 4x4: 50 KB ->
// 26 KB of LDS, 3 -> 6 workgroups per CU).
uint32_t bcoo_ipt(int dt, uint32_t R) { return acc_size(dt) * R >= 32 ? 2 : 4; }

// Staged-x slots of the hot-x kernel for a value type.
uint32_t hot_slots(int dt) {
    switch (dt) {
    case SP_I8: return spd::Hot<int8_t>::SLOTS;
    case SP_I16: return spd::Hot<int16_t>::SLOTS;
    case SP_I32: return spd::Hot<int32_t>::SLOTS;
    case SP_I64: return spd::Hot<int64_t>::SLOTS;
    case SP_F32: return spd::Hot<float>::SLOTS;
    default: return spd::Hot<double>::SLOTS;
    }
}

// Hot-x tiles: the staged columns are the up-to-`slots` most frequent ones
// used at least twice (ties to the lower column), numbered in column order.
// Each merge tile stores its staged entries first, ordered by slot, then the
// others by column; the key is id << pb | position in the tile, with id the
// slot or the column.  nhot[t] counts tile t's staged entries.
void hot_tiles(const sp_csr *A, const std::vector<uint32_t> &cnnz, uint32_t pb, size_t sv, uint32_t slots,
               std::vector<uint32_t> &hcol, std::vector<uint32_t> &nhot, std::vector<uint32_t> &key,
               std::vector<char> &val) {
    std::vector<uint32_t> cnt(A->ncols, 0);
    for (uint32_t k = 0; k < A->nnz; k++) cnt[A->colind[k]]++;
    std::vector<uint32_t> cols;
    for (uint32_t c = 0; c < A->ncols; c++)
        if (cnt[c] >= 2) cols.push_back(c);
    if (cols.size() > slots) {
        std::nth_element(cols.begin(), cols.begin() + slots, cols.end(), [&](uint32_t a, uint32_t b) {
            return cnt[a] != cnt[b] ? cnt[a] > cnt[b] : a < b;
        });
        cols.resize(slots);
    }
    std::sort(cols.begin(), cols.end());
    hcol = cols;
    std::vector<uint32_t> slot(A->ncols, spd::NONE);
    for (uint32_t s = 0; s < (uint32_t)hcol.size(); s++) slot[hcol[s]] = s;
    key.resize(A->nnz);
    val.resize((size_t)A->nnz * sv);
    const size_t nt = cnnz.empty() ? 0 : cnnz.size() - 1;
    nhot.assign(nt + 1, 0);
    const char *src = (const char *)A->values;
    const uint32_t mask = (1u << pb) - 1;
    auto work = [&](size_t t0, size_t step) {
        std::vector<uint32_t> hot, cold;
        for (size_t t = t0; t < nt; t += step) {
            const uint32_t j0 = cnnz[t], j1 = cnnz[t + 1];
            hot.clear();
            cold.clear();
            for (uint32_t k = j0; k < j1; k++) {
                const uint32_t c = A->colind[k], sl = slot[c];
                if (sl != spd::NONE) hot.push_back((sl << pb) | (k - j0));
                else cold.push_back((c << pb) | (k - j0));
            }
            std::sort(hot.begin(), hot.end());
            std::sort(cold.begin(), cold.end());
            nhot[t] = (uint32_t)hot.size();
            std::copy(hot.begin(), hot.end(), key.begin() + j0);
            std::copy(cold.begin(), cold.end(), key.begin() + j0 + hot.size());
            for (uint32_t k = j0; k < j1; k++)
                memcpy(&val[(size_t)k * sv], src + (size_t)(j0 + (key[k] & mask)) * sv, sv);
        }
    };
    const size_t nth = std::min<size_t>({16, std::max(1u, std::thread::hardware_concurrency()), nt / 64 + 1});
    std::vector<std::thread> th;
    for (size_t i = 1; i < nth; i++) th.emplace_back(work, i, nth);
    work(0, nth);
    for (auto &t : th) t.join();
}

bool coo_sorted(const sp_coo *A) {
    for (uint32_t k = 1; k < A->nnz; k++) {
        if (A->rowind[k] < A->rowind[k - 1]) return false;
        if (A->rowind[k] == A->rowind[k - 1] && A->colind[k] < A->colind[k - 1]) return false;
    }
    return true;
}

}  // namespace

// ---------------------------------------------------------------------------
// CSR (ids 0, 1)
// ---------------------------------------------------------------------------
int build_csr_plan(sp_plan *p, const sp_csr *A) {
    const size_t sv = sp_dtype_size(A->dtype);
    if (!A->rowptr || (A->nnz && (!A->colind || !A->values))) return SP_ERR_ARG;
    if (A->rowptr[0] != 0 || A->rowptr[A->nrows] != A->nnz) return SP_ERR_ARG;
    for (uint32_t i = 0; i < A->nrows; i++)
        if (A->rowptr[i + 1] < A->rowptr[i]) return SP_ERR_ARG;
    for (uint32_t k = 0; k < A->nnz; k++)
        if (A->colind[k] >= A->ncols) return SP_ERR_ARG;
    if (p->algo == SP_ALGO_AUTO) {
        uint32_t maxlen = 0;
        for (uint32_t i = 0; i < A->nrows; i++) maxlen = std::max(maxlen, A->rowptr[i + 1] - A->rowptr[i]);
        const double avg = A->nrows ? (double)A->nnz / A->nrows : 0.0;
        // Row-granular kernel for even row lengths up to a few million entries (one launch,
        // no carries); the merge-path kernel for skewed rows, and for large matrices, where
        // its deeper per-thread load batches win (27-point stencil, 113M entries: 1.3x).
        p->algo = (maxlen <= 4 * avg + 64 && A->nnz <= (16u << 20)) ? SP_ALGO_VECTOR : SP_ALGO_SEGMENTED;
        // Even rows of 16 entries or more: the row kernel with 4-entry chunk loads, at any size
        // (27-point stencil, 113M entries: 245 us against 303 us for the merge path)
        if (maxlen <= 4 * avg + 64 && avg >= 16.0 && p->opts.items_per_lane == 0) {
            p->algo = SP_ALGO_VECTOR;
            p->opts.items_per_lane = 4;
        }
        // Skewed rows and more than 8M entries: 4096-item tiles (16 items per lane, run as
        // 512 x 8), whose column order shares the most lines (R-MAT 2^20: 77 vs 96 us); below
        // that there are too few tiles per CU, and regular rows do better with 2048.
        // (only where they can be column ordered, ncols <= 2^20: row-ordered 4096-item tiles
        // are slower than column-ordered 2048-item ones)
        if (p->opts.items_per_lane == 0 && maxlen > 4 * avg + 64 && A->nnz > (8u << 20) && A->ncols <= (1u << 20))
            p->opts.items_per_lane = 16;
        // Column-ordered tiles give the same sums in the same order and gather fewer cache
        // lines (R-MAT 2^20: 8% faster; 27-point stencil: 6%); they need col << log2(tile)
        // to fit 32 bits.
        if (p->algo == SP_ALGO_SEGMENTED) {
            const uint32_t I = pick_ipt(p->opts, 8);
            uint32_t pb = 0;
            while (I && (1u << pb) < TPB * I) pb++;
            if (I && A->ncols <= (1ull << (32 - pb))) p->algo = SP_ALGO_SEGMENTED_SORTED;
        }
    }
    Prob P{};
    int rc;
    void *d;
    if ((rc = plan_upload(p, A->rowptr, ((size_t)A->nrows + 1) * 4, &d))) return rc;
    P.ptr = (const uint32_t *)d;
    std::vector<uint32_t> crow, cnnz;
    if (p->algo == SP_ALGO_SEGMENTED_HOT) {
        const uint32_t tile = spd::HOT_TPB * spd::HOT_IPT;
        uint32_t pb = 0;
        while ((1u << pb) < tile) pb++;
        if (A->ncols > (1ull << (32 - pb))) return SP_ERR_UNSUPPORTED;  // col << pb must fit 32 bits
        if (A->nrows) merge_coords(A->rowptr, A->nrows, A->nnz, tile, crow, cnnz);
        std::vector<uint32_t> hcol, nhot, key;
        std::vector<char> val;
        hot_tiles(A, cnnz, pb, sv, hot_slots(A->dtype), hcol, nhot, key, val);
        if ((rc = plan_upload(p, key.data(), (size_t)A->nnz * 4, &d))) return rc;
        P.cind = (const uint32_t *)d;
        if ((rc = plan_upload(p, val.data(), (size_t)A->nnz * sv, &d))) return rc;
        P.val = d;
        if ((rc = plan_upload(p, hcol.data(), hcol.size() * 4, &d))) return rc;
        P.hcol = (const uint32_t *)d;
        if ((rc = plan_alloc(p, (hcol.size() * sv + 15) / 16 * 16, &d))) return rc;
        P.xh = d;
        P.nhcol = (uint32_t)hcol.size();
        P.pbits = pb;
        P.nrows = P.nrows_elem = A->nrows;
        P.nent = A->nnz;
        P.ncols_elem = A->ncols;
        p->ipt = p->work.ipt = spd::HOT_IPT;
        p->nparts = p->ntiles = crow.empty() ? 0 : (uint32_t)crow.size() - 1;
        p->nsub = spd::HOT_TPB;
        p->work.hcol_max = P.nhcol;
        p->work.grid_cap = (uint32_t)std::max(1, p->ctx->num_cus);  // one workgroup per CU
        Builder b(p);
        b.add(P, crow, &cnnz, &nhot);
        return b.finish(WK_CSR_HOT, true, 1);
    }
    if (p->algo == SP_ALGO_SEGMENTED_SORTED) {
        const uint32_t I = pick_ipt(p->opts, 8);
        if (!I) return SP_ERR_ARG;
        uint32_t pb = 0;
        while ((1u << pb) < TPB * I) pb++;
        if (A->ncols > (1ull << (32 - pb))) return SP_ERR_UNSUPPORTED;  // col << pb must fit 32 bits
        if (A->nrows) merge_coords(A->rowptr, A->nrows, A->nnz, TPB * I, crow, cnnz);
        std::vector<uint32_t> key;
        std::vector<char> val;
        column_ordered_tiles(A->colind, A->values, A->nnz, cnnz, pb, sv, key, val);
        if ((rc = plan_upload(p, key.data(), (size_t)A->nnz * 4, &d))) return rc;
        P.cind = (const uint32_t *)d;
        if ((rc = plan_upload(p, val.data(), (size_t)A->nnz * sv, &d))) return rc;
        P.val = d;
        P.pbits = pb;
    } else {
        if ((rc = plan_upload(p, A->colind, (size_t)A->nnz * 4, &d))) return rc;
        P.cind = (const uint32_t *)d;
        if ((rc = plan_upload(p, A->values, (size_t)A->nnz * sv, &d))) return rc;
        P.val = d;
    }
    P.nrows = P.nrows_elem = A->nrows;
    P.nent = A->nnz;
    P.ncols_elem = A->ncols;
    Builder b(p);
    if (p->algo == SP_ALGO_VECTOR) {
        // items_per_lane 4: each lane loads whole 4-entry chunks (csr_chunk_kernel)
        const bool ch = p->opts.items_per_lane == 4;
        const double avg = A->nrows ? (double)A->nnz / A->nrows : 1.0;
        const uint32_t L = p->opts.lanes_per_row ? p->opts.lanes_per_row : ch ? chunk_lanes(avg) : auto_lanes_1d(avg);
        if (L > 64 || (L & (L - 1)) || (ch && L < 4)) return SP_ERR_ARG;
        p->lanes = p->work.lanes = L;
        if (ch) {
            p->ipt = p->work.ipt = 4;
            p->work.chunk = true;  // the plan's own arrays: aligned and padded (plan_alloc)
            p->work.quad = quad_rows(A->rowptr, A->nrows, A->colind, A->nnz, A->ncols);
        }
        const uint32_t P1 = p->opts.nparts ? p->opts.nparts : 1;
        std::vector<uint32_t> split(P1 + 1);
        rc = p->id == SP_K_CSR_NNZ ? sp_partition_by_nnz(A->rowptr, A->nrows, P1, split.data())
                                   : sp_partition_by_row(A->nrows, P1, split.data());
        if (rc) return rc;
        split[P1] = A->nrows;
        const uint32_t gran = (uint32_t)(8 / sv);
        const uint64_t rows_per_unit = 4ULL * (TPB / L);
        std::vector<uint32_t> units;
        push_cut(units, 0);
        for (uint32_t q = 0; q < P1; q++) {
            const uint32_t r0 = split[q], r1 = std::max(split[q + 1], r0);
            const uint32_t T = p->opts.nsub ? p->opts.nsub
                                            : (uint32_t)std::max<uint64_t>(1, (r1 - r0 + rows_per_unit - 1) / rows_per_unit);
            if ((rc = row_units(A->rowptr, r0, r1, T, p->opts.sub_balance == SP_BAL_NNZ, gran,
                                (uint32_t)(4 * rows_per_unit), units)))
                return rc;
        }
        push_cut(units, A->nrows);
        if (A->nrows == 0) units.clear();
        p->nparts = P1;
        p->nsub = units.empty() ? 0 : (uint32_t)units.size() - 1;
        b.add(P, units, nullptr);
        return b.finish(WK_CSR_VEC, false, 1);
    }
    const uint32_t I = pick_ipt(p->opts, 8);
    if (!I) return SP_ERR_ARG;
    p->ipt = p->work.ipt = I;
    if (p->algo != SP_ALGO_SEGMENTED_SORTED && A->nrows) merge_coords(A->rowptr, A->nrows, A->nnz, TPB * I, crow, cnnz);
    p->nparts = p->ntiles = crow.empty() ? 0 : (uint32_t)crow.size() - 1;
    p->nsub = TPB;
    b.add(P, crow, &cnnz);
    p->work.chunk = true;  // the plan's own arrays: aligned and padded (plan_alloc)
    return b.finish(WK_CSR_MERGE, true, 1);
}

// ---------------------------------------------------------------------------
// COO (ids 2, 3, 4)
// ---------------------------------------------------------------------------
int build_coo_plan(sp_plan *p, const sp_coo *A0) {
    const sp_coo *A = A0;
    sp_coo *sorted = nullptr;
    for (uint32_t k = 0; k < A0->nnz; k++)
        if (A0->rowind[k] >= A0->nrows || A0->colind[k] >= A0->ncols) return SP_ERR_ARG;
    if (!coo_sorted(A0)) {  // every COO app.c sorts first (spmv/1D/COO-row/host/app.c calls sortCOOMatrix)
        int rc = sp_coo_alloc(A0->nrows, A0->ncols, A0->nnz, A0->dtype, &sorted);
        if (rc) return rc;
        memcpy(sorted->rowind, A0->rowind, (size_t)A0->nnz * 4);
        memcpy(sorted->colind, A0->colind, (size_t)A0->nnz * 4);
        memcpy(sorted->values, A0->values, (size_t)A0->nnz * sp_dtype_size(A0->dtype));
        if ((rc = sp_sort_coo(sorted))) { sp_free_coo(sorted); return rc; }
        A = sorted;
    }
    const size_t sv = sp_dtype_size(A->dtype);
    const uint32_t n = A->nrows, nnz = A->nnz;
    std::vector<uint32_t> rowptr(n + 1, 0);
    for (uint32_t k = 0; k < nnz; k++) rowptr[A->rowind[k] + 1]++;
    for (uint32_t i = 0; i < n; i++) rowptr[i + 1] += rowptr[i];
    Prob P{};
    int rc = SP_OK;
    void *d;
    Builder b(p);
    const bool atom = p->algo == SP_ALGO_SEGMENTED_ATOMIC;
    do {
        P.nrows = P.nrows_elem = n;
        P.nent = nnz;
        P.ncols_elem = A->ncols;
        // 4 items per lane, the tuner's usual winner for 1D COO; skewed rows and more than 8M
        // entries take 4096-entry tiles (16, run as 512 x 8: R-MAT 2^20 86 vs 109 us), as for CSR
        uint32_t maxlen = 0;
        for (uint32_t i = 0; i < n; i++) maxlen = std::max(maxlen, rowptr[i + 1] - rowptr[i]);
        if (p->opts.items_per_lane == 0 && maxlen > 4.0 * nnz / std::max(n, 1u) + 64 && nnz > (8u << 20) &&
            A->ncols <= (1u << 20) && p->algo != SP_ALGO_SEGMENTED)  // column-ordered 4096-entry tiles only
            p->opts.items_per_lane = 16;
        // rows made of aligned runs of four columns: row-ordered chunk tiles reading x four
        // values per load (unless column order is asked for, or atomics), 8 items per lane
        // (27-point stencil 64^3 x 4: FP32 295-297 us against 363-366 us at 4 and 360 us for
        // column-ordered tiles; FP64 401-408 against 446-456 and 437-441 us)
        const bool quads = !atom && !p->opts.lds_slice && p->algo != SP_ALGO_SEGMENTED_SORTED &&
                           quad_rows(rowptr.data(), n, A->colind, nnz, A->ncols);
        const uint32_t I = pick_ipt(p->opts, quads ? 8 : 4);
        if (!I) { rc = SP_ERR_ARG; break; }
        p->ipt = p->work.ipt = I;
        const uint32_t tile = TPB * I;
        // Column-ordered tiles (as for CSR) unless the plain segmented order is asked for or the
        // columns do not fit next to the position bits.
        uint32_t pb = 0;
        while ((1u << pb) < tile) pb++;
        bool col_order = p->algo != SP_ALGO_SEGMENTED && A->ncols <= (1ull << (32 - pb));
        if (quads) col_order = false;
        const uint32_t P1 = p->opts.nparts ? p->opts.nparts : 1;
        // part entry ranges
        std::vector<uint32_t> pe0(P1), pe1(P1);
        if (p->id == SP_K_COO_NNZ) {
            std::vector<uint32_t> prow(P1), rpp(P1), merge(P1);
            std::vector<uint32_t> rows_nnz(n);
            for (uint32_t i = 0; i < n; i++) rows_nnz[i] = rowptr[i + 1] - rowptr[i];
            if (nnz == 0) {
                for (uint32_t q = 0; q < P1; q++) pe0[q] = pe1[q] = 0;
            } else {
                std::vector<uint32_t> cnt(P1);
                if ((rc = sp_partition_coo_nnz(rows_nnz.data(), n, nnz, P1, prow.data(), rpp.data(), pe0.data(),
                                               cnt.data(), merge.data())))
                    break;
                for (uint32_t q = 0; q < P1; q++) pe1[q] = pe0[q] + cnt[q];
            }
        } else {
            std::vector<uint32_t> split(P1 + 1);
            rc = p->id == SP_K_COO_NNZ_RGRN ? sp_partition_by_nnz(rowptr.data(), n, P1, split.data())
                                            : sp_partition_by_row(n, P1, split.data());
            if (rc) break;
            split[P1] = n;
            for (uint32_t q = 0; q < P1; q++) {
                pe0[q] = rowptr[split[q]];
                pe1[q] = rowptr[std::max(split[q + 1], split[q])];
            }
        }
        // sub-parts: the reference tasklet split of each part, then fixed tiles
        std::vector<uint32_t> units;
        push_cut(units, 0);
        for (uint32_t q = 0; q < P1; q++) {
            const uint32_t T = p->opts.nsub ? p->opts.nsub : 1;
            const uint32_t cnt = pe1[q] - pe0[q];
            if (cnt == 0) continue;
            std::vector<uint32_t> st(T), ct(T);
            if ((rc = sp_partition_even(cnt, T, st.data(), ct.data()))) break;
            for (uint32_t s = 0; s < T; s++)
                if (ct[s]) tile_range(pe0[q] + st[s], pe0[q] + st[s] + ct[s], tile, units);
        }
        if (rc) break;
        push_cut(units, nnz);
        if (nnz == 0) units.clear();
        for (size_t u = 0; u + 1 < units.size(); u++)
            if (units[u + 1] - units[u] > tile) col_order = false;  // positions must fit pb bits
        p->algo = atom ? SP_ALGO_SEGMENTED_ATOMIC : col_order ? SP_ALGO_SEGMENTED_SORTED : SP_ALGO_SEGMENTED;
        if ((rc = plan_upload(p, A->rowind, (size_t)nnz * 4, &d))) break;
        P.rind = (const uint32_t *)d;
        if (col_order) {
            std::vector<uint32_t> key;
            std::vector<char> val;
            column_ordered_tiles(A->colind, A->values, nnz, units, pb, sv, key, val);
            if ((rc = plan_upload(p, key.data(), (size_t)nnz * 4, &d))) break;
            P.cind = (const uint32_t *)d;
            if ((rc = plan_upload(p, val.data(), (size_t)nnz * sv, &d))) break;
            P.val = d;
            P.pbits = pb;
        } else {
            if ((rc = plan_upload(p, A->colind, (size_t)nnz * 4, &d))) break;
            P.cind = (const uint32_t *)d;
            if ((rc = plan_upload(p, A->values, (size_t)nnz * sv, &d))) break;
            P.val = d;
        }
        p->nparts = P1;
        p->ntiles = units.empty() ? 0 : (uint32_t)units.size() - 1;
        for (uint32_t i = 0; i < n; i++)
            if (rowptr[i + 1] == rowptr[i]) { b.zpid.push_back(0); b.zrow.push_back(i); }
        b.add(P, units, nullptr);
        p->work.chunk = true;  // the plan's own arrays: aligned and padded (plan_alloc)
        p->work.quad = quads;
        rc = b.finish(WK_COO_SEG, true, 1);
        if (!rc && atom) rc = atomic_prezero(p, A->rowind, units, 1, n, b.zrow);
    } while (0);
    if (sorted) sp_free_coo(sorted);
    return rc;
}

// ---------------------------------------------------------------------------
// BCSR (ids 5, 6)
// ---------------------------------------------------------------------------
int build_bcsr_plan(sp_plan *p, const sp_bcsr *B) {
    const size_t sv = sp_dtype_size(B->dtype);
    const uint32_t R = B->row_block_size, C = B->col_block_size, nbr = B->num_block_rows;
    const size_t rc2 = (size_t)R * C;
    Prob P{};
    int rc;
    void *d;
    if ((rc = plan_upload(p, B->browptr, ((size_t)nbr + 1) * 4, &d))) return rc;
    P.ptr = (const uint32_t *)d;
    if ((rc = plan_upload(p, B->bcolind, B->num_blocks * 4, &d))) return rc;
    P.cind = (const uint32_t *)d;
    if ((rc = plan_upload(p, B->bval, B->num_blocks * rc2 * sv, &d))) return rc;
    P.val = d;
    P.nrows = nbr;
    P.nent = (uint32_t)B->num_blocks;
    P.nrows_elem = B->nrows;
    P.ncols_elem = B->ncols;
    // Entry balanced (merge-path tiles over block-row ends and blocks) when asked, or,
    // automatically, when block rows are skewed (as for CSR): a block row per wave or
    // lane group would leave the hub rows to a few waves.
    uint32_t maxlen = 0;
    for (uint32_t I = 0; I < nbr; I++) maxlen = std::max(maxlen, B->browptr[I + 1] - B->browptr[I]);
    const double avg_b = nbr ? (double)B->num_blocks / nbr : 0.0;
    const bool seg = B->num_blocks <= 0xFFFFFFFFULL &&
                     (p->algo == SP_ALGO_SEGMENTED ||
                      (p->algo == SP_ALGO_AUTO && p->opts.lanes_per_row == 0 && maxlen > 4 * avg_b + 64));
    if (seg) {
        p->algo = SP_ALGO_SEGMENTED;
        const uint32_t I = bcoo_ipt(B->dtype, R);
        p->ipt = p->work.ipt = I;
        p->work.R = R;
        p->work.C = C;
        std::vector<uint32_t> crow, cnnz;
        if (nbr) merge_coords(B->browptr, nbr, (uint32_t)B->num_blocks, TPB * I, crow, cnnz);
        p->nparts = p->ntiles = crow.empty() ? 0 : (uint32_t)crow.size() - 1;
        p->nsub = TPB;
        Builder b(p);
        // element rows of block rows without blocks are zeroed by the fix-up
        for (uint32_t I2 = 0; I2 < nbr; I2++)
            if (B->browptr[I2 + 1] == B->browptr[I2])
                for (uint32_t i = 0; i < R && (uint64_t)I2 * R + i < B->nrows; i++) {
                    b.zpid.push_back(0);
                    b.zrow.push_back(I2 * R + i);
                }
        b.add(P, crow, &cnnz);
        return b.finish(WK_BCSR_SEG, true, R);
    }
    p->algo = SP_ALGO_VECTOR;
    uint32_t L = p->opts.lanes_per_row;
    const double avg = nbr ? (double)B->num_blocks / nbr : 1.0;
    const bool coal = L == 0 && coal_ok(B->dtype, R, C, avg);
    if (coal) {
        L = 64;  // a wave per block row; work.lanes 0 selects the coalesced kernel
    } else if (L == 0) {
        L = avg >= 48 ? 64 : avg >= 12 ? 16 : avg >= 3 ? 4 : 1;
    }
    if (L != 1 && L != 4 && L != 16 && L != 64) return SP_ERR_ARG;
    p->lanes = p->work.lanes = coal ? 0 : L;
    p->work.R = R;
    p->work.C = C;
    const uint32_t P1 = p->opts.nparts ? p->opts.nparts : 1;
    std::vector<uint32_t> bs(P1 + 1), bl(P1), ns(P1);
    if ((rc = sp_partition_bcsr(B->browptr, B->nnz_per_block, nbr, B->num_blocks, B->nnz, P1,
                                p->id == SP_K_BCSR_NNZ, bs.data(), bl.data(), ns.data())))
        return rc;
    bs[P1] = nbr;
    const uint64_t brows_per_unit = 4ULL * (TPB / L);
    std::vector<uint32_t> units;
    push_cut(units, 0);
    for (uint32_t q = 0; q < P1; q++) {
        const uint32_t r0 = bs[q], r1 = std::max(bs[q + 1], r0);
        if (r1 == r0) continue;
        const uint32_t T = p->opts.nsub ? p->opts.nsub
                                        : (uint32_t)std::max<uint64_t>(1, (r1 - r0 + brows_per_unit - 1) / brows_per_unit);
        std::vector<uint32_t> ts(T + 1);
        if ((rc = sp_partition_tsklt_bcsr(B->browptr, B->nnz_per_block, bs.data(), bl.data(), ns.data(), q, T,
                                          p->opts.sub_balance == SP_BAL_NNZ, ts.data())))
            return rc;
        ts[T] = r1;
        uint32_t prev = r0;
        for (uint32_t i = 0; i <= T; i++) {
            const uint32_t cut = std::min(std::max(ts[i], prev), r1);
            // cap a unit at 4x the even share (trailing empty block rows land in one tasklet)
            for (uint64_t c = (uint64_t)prev + 4 * brows_per_unit; c < cut; c += 4 * brows_per_unit)
                push_cut(units, (uint32_t)c);
            push_cut(units, cut);
            prev = cut;
        }
    }
    push_cut(units, nbr);
    if (nbr == 0) units.clear();
    p->nparts = P1;
    p->nsub = units.empty() ? 0 : (uint32_t)units.size() - 1;
    Builder b(p);
    b.add(P, units, nullptr);
    return b.finish(WK_BCSR_VEC, false, 1);
}

// ---------------------------------------------------------------------------
// BCOO (ids 7, 8)
// ---------------------------------------------------------------------------
int build_bcoo_plan(sp_plan *p, const sp_bcoo *B) {
    const size_t sv = sp_dtype_size(B->dtype);
    const uint32_t R = B->row_block_size, C = B->col_block_size;
    const size_t rc2 = (size_t)R * C;
    const uint64_t nb = B->num_blocks;
    if (nb > 0xFFFFFFFFULL) return SP_ERR_UNSUPPORTED;
    Prob P{};
    int rc;
    void *d;
    if ((rc = plan_upload(p, B->browind, nb * 4, &d))) return rc;
    P.rind = (const uint32_t *)d;
    if ((rc = plan_upload(p, B->bcolind, nb * 4, &d))) return rc;
    P.cind = (const uint32_t *)d;
    if ((rc = plan_upload(p, B->bval, nb * rc2 * sv, &d))) return rc;
    P.val = d;
    P.nrows = B->num_block_rows;
    P.nent = (uint32_t)nb;
    P.nrows_elem = B->nrows;
    P.ncols_elem = B->ncols;
    const bool atom = p->algo == SP_ALGO_SEGMENTED_ATOMIC;
    p->algo = atom ? SP_ALGO_SEGMENTED_ATOMIC : SP_ALGO_SEGMENTED;
    p->work.R = R;
    p->work.C = C;
    p->ipt = p->work.ipt = bcoo_ipt(B->dtype, R);
    const uint32_t tile = TPB * p->work.ipt;
    const uint32_t P1 = p->opts.nparts ? p->opts.nparts : 1;
    std::vector<uint32_t> bs(P1 + 1), ns(P1);
    if ((rc = sp_partition_bcoo(B->nnz_per_block, nb, B->nnz, P1, p->id == SP_K_BCOO_NNZ, bs.data(), ns.data())))
        return rc;
    bs[P1] = (uint32_t)nb;
    std::vector<uint32_t> units;
    push_cut(units, 0);
    for (uint32_t q = 0; q < P1; q++) {
        const uint32_t b0 = bs[q], b1 = std::max(bs[q + 1], b0);
        if (b1 == b0) continue;
        const uint32_t T = p->opts.nsub ? p->opts.nsub : 1;
        std::vector<uint32_t> ts(T + 1);
        if ((rc = sp_partition_tsklt_bcoo(B->nnz_per_block, bs.data(), ns.data(), q, T,
                                          p->opts.sub_balance == SP_BAL_NNZ, ts.data())))
            return rc;
        ts[T] = b1;
        for (uint32_t i = 0; i < T; i++) {
            const uint32_t s0 = std::min(std::max(ts[i], b0), b1), s1 = std::min(std::max(ts[i + 1], s0), b1);
            if (s1 > s0) tile_range(s0, s1, tile, units);
        }
    }
    push_cut(units, (uint32_t)nb);
    if (nb == 0) units.clear();
    p->nparts = P1;
    p->ntiles = units.empty() ? 0 : (uint32_t)units.size() - 1;
    Builder b(p);
    // element rows of block rows without blocks
    std::vector<uint8_t> has(B->num_block_rows, 0);
    for (uint64_t k = 0; k < nb; k++) has[B->browind[k]] = 1;
    for (uint32_t I = 0; I < B->num_block_rows; I++)
        if (!has[I])
            for (uint32_t i = 0; i < R && (uint64_t)I * R + i < B->nrows; i++) {
                b.zpid.push_back(0);
                b.zrow.push_back(I * R + i);
            }
    b.add(P, units, nullptr);
    if ((rc = b.finish(WK_BCOO_SEG, true, R))) return rc;
    return atom ? atomic_prezero(p, B->browind, units, R, B->nrows, b.zrow) : SP_OK;
}

// ---------------------------------------------------------------------------
// Execution
// ---------------------------------------------------------------------------
int run_plan(sp_plan *p, const void *x, void *y, int k, int ldx, int ldy) {
    if (p->nrows == 0) return SP_OK;
    if (!p->tiled) return run_work(p->work, p->ctx, x, y, nullptr, nullptr, k, ldx, ldy);
    // 2D: partial rows per stripe into ypart, then the stripe-order reduction.
    const uint64_t need = (uint64_t)p->vert * p->nrows * (uint64_t)k * sp_dtype_size(p->dtype);
    if (need > p->ypart_bytes) {
        void *d = nullptr;
        int rc = plan_alloc(p, need, &d);
        if (rc) return rc;
        p->d_ypart = d;
        p->ypart_bytes = need;
    }
    Work w = p->work;
    w.zero_rows = w.zero_all ? (uint64_t)p->vert * p->nrows : 0;
    int rc = run_work(w, p->ctx, x, p->d_ypart, nullptr, nullptr, k, ldx, k);
    if (rc) return rc;
    return launch_reduce_parts(p->ctx->stream, p->dtype, p->d_ypart, p->vert, (uint64_t)p->nrows * k, y,
                               p->nrows, k, ldy);
}

int run_work(const Work &w, sp_ctx *ctx, const void *xb, void *yb, const uintptr_t *xtab, const uintptr_t *ytab,
             int k, int ldx, int ldy) {
    switch (w.dtype) {
    case SP_I8: return run_work_t<int8_t>(w, ctx, xb, yb, xtab, ytab, k, ldx, ldy);
    case SP_I16: return run_work_t<int16_t>(w, ctx, xb, yb, xtab, ytab, k, ldx, ldy);
    case SP_I32: return run_work_t<int32_t>(w, ctx, xb, yb, xtab, ytab, k, ldx, ldy);
    case SP_I64: return run_work_t<int64_t>(w, ctx, xb, yb, xtab, ytab, k, ldx, ldy);
    case SP_F32: return run_work_t<float>(w, ctx, xb, yb, xtab, ytab, k, ldx, ldy);
    case SP_F64: return run_work_t<double>(w, ctx, xb, yb, xtab, ytab, k, ldx, ldy);
    default: return SP_ERR_ARG;
    }
}

// ---------------------------------------------------------------------------
// 2D tilings (ids 9..24): one sub-problem per tile, each writing its
// stripe's partial rows into ypart[stripe][row]; the reduction then adds
// stripes in order (the reference host merge, spmv/2D/DCSR/host/app.c:340-351).
// ---------------------------------------------------------------------------
namespace {

struct TiledKind {
    int tiling, format, by_nnz;
};

TiledKind tiled_kind(int id) {
    switch (id) {
    case SP_K_DCSR: return {SP_TILE_D, SP_FMT_CSR, 0};
    case SP_K_DCOO: return {SP_TILE_D, SP_FMT_COO, 0};
    case SP_K_DBCSR: return {SP_TILE_D, SP_FMT_BCSR, 0};
    case SP_K_DBCOO: return {SP_TILE_D, SP_FMT_BCOO, 0};
    case SP_K_RBDCSR: return {SP_TILE_RB, SP_FMT_CSR, 0};
    case SP_K_RBDCOO: return {SP_TILE_RB, SP_FMT_COO, 0};
    case SP_K_RBDBCSR_BLOCK: return {SP_TILE_RB, SP_FMT_BCSR, 0};
    case SP_K_RBDBCSR_NNZ: return {SP_TILE_RB, SP_FMT_BCSR, 1};
    case SP_K_RBDBCOO_BLOCK: return {SP_TILE_RB, SP_FMT_BCOO, 0};
    case SP_K_RBDBCOO_NNZ: return {SP_TILE_RB, SP_FMT_BCOO, 1};
    case SP_K_BDCSR: return {SP_TILE_BD, SP_FMT_CSR, 0};
    case SP_K_BDCOO: return {SP_TILE_BD, SP_FMT_COO, 0};
    case SP_K_BDBCSR_BLOCK: return {SP_TILE_BD, SP_FMT_BCSR, 0};
    case SP_K_BDBCSR_NNZ: return {SP_TILE_BD, SP_FMT_BCSR, 1};
    case SP_K_BDBCOO_BLOCK: return {SP_TILE_BD, SP_FMT_BCOO, 0};
    default: return {SP_TILE_BD, SP_FMT_BCOO, 1};
    }
}

}  // namespace

int build_tiled_plan(sp_plan *p, const sp_coo *A0) {
    const TiledKind tk = tiled_kind(p->id);
    const uint32_t n = A0->nrows;
    const size_t sv = sp_dtype_size(A0->dtype);
    uint32_t vert = p->opts.vert ? p->opts.vert : 2;
    if (vert > A0->ncols && A0->ncols) vert = A0->ncols;
    if (vert == 0) vert = 1;
    // horz: D tiles per column stripe (ex NR_DPUS / vert); RB/BD row parts per stripe.
    uint32_t horz;
    if (tk.tiling == SP_TILE_D) {
        horz = p->opts.nparts ? std::max<uint32_t>(1, sp_find_partitions(p->opts.nparts, vert))
                              : std::max<uint32_t>(1, (n + 8191) / 8192);
        if (horz > n && n) horz = n;
    } else {
        horz = p->opts.nparts ? std::max<uint32_t>(1, sp_find_partitions(p->opts.nparts, vert)) : 1;
    }
    // tilings need the sorted COO
    sp_coo *sorted = nullptr;
    const sp_coo *A = A0;
    if (!coo_sorted(A0)) {
        int rc = sp_coo_alloc(A0->nrows, A0->ncols, A0->nnz, A0->dtype, &sorted);
        if (rc) return rc;
        memcpy(sorted->rowind, A0->rowind, (size_t)A0->nnz * 4);
        memcpy(sorted->colind, A0->colind, (size_t)A0->nnz * 4);
        memcpy(sorted->values, A0->values, (size_t)A0->nnz * sv);
        if ((rc = sp_sort_coo(sorted))) { sp_free_coo(sorted); return rc; }
        A = sorted;
    }
    sp_tiled *T = nullptr;
    int rc = sp_coo2tiled(A, tk.tiling, tk.format, horz, vert, p->opts.row_block_size, p->opts.col_block_size,
                          p->opts.count_values, &T);
    // row statistics for the CSR algorithm and the tile size choice
    uint32_t maxlen = 0;
    if (!rc && (tk.format == SP_FMT_CSR || tk.format == SP_FMT_COO)) {
        std::vector<uint32_t> cnt(n + 1, 0);
        for (uint32_t k = 0; k < A->nnz; k++) cnt[A->rowind[k]]++;
        for (uint32_t i = 0; i < n; i++) maxlen = std::max(maxlen, cnt[i]);
    }
    if (sorted) sp_free_coo(sorted);
    if (rc) return rc;

    p->tiled = true;
    p->vert = vert;
    p->ntiles = T->ntiles;
    p->nparts = tk.tiling == SP_TILE_D ? T->ntiles : vert * horz;
    const uint32_t R = p->opts.row_block_size, C = p->opts.col_block_size;
    const uint32_t nt = T->ntiles, th = T->tile_height;
    Builder b(p);
    void *d;
    // every tile's rows made of aligned runs of four columns, in stripes starting at a multiple
    // of 4: the chunked row kernel reads x four values per load
    bool quads = tk.format == SP_FMT_CSR && !p->opts.lds_slice &&
                 (p->algo == SP_ALGO_AUTO || (p->algo == SP_ALGO_VECTOR && p->opts.items_per_lane == 4));
    for (uint32_t t = 0; t < nt && quads; t++) {
        const uint32_t c = t % vert, w0 = T->vert_tile_widths[c], w = T->vert_tile_widths[c + 1] - w0;
        const uint32_t cnt = T->nnzs_per_partition[t];
        quads = w0 % 4 == 0 &&
                (cnt == 0 || quad_rows(T->drowptr + (size_t)t * (th + 1), th, T->dcolind + T->nnz_offset[t], cnt, w));
    }
    if (quads && p->algo == SP_ALGO_AUTO) {
        p->algo = SP_ALGO_VECTOR;
        p->opts.items_per_lane = 4;
    }
    // 2D CSR tiles in the chunked row kernel (VECTOR with items_per_lane 4: asked for, or AUTO
    // with quads; otherwise AUTO keeps the merge path, which measured the same on the 27-point
    // stencil at two stripes, 313 us each, with four gathers per chunk), with every tile's
    // entries starting at a multiple of 4 in the arrays
    const bool csr_chunk = tk.format == SP_FMT_CSR && p->algo == SP_ALGO_VECTOR && p->opts.items_per_lane == 4 &&
                           !p->opts.lds_slice;
    // 2D COO tiles of aligned four-column runs (every group of four entries of a tile one row's
    // columns c .. c+3, c a multiple of 4, stripes starting at a multiple of 4): row-ordered
    // 4-entry chunk tiles reading x four values per load, in the same padded layout
    bool coo_quads = tk.format == SP_FMT_COO && !p->opts.lds_slice && p->algo != SP_ALGO_SEGMENTED_ATOMIC &&
                     p->algo != SP_ALGO_SEGMENTED_SORTED;
    for (uint32_t t = 0; t < nt && coo_quads; t++) {
        const uint32_t w0 = T->vert_tile_widths[t % vert], cnt = T->nnzs_per_partition[t];
        const uint32_t *r = T->drowind + T->nnz_offset[t], *c = T->dcolind + T->nnz_offset[t];
        coo_quads = w0 % 4 == 0 && cnt % 4 == 0;
        for (uint32_t j = 0; j < cnt && coo_quads; j += 4)
            coo_quads = r[j + 1] == r[j] && r[j + 2] == r[j] && r[j + 3] == r[j] && c[j] % 4 == 0 &&
                        c[j + 1] == c[j] + 1 && c[j + 2] == c[j] + 2 && c[j + 3] == c[j] + 3;
    }
    const bool padded = csr_chunk || coo_quads;
    std::vector<uint64_t> e0c;  // chunked layout: each tile's first entry in the uploaded arrays
    do {
        // upload the tiling's arrays once; tiles point into them
        const uint32_t *d_ptr = nullptr, *d_rind = nullptr, *d_cind = nullptr;
        const char *d_val = nullptr;
        const bool blk = tk.format >= SP_FMT_BCSR;
        const uint64_t nent = blk ? T->num_blocks : T->nnz;
        const size_t esz = blk ? (size_t)R * C * sv : sv;
        if (tk.format == SP_FMT_CSR) {
            if ((rc = plan_upload(p, T->drowptr, (size_t)nt * (th + 1) * 4, &d))) break;
            d_ptr = (const uint32_t *)d;
        } else if (tk.format == SP_FMT_BCSR) {
            if ((rc = plan_upload(p, T->browptr, (size_t)nt * (T->num_block_rows + 1) * 4, &d))) break;
            d_ptr = (const uint32_t *)d;
        }
        if ((tk.format == SP_FMT_COO && !coo_quads) || tk.format == SP_FMT_BCOO) {
            if ((rc = plan_upload(p, tk.format == SP_FMT_COO ? T->drowind : T->browind, nent * 4, &d))) break;
            d_rind = (const uint32_t *)d;
        }
        if (padded) {
            // tile t's entries at e0c[t], a multiple of 4 (the gaps hold zeros)
            e0c.resize(nt);
            uint64_t tot = 0;
            for (uint32_t t = 0; t < nt; t++) {
                e0c[t] = tot;
                tot += ((uint64_t)T->nnzs_per_partition[t] + 3) & ~(uint64_t)3;
            }
            std::vector<uint32_t> ci(tot, 0);
            std::vector<char> va(tot * sv, 0);
            for (uint32_t t = 0; t < nt; t++) {
                const uint64_t src = T->nnz_offset[t], cnt = T->nnzs_per_partition[t];
                memcpy(ci.data() + e0c[t], T->dcolind + src, cnt * 4);
                memcpy(va.data() + e0c[t] * sv, (const char *)T->dval + src * sv, cnt * sv);
            }
            if (coo_quads) {
                std::vector<uint32_t> ri(tot, 0);
                for (uint32_t t = 0; t < nt; t++)
                    memcpy(ri.data() + e0c[t], T->drowind + T->nnz_offset[t], (size_t)T->nnzs_per_partition[t] * 4);
                if ((rc = plan_upload(p, ri.data(), tot * 4, &d))) break;
                d_rind = (const uint32_t *)d;
            }
            if ((rc = plan_upload(p, ci.data(), tot * 4, &d))) break;
            d_cind = (const uint32_t *)d;
            if ((rc = plan_upload(p, va.data(), tot * sv, &d))) break;
            d_val = (const char *)d;
        } else {
            if ((rc = plan_upload(p, blk ? T->bcolind : T->dcolind, nent * 4, &d))) break;
            d_cind = (const uint32_t *)d;
            if ((rc = plan_upload(p, blk ? T->bval : T->dval, nent * esz, &d))) break;
            d_val = (const char *)d;
        }

        std::vector<uint32_t> split;
        if (tk.tiling != SP_TILE_D) {
            split.resize((size_t)vert * (horz + 1));
            if ((rc = sp_partition_tiled(T, horz, tk.by_nnz, split.data()))) break;
        }
        int kind;
        const bool atom = p->algo == SP_ALGO_SEGMENTED_ATOMIC;  // COO/BCOO tiles (the id was checked)
        bool col_order = false;  // column-ordered tiles (CSR merge path, COO), as for 1D
        if (tk.format == SP_FMT_CSR) {
            if (p->algo == SP_ALGO_AUTO) {
                const double avg = n ? (double)A0->nnz / n : 0.0;
                // as for 1D CSR; stripes also leave row ranges empty, which the merge path
                // skips at no cost
                p->algo = (maxlen <= 4 * avg + 64 && A0->nnz <= (16u << 20)) ? SP_ALGO_VECTOR
                                                                             : SP_ALGO_SEGMENTED_SORTED;
            }
            if (p->algo == SP_ALGO_SEGMENTED_HOT) p->algo = SP_ALGO_SEGMENTED_SORTED;  // hot x: 1D CSR only
            kind = p->algo == SP_ALGO_VECTOR ? WK_CSR_VEC : WK_CSR_MERGE;
            col_order = p->algo == SP_ALGO_SEGMENTED_SORTED;
        } else if (tk.format == SP_FMT_COO) {
            kind = WK_COO_SEG;
            col_order = p->algo != SP_ALGO_SEGMENTED && !coo_quads;  // as for 1D COO
            p->algo = atom ? SP_ALGO_SEGMENTED_ATOMIC : SP_ALGO_SEGMENTED;
        } else if (tk.format == SP_FMT_BCSR) {
            kind = WK_BCSR_VEC;
            p->algo = SP_ALGO_VECTOR;
        } else {
            kind = WK_BCOO_SEG;
            p->algo = atom ? SP_ALGO_SEGMENTED_ATOMIC : SP_ALGO_SEGMENTED;
        }
        // atomic merges need no pre-zero list here: the partial buffers are zeroed first
        if (atom && p->opts.lds_slice) { rc = SP_ERR_UNSUPPORTED; break; }
        // skewed rows and more than 8M entries: 4096-entry tiles, as for 1D CSR/COO
        if ((kind == WK_CSR_MERGE || kind == WK_COO_SEG) && col_order && p->opts.items_per_lane == 0 &&
            !p->opts.lds_slice && maxlen > 4.0 * A0->nnz / std::max(n, 1u) + 64 && A0->nnz > (8u << 20) &&
            A0->ncols <= (1u << 20))
            p->opts.items_per_lane = 16;
        const uint32_t I = pick_ipt(p->opts, 8);
        if (!I) { rc = SP_ERR_ARG; break; }
        // x slices staged in LDS (2D CSR/COO, the reference's per-DPU x slice,
        // spmv/2D/DCSR/host/app.c:308-315): every stripe must fit the asked slice
        // and the kernels' LDS budget; the merge-path kernel carries the CSR case
        if (p->opts.lds_slice && (tk.format == SP_FMT_CSR || tk.format == SP_FMT_COO)) {
            uint32_t maxw = 0;
            for (uint32_t c = 0; c < vert; c++)
                maxw = std::max(maxw, T->vert_tile_widths[c + 1] - T->vert_tile_widths[c]);
            if (maxw > p->opts.lds_slice || maxw > spd::SLICE_BYTES / sv || I > 8) {
                rc = SP_ERR_UNSUPPORTED;
                break;
            }
            if (kind == WK_CSR_VEC) {
                kind = WK_CSR_MERGE;
                p->algo = SP_ALGO_SEGMENTED;
            }
            p->work.slice = true;
            p->work.grid_cap = 3 * (uint32_t)std::max(1, p->ctx->num_cus);  // three workgroups per CU fit
        }
        uint32_t L = p->opts.lanes_per_row;
        if (kind == WK_CSR_VEC && L == 0)
            L = csr_chunk ? chunk_lanes(n ? (double)A0->nnz / n / vert : 1.0) : auto_lanes(n ? (double)A0->nnz / n / vert : 1.0);
        if (csr_chunk && (L < 4 || L > 64 || (L & (L - 1)))) { rc = SP_ERR_ARG; break; }
        p->work.chunk = padded;
        p->work.quad = (csr_chunk && quads) || coo_quads;
        bool coal = false;
        if (kind == WK_BCSR_VEC && L == 0) {
            const double avg = T->num_block_rows ? (double)T->num_blocks / ((double)T->num_block_rows * nt) : 1.0;
            coal = coal_ok(A0->dtype, R, C, avg);
            L = coal ? 64 : avg >= 48 ? 64 : avg >= 12 ? 16 : avg >= 3 ? 4 : 1;
        }
        p->lanes = p->work.lanes = coal ? 0 : L;
        p->ipt = p->work.ipt = kind == WK_BCOO_SEG ? bcoo_ipt(A0->dtype, R) : I;
        p->work.R = R;
        p->work.C = C;
        const uint32_t tile_items = TPB * p->work.ipt;
        std::vector<uint32_t> order_cuts;  // column-ordered tiles: unit cuts over all entries
        for (uint32_t t = 0; t < nt && !rc; t++) {
            const uint32_t c = t % vert;
            const uint32_t row_off = tk.tiling == SP_TILE_D ? (t / vert) * th : 0;
            const uint32_t rows = row_off < n ? std::min(th, n - row_off) : 0;
            Prob P{};
            P.nrows_elem = rows;
            P.ncols_elem = T->vert_tile_widths[c + 1] - T->vert_tile_widths[c];
            P.xoff = T->vert_tile_widths[c];
            P.yoff = (int64_t)c * n + row_off;
            std::vector<uint32_t> k0, k1;
            if (!blk) {
                const uint64_t e0 = T->nnz_offset[t], e0u = padded ? e0c[t] : e0;
                P.nent = T->nnzs_per_partition[t];
                P.cind = d_cind + e0u;
                P.val = d_val + e0u * sv;
                P.nrows = rows;
                if (tk.format == SP_FMT_CSR) {
                    P.ptr = d_ptr + (size_t)t * (th + 1);
                    const uint32_t *hrp = T->drowptr + (size_t)t * (th + 1);
                    if (rows == 0) {
                    } else if (kind == WK_CSR_MERGE) {
                        merge_coords(hrp, rows, hrp[rows], tile_items, k0, k1);
                    } else {
                        const uint32_t gran = (uint32_t)(8 / sv);
                        const uint64_t rpu = 4ULL * (TPB / L);
                        push_cut(k0, 0);
                        const uint32_t parts = tk.tiling == SP_TILE_D ? 1 : horz;
                        for (uint32_t q = 0; q < parts && !rc; q++) {
                            const uint32_t r0 = tk.tiling == SP_TILE_D ? 0 : std::min(split[(size_t)c * (horz + 1) + q], rows);
                            const uint32_t r1 = tk.tiling == SP_TILE_D ? rows : std::min(std::max(split[(size_t)c * (horz + 1) + q + 1], r0), rows);
                            if (r1 == r0) continue;
                            const uint32_t Tn = p->opts.nsub ? p->opts.nsub
                                                             : (uint32_t)std::max<uint64_t>(1, (r1 - r0 + rpu - 1) / rpu);
                            rc = row_units(hrp, r0, r1, Tn, p->opts.sub_balance == SP_BAL_NNZ, gran, (uint32_t)(4 * rpu), k0);
                        }
                        push_cut(k0, rows);
                    }
                } else {  // COO tile
                    P.rind = d_rind + e0u;
                    const uint32_t cnt = P.nent;
                    if (cnt) {
                        if (tk.tiling == SP_TILE_D) {
                            tile_range(0, cnt, tile_items, k0);
                        } else {
                            push_cut(k0, 0);
                            for (uint32_t q = 0; q < horz; q++) {
                                const uint32_t s0 = std::min(split[(size_t)c * (horz + 1) + q], cnt);
                                const uint32_t s1 = std::min(std::max(split[(size_t)c * (horz + 1) + q + 1], s0), cnt);
                                if (s1 > s0) tile_range(s0, s1, tile_items, k0);
                            }
                            push_cut(k0, cnt);
                        }
                    }
                }
            } else {
                const uint64_t b0 = T->block_offset[t];
                const uint32_t nbr_t = (rows + R - 1) / R;
                P.nent = T->blocks_per_partition[t];
                P.cind = d_cind + b0;
                P.val = d_val + b0 * esz;
                P.nrows = nbr_t;
                if (tk.format == SP_FMT_BCSR) {
                    P.ptr = d_ptr + (size_t)t * (T->num_block_rows + 1);
                    const uint32_t *hbp = T->browptr + (size_t)t * (T->num_block_rows + 1);
                    if (nbr_t) {
                        const uint64_t bpu = 4ULL * (TPB / L);
                        push_cut(k0, 0);
                        const uint32_t parts = tk.tiling == SP_TILE_D ? 1 : horz;
                        for (uint32_t q = 0; q < parts; q++) {
                            const uint32_t r0 = tk.tiling == SP_TILE_D ? 0 : std::min(split[(size_t)c * (horz + 1) + q], nbr_t);
                            const uint32_t r1 = tk.tiling == SP_TILE_D ? nbr_t : std::min(std::max(split[(size_t)c * (horz + 1) + q + 1], r0), nbr_t);
                            if (r1 == r0) continue;
                            const uint32_t Tn = p->opts.nsub ? p->opts.nsub
                                                             : (uint32_t)std::max<uint64_t>(1, (r1 - r0 + bpu - 1) / bpu);
                            // even block-row sub-parts (the tasklet split is a work assignment only)
                            for (uint32_t s = 0; s <= Tn; s++) push_cut(k0, r0 + (uint32_t)((uint64_t)(r1 - r0) * s / Tn));
                        }
                        push_cut(k0, nbr_t);
                        (void)hbp;
                    }
                } else {  // BCOO tile
                    P.rind = d_rind + b0;
                    const uint32_t cnt = P.nent;
                    if (cnt) {
                        if (tk.tiling == SP_TILE_D) {
                            tile_range(0, cnt, tile_items, k0);
                        } else {
                            push_cut(k0, 0);
                            for (uint32_t q = 0; q < horz; q++) {
                                const uint32_t s0 = std::min(split[(size_t)c * (horz + 1) + q], cnt);
                                const uint32_t s1 = std::min(std::max(split[(size_t)c * (horz + 1) + q + 1], s0), cnt);
                                if (s1 > s0) tile_range(s0, s1, tile_items, k0);
                            }
                            push_cut(k0, cnt);
                        }
                    }
                }
            }
            if (col_order && !blk) {  // the tile's units as absolute entry cuts
                const std::vector<uint32_t> &cuts = kind == WK_CSR_MERGE ? k1 : k0;
                const uint64_t e0 = T->nnz_offset[t];
                for (size_t u = 0; u < cuts.size(); u++) {
                    if (u + 1 < cuts.size() && cuts[u + 1] - cuts[u] > tile_items) col_order = false;
                    push_cut(order_cuts, (uint32_t)(e0 + cuts[u]));
                }
            }
            b.add(P, k0, kind == WK_CSR_MERGE ? &k1 : nullptr);
        }
        if (rc) break;
        // Column-ordered tiles (as for 1D CSR/COO): every unit's entries sorted by local column,
        // with the row-order position in the low bits; the arrays uploaded above are replaced.
        if (col_order && !blk && T->nnz) {
            uint32_t pb = 0;
            while ((1u << pb) < tile_items) pb++;
            if (A0->ncols <= (1ull << (32 - pb)) && !order_cuts.empty() && order_cuts.front() == 0 &&
                order_cuts.back() == T->nnz) {
                std::vector<uint32_t> key;
                std::vector<char> val;
                column_ordered_tiles(T->dcolind, T->dval, (uint32_t)T->nnz, order_cuts, pb, sv, key, val);
                if (hipMemcpy((void *)d_cind, key.data(), (size_t)T->nnz * 4, hipMemcpyHostToDevice) != hipSuccess ||
                    hipMemcpy((void *)d_val, val.data(), (size_t)T->nnz * sv, hipMemcpyHostToDevice) != hipSuccess) {
                    rc = SP_ERR_HIP;
                    break;
                }
                for (Prob &Q : b.probs) Q.pbits = pb;
            } else {
                col_order = false;
            }
        }
        if (tk.format == SP_FMT_CSR && kind == WK_CSR_MERGE)
            p->algo = col_order ? SP_ALGO_SEGMENTED_SORTED : SP_ALGO_SEGMENTED;
        else if (tk.format == SP_FMT_COO)
            p->algo = atom ? SP_ALGO_SEGMENTED_ATOMIC : col_order ? SP_ALGO_SEGMENTED_SORTED : SP_ALGO_SEGMENTED;
        // algorithmic bytes: the tile formats' arrays, x once and y once (the partial
        // vectors and their reduction are overhead, not algorithm)
        {
            const uint64_t xy = ((uint64_t)n + A0->ncols) * sv;
            if (tk.format == SP_FMT_CSR)
                p->algo_bytes = (uint64_t)nt * (th + 1) * 4 + (uint64_t)T->nnz * (4 + sv) + xy;
            else if (tk.format == SP_FMT_COO)
                p->algo_bytes = (uint64_t)T->nnz * (8 + sv) + xy;
            else if (tk.format == SP_FMT_BCSR)
                p->algo_bytes = T->num_blocks * ((uint64_t)R * C * sv + 4) +
                                (uint64_t)nt * (T->num_block_rows + 1) * 4 + xy;
            else
                p->algo_bytes = T->num_blocks * ((uint64_t)R * C * sv + 8) + xy;
        }
        const bool carries = kind == WK_CSR_MERGE || kind == WK_COO_SEG || kind == WK_BCOO_SEG;
        rc = b.finish(kind, carries, kind == WK_BCOO_SEG ? R : 1);
        p->work.zero_all = (kind == WK_COO_SEG || kind == WK_BCOO_SEG);
        if (atom) {
            p->work.atomic = true;
            p->launches = 1;
        }
        p->nsub = p->work.nunits;
    } while (0);
    sp_free_tiled(T);
    return rc;
}
