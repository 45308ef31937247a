"""Pins of the oracle's 2D tile-major spmv_host (orc_spmv_tiled, the
restatement of spmv/2D/DCSR/host/app.c:84-103, DCOO:80-95, RBDCSR:88-103,
BDBCSR-block:92-114 and their siblings) and of its vertical merge
(orc_merge_parts, spmv/2D/DCSR/host/app.c:340-351) for floating point.

What fixes the answer, independently of the restatement:
  * integers: exact modular dense arithmetic (Python integers reduced mod
    2^bits), for every tiling (D, RB, BD) x format (CSR, COO, BCSR, BCOO);
  * floats with small-integer values: every product and partial sum is an
    integer below 2^24, so any summation order is exact and the FP32 / FP64
    results must equal the integer result bit for bit;
  * general floats: the rounding bound gamma(len+2) * sum|a x| against the
    long-double reference;
  * FP32 CSR tilings: the stripe-sum identity y = sum over tiles (in tile
    order) of the 1D CSR spmv_host of each tile on its x slice, computed
    from the tile arrays with the pinned 1D oracle, must hold bit for bit.
Planted mistakes (every tile read at column offset 0, rows of a D tiling
all at offset 0) must fail these checks."""
import numpy as np
import pytest

import oracle_lib as orc
from oracle_lib import tile_args

BITS = {0: 8, 1: 16, 2: 32, 3: 64}
TILINGS = (0, 1, 2)   # D, RB, BD
FORMATS = (0, 1, 2, 3)  # CSR, COO, BCSR, BCOO
SHAPES = [(64, 80, 900, 4, 2), (100, 37, 700, 3, 5), (200, 300, 3000, 8, 8), (33, 1000, 500, 2, 16),
          (57, 61, 1, 2, 3)]


def wrap(acc, bits, dt):
    m = 1 << bits
    out = []
    for v in acc:
        v %= m
        out.append(v - m if v >= m // 2 else v)
    return np.array(out, dtype=orc.NP[dt])


def dense_modular(nrows, r, c, v, x, dt):
    acc = [0] * nrows
    for i, j, a in zip(r.tolist(), c.tolist(), v.tolist()):
        acc[i] += int(a) * int(x[j])
    return wrap(acc, BITS[dt], dt)


def coo_case(sp, rng, dt, nrows, ncols, nnz, small_ints=False):
    keys = np.sort(rng.choice(nrows * ncols, nnz, replace=False))
    r, c = (keys // ncols).astype(np.uint32), (keys % ncols).astype(np.uint32)
    if small_ints:
        v = rng.integers(-4, 5, nnz).astype(orc.NP[dt])
        x = rng.integers(-4, 5, ncols).astype(orc.NP[dt])
    elif dt <= 3:
        info = np.iinfo(orc.NP[dt])
        v = rng.integers(info.min, info.max, nnz, dtype=np.int64, endpoint=True).astype(orc.NP[dt])
        x = rng.integers(info.min, info.max, ncols, dtype=np.int64, endpoint=True).astype(orc.NP[dt])
    else:
        v = rng.uniform(-1, 1, nnz).astype(orc.NP[dt])
        x = rng.uniform(-1, 1, ncols).astype(orc.NP[dt])
    return r, c, v, x, sp.Coo.from_arrays(nrows, ncols, r, c, v)


def tiled_y(dt, T, nrows, ncols, x, **override):
    a = tile_args(T, nrows)
    a.update(override)
    return orc.spmv_tiled(dt, T.format, nrows, ncols, a["row_off"], a["rows"], a["col_off"], a["col_w"],
                          a["ptr_off"], a["ent_off"], a["ptr"], a["ri"], a["ci"], a["R"], a["C"], a["vals"], x)


def build(sp, A, tiling, fmt, horz, vert, rc=(4, 4)):
    return sp.coo2tiled(A, tiling, fmt, horz, vert, rc[0], rc[1], count_values=False)


@pytest.mark.parametrize("dt", [0, 1, 2, 3])
@pytest.mark.parametrize("fmt", FORMATS)
def test_tiled_integers_match_modular_dense(sp, dt, fmt):
    rng = np.random.default_rng(100 + 10 * dt + fmt)
    for (nr, nc, nnz, horz, vert) in SHAPES:
        r, c, v, x, A = coo_case(sp, rng, dt, nr, nc, nnz)
        want = dense_modular(nr, r, c, v, x, dt)
        for tiling in TILINGS:
            for rc in ((4, 4), (2, 3)) if fmt >= 2 else ((4, 4),):
                T = build(sp, A, tiling, fmt, horz, vert, rc)
                got = tiled_y(dt, T, nr, nc, x)
                assert np.array_equal(got, want), (tiling, fmt, rc, (nr, nc, horz, vert))


@pytest.mark.parametrize("fmt", FORMATS)
def test_planted_tiling_mistakes_fail(sp, fmt):
    rng = np.random.default_rng(130 + fmt)
    nr, nc, nnz, horz, vert = 64, 80, 900, 4, 4
    r, c, v, x, A = coo_case(sp, rng, 2, nr, nc, nnz)
    want = dense_modular(nr, r, c, v, x, 2)
    for tiling in TILINGS:
        T = build(sp, A, tiling, fmt, horz, vert)
        a = tile_args(T, nr)
        assert np.array_equal(tiled_y(2, T, nr, nc, x), want)
        # every tile read as if it started at column 0
        wrong = tiled_y(2, T, nr, nc, x, col_off=np.zeros_like(a["col_off"]))
        assert not np.array_equal(wrong, want), ("col_off", tiling, fmt)
        if tiling == 0:
            wrong = tiled_y(2, T, nr, nc, x, row_off=np.zeros_like(a["row_off"]))
            assert not np.array_equal(wrong, want), ("row_off", fmt)


@pytest.mark.parametrize("dt", [4, 5])
@pytest.mark.parametrize("fmt", FORMATS)
def test_tiled_floats_exact_on_small_integers(sp, dt, fmt):
    rng = np.random.default_rng(140 + 10 * dt + fmt)
    for (nr, nc, nnz, horz, vert) in SHAPES:
        r, c, v, x, A = coo_case(sp, rng, dt, nr, nc, nnz, small_ints=True)
        want = dense_modular(nr, r, c, v.astype(np.int64), x.astype(np.int64), 3).astype(orc.NP[dt])
        for tiling in TILINGS:
            T = build(sp, A, tiling, fmt, horz, vert)
            assert np.array_equal(tiled_y(dt, T, nr, nc, x), want), (tiling, fmt, (nr, nc, horz, vert))


@pytest.mark.parametrize("dt", [4, 5])
@pytest.mark.parametrize("fmt", FORMATS)
def test_tiled_floats_within_bound(sp, dt, fmt):
    rng = np.random.default_rng(160 + 10 * dt + fmt)
    for (nr, nc, nnz, horz, vert) in SHAPES:
        r, c, v, x, A = coo_case(sp, rng, dt, nr, nc, nnz)
        C = A.to_csr()
        yr, b = orc.ref_csr(dt, C.rowptr, C.colind, C.values, x)
        tol = orc.fp_tolerance(dt, np.diff(C.rowptr.astype(np.int64)), b)
        for tiling in TILINGS:
            T = build(sp, A, tiling, fmt, horz, vert)
            got = tiled_y(dt, T, nr, nc, x).astype(np.float64)
            assert np.all(np.abs(got - yr) <= tol), (tiling, fmt)


@pytest.mark.parametrize("tiling", TILINGS)
def test_fp32_csr_stripe_sum_identity(sp, tiling):
    """Tile-major CSR: each tile's per-row sum is formed in entry order and
    added to y once, tiles in order.  The same sums come from the pinned 1D
    CSR oracle run on every tile alone (its rowptr slice, its columns
    shifted by the tile's column offset) and added up in tile order."""
    rng = np.random.default_rng(170 + tiling)
    for (nr, nc, nnz, horz, vert) in SHAPES:
        r, c, v, x, A = coo_case(sp, rng, 4, nr, nc, nnz)
        T = build(sp, A, tiling, 0, horz, vert)
        a = tile_args(T, nr)
        acc = np.zeros(nr, np.float32)
        for t in range(T.ntiles):
            rp = T.drowptr[a["ptr_off"][t]:a["ptr_off"][t] + T.tile_height + 1].astype(np.uint32)
            e0 = int(a["ent_off"][t])
            ci = (T.dcolind[e0:e0 + int(rp[-1])] + np.uint32(a["col_off"][t])).astype(np.uint32)
            vt = np.ascontiguousarray(T.dval[e0:e0 + int(rp[-1])])
            part = orc.spmv_csr(4, rp, ci, vt, x)
            ro = int(a["row_off"][t])
            n = min(T.tile_height, nr - ro)
            if n > 0:
                acc[ro:ro + n] = acc[ro:ro + n] + part[:n]
        assert np.array_equal(tiled_y(4, T, nr, nc, x), acc), (tiling, (nr, nc, horz, vert))


@pytest.mark.parametrize("dt", [4, 5])
def test_float_merge_parts(sp, dt):
    """orc_merge_parts adds part c to part 0 for c = 1, 2, ... in order.
    Integer-valued floats make any order exact; general floats must match
    the left-to-right sum of IEEE adds, and the reverse order (a planted
    mistake) must not."""
    rng = np.random.default_rng(180 + dt)
    ip = rng.integers(-1000, 1000, (7, 500))
    got = orc.merge_parts(dt, ip.astype(orc.NP[dt]))
    assert np.array_equal(got, ip.sum(axis=0).astype(orc.NP[dt]))
    parts = (rng.standard_normal((7, 4000)) * 10.0 ** rng.uniform(-4, 4, (7, 4000))).astype(orc.NP[dt])
    fwd = parts[0].copy()
    for c in range(1, 7):
        fwd = fwd + parts[c]
    rev = parts[6].copy()
    for c in range(5, -1, -1):
        rev = rev + parts[c]
    assert not np.array_equal(fwd, rev)
    for nth in (1, 4):
        got = orc.merge_parts(dt, parts.copy(), nthreads=nth)
        assert np.array_equal(got, fwd)
