"""Pins of the CPU oracle (oracle/sp_oracle.c) by what the mathematics fixes:
integer SpMV against exact modular dense arithmetic (Python integers reduced
mod 2^bits) for every format, float SpMV against closed forms (identity,
diagonal, stencil row sums), against scipy.sparse within the rounding bound,
and the reference's sequential per-row order against an explicit float32
loop; a planted mistake (transposed operand) must fail the dense check."""
import math

import numpy as np
import pytest

import oracle_lib as orc

BITS = {0: 8, 1: 16, 2: 32, 3: 64}


def wrap(vals, bits):
    m = 1 << bits
    out = []
    for v in vals:
        v %= m
        out.append(v - m if v >= m // 2 else v)
    return np.array(out, dtype=orc.NP[{8: 0, 16: 1, 32: 2, 64: 3}[bits]])


def dense_modular(nrows, r, c, v, x, bits):
    acc = [0] * nrows
    for i, j, a in zip(r.tolist(), c.tolist(), v.tolist()):
        acc[i] += int(a) * int(x[j])
    return wrap(acc, bits)


def int_case(sp, rng, dt, nrows=37, ncols=29, nnz=400):
    info = np.iinfo(orc.NP[dt])
    keys = np.sort(rng.choice(nrows * ncols, nnz, replace=False))
    r, c = (keys // ncols).astype(np.uint32), (keys % ncols).astype(np.uint32)
    v = rng.integers(info.min, info.max, nnz, dtype=np.int64, endpoint=True).astype(orc.NP[dt])
    x = rng.integers(info.min, info.max, ncols, dtype=np.int64, endpoint=True).astype(orc.NP[dt])
    return r, c, v, x, sp.Coo.from_arrays(nrows, ncols, r, c, v)


@pytest.mark.parametrize("dt", [0, 1, 2, 3])
def test_integer_formats_match_modular_dense(sp, dt):
    rng = np.random.default_rng(50 + dt)
    r, c, v, x, A = int_case(sp, rng, dt)
    want = dense_modular(A.nrows, r, c, v, x, BITS[dt])
    C = A.to_csr()
    assert np.array_equal(orc.spmv_csr(dt, C.rowptr, C.colind, C.values, x), want)
    assert np.array_equal(orc.spmv_coo(dt, A.nrows, A.rowind, A.colind, A.values, x), want)
    for R, Cb in ((4, 4), (2, 3)):
        B = C.to_bcsr(R, Cb)
        assert np.array_equal(orc.spmv_bcsr(dt, A.nrows, A.ncols, R, Cb, B.browptr, B.bcolind, B.bval, x), want)
        O = B.to_bcoo()
        assert np.array_equal(orc.spmv_bcoo(dt, A.nrows, A.ncols, R, Cb, O.browind, O.bcolind, O.bval, x), want)


def test_planted_transpose_fails_dense_check(sp):
    # a plausible mistake (A^T x instead of A x) must not pass the modular check
    rng = np.random.default_rng(60)
    r, c, v, x, A = int_case(sp, rng, 2, nrows=29, ncols=29, nnz=300)
    want = dense_modular(A.nrows, r, c, v, x, 32)
    wrong = dense_modular(A.ncols, c, r, v, x, 32)
    assert not np.array_equal(wrong, want)
    C = A.to_csr()
    assert not np.array_equal(orc.spmv_csr(2, C.rowptr, C.colind, C.values, x), wrong)


def test_integer_partial_merge_is_modular(sp):
    rng = np.random.default_rng(61)
    parts = rng.integers(-2 ** 31, 2 ** 31 - 1, (5, 100)).astype(np.int32)
    want = wrap(parts.astype(object).sum(axis=0).tolist(), 32)
    assert np.array_equal(orc.merge_parts(2, parts.copy()), want)


@pytest.mark.parametrize("dt", [4, 5])
def test_identity_and_diagonal_exact(sp, dt):
    n = 300
    rng = np.random.default_rng(62)
    x = rng.standard_normal(n).astype(orc.NP[dt])
    d = rng.standard_normal(n).astype(orc.NP[dt])
    idx = np.arange(n, dtype=np.uint32)
    for vals, want in ((np.ones(n, orc.NP[dt]), x), (d, d * x)):
        A = sp.Coo.from_arrays(n, n, idx, idx, vals)
        C = A.to_csr()
        assert np.array_equal(orc.spmv_csr(dt, C.rowptr, C.colind, C.values, x), want)
        assert np.array_equal(orc.spmv_coo(dt, n, A.rowind, A.colind, A.values, x), want)


@pytest.mark.parametrize("dt", [4, 5])
def test_laplacian_row_sums_closed_form(sp, dt):
    g = 40
    A = sp.gen_stencil5(g, g, dt, offdiag=sp.dist(sp.DIST_CONST, -1.0, -1.0), diag=4.0, seed=1)
    C = A.to_csr()
    y = orc.spmv_csr(dt, C.rowptr, C.colind, C.values, np.ones(A.ncols, orc.NP[dt]))
    gx, gy = np.meshgrid(np.arange(g), np.arange(g))
    missing = (gx == 0).astype(int) + (gx == g - 1) + (gy == 0) + (gy == g - 1)
    assert np.array_equal(y, missing.ravel().astype(orc.NP[dt]))


@pytest.mark.parametrize("dt", [4, 5])
def test_float_vs_scipy_within_bound(sp, dt):
    scipy_sparse = pytest.importorskip("scipy.sparse")
    A = sp.gen_rmat(12, 16 * 4096, dtype=dt, seed=63)
    x = sp.gen_vector(dt, A.ncols, seed=64)
    C = A.to_csr()
    y = orc.spmv_csr(dt, C.rowptr, C.colind, C.values, x)
    M = scipy_sparse.csr_matrix((C.values.astype(np.float64), C.colind, C.rowptr), shape=(A.nrows, A.ncols))
    yr = M @ x.astype(np.float64)
    absb = abs(M) @ np.abs(x.astype(np.float64))
    lens = np.diff(C.rowptr.astype(np.int64))
    tol = orc.fp_tolerance(dt, lens, absb) * 2  # both sides round
    assert np.all(np.abs(y.astype(np.float64) - yr) <= tol)
    # the long-double reference agrees with exact-sum float64 products
    yl, b = orc.ref_csr(dt, C.rowptr, C.colind, C.values, x)
    for i in range(0, A.nrows, 97):
        s = slice(C.rowptr[i], C.rowptr[i + 1])
        prods = C.values[s].astype(np.float64) * x[C.colind[s]].astype(np.float64)
        assert abs(yl[i] - math.fsum(prods.tolist())) <= 4 * 2 ** -53 * (np.abs(prods).sum() + 1e-300)
        assert abs(b[i] - np.abs(prods).sum()) <= 1e-12 * (np.abs(prods).sum() + 1e-300)


def test_float32_order_is_sequential_per_row(sp):
    rng = np.random.default_rng(65)
    n, m, nnz = 20, 50, 400
    keys = np.sort(rng.choice(n * m, nnz, replace=False))
    r, c = (keys // m).astype(np.uint32), (keys % m).astype(np.uint32)
    v = (rng.standard_normal(nnz) * 10 ** rng.uniform(-3, 3, nnz)).astype(np.float32)
    x = (rng.standard_normal(m) * 10 ** rng.uniform(-3, 3, m)).astype(np.float32)
    C = sp.Coo.from_arrays(n, m, r, c, v).to_csr()
    y = orc.spmv_csr(4, C.rowptr, C.colind, C.values, x)
    for i in range(n):
        acc = np.float32(0)
        for k in range(C.rowptr[i], C.rowptr[i + 1]):
            acc = np.float32(acc + np.float32(C.values[k] * x[C.colind[k]]))
        assert y[i] == acc, i


def test_cpu_baseline_timers_time_the_restatements(sp, orc):
    """The per-format CPU-baseline timers run the same restatements the
    parity checks use: the y of a timed run equals the oracle's y."""
    import numpy as np
    A = sp.gen_rmat(12, 40000, dtype=2, seed=9)
    x = sp.gen_vector(2, A.ncols, seed=10)
    C = A.to_csr()
    want = orc.spmv_csr(2, C.rowptr, C.colind, C.values, x)
    ms, done, y = orc.time_coo(2, A.nrows, A.rowind, A.colind, A.values, x, iters=3, warmup=1)
    assert ms > 0 and done == 3 and np.array_equal(y, want)
    B = C.to_bcsr(4, 4)
    ms, done, y = orc.time_bcsr(2, A.nrows, A.ncols, 4, 4, B.browptr, B.bcolind, B.bval, x, iters=3, warmup=1)
    assert ms > 0 and done == 3 and np.array_equal(y, want)
    O = B.to_bcoo()
    ms, done, y = orc.time_bcoo(2, A.nrows, A.ncols, 4, 4, O.browind, O.bcolind, O.bval, x, iters=3, warmup=1)
    assert ms > 0 and done == 3 and np.array_equal(y, want)
    for tiling in (0, 1, 2):
        for fmt in (0, 1, 2, 3):
            T = sp.coo2tiled(A, tiling, fmt, 4, 4)
            ms, done, y = orc.time_tiled(2, T, A.nrows, A.ncols, x, iters=2, warmup=0)
            assert ms > 0 and done == 2 and np.array_equal(y, want), (tiling, fmt)
    ms, done = orc.time_merge(4, 4, 1000, 2, iters=3, warmup=0)
    assert ms > 0 and done == 3
