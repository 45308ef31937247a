"""The allocation-free entry point sp_spmv_csr_noalloc against sp_spmv of a
CSR-nnz plan (the entry point it is a variant of) and the oracle, through
the C ABI: integers bit-exact, floats within the gamma(len+2)*sum|a x|
bound.  Two calls in a row share the static scratch; the second is smaller.
"""
import numpy as np
import pytest

from gpu_util import assert_parity, run

pytestmark = pytest.mark.gpu


def _noalloc(sp, ctx, C, x, dt):
    rp = sp.DeviceArray.from_numpy(ctx, C.rowptr)
    ci = sp.DeviceArray.from_numpy(ctx, C.colind[:max(C.nnz, 1)] if C.nnz else np.zeros(1, np.uint32))
    va = sp.DeviceArray.from_numpy(ctx, C.values[:max(C.nnz, 1)] if C.nnz else np.zeros(1, sp.NP_DTYPE[dt]))
    xd = sp.DeviceArray.from_numpy(ctx, x)
    yd = sp.DeviceArray(ctx, (C.nrows,), sp.NP_DTYPE[dt])
    yd.fill_bytes(0x5A)  # poison: every row must be written
    rc = sp.spmv_csr_noalloc(ctx, dt, C.nrows, C.ncols, C.nnz, rp.addr, ci.addr, va.addr, xd.addr, yd.addr)
    assert rc == 0, f"sp_spmv_csr_noalloc rc={rc}"
    ctx.sync()
    return yd.numpy()


def _heavy_row(sp, dt):
    # one row of 9000 entries (five tiles), empty rows around it, short rows after
    rows = [3] * 9000 + [7, 7, 9] + list(range(10, 400))
    cols = list(range(9000)) + [1, 5, 2] + list(range(10, 400))
    vals = np.random.default_rng(3).integers(-4, 5, size=len(rows)).astype(sp.NP_DTYPE[dt])
    return sp.Coo.from_arrays(400, 9000, np.asarray(rows), np.asarray(cols), vals, sort=True)


@pytest.mark.parametrize("dt", [0, 1, 2, 3, 4, 5])
def test_noalloc_matches_plan_twice(sp, ctx, dt):
    big = sp.gen_rmat(15, 400000, dtype=dt, seed=21)
    small = _heavy_row(sp, dt)
    for A in (big, small):  # the second call is smaller than the first
        C = A.to_csr()
        x = sp.gen_vector(dt, A.ncols, seed=22)
        y = _noalloc(sp, ctx, C, x, dt)
        assert_parity(dt, y, A, x, f"noalloc dt{dt} n{A.nrows}")
        yp, _ = run(sp, ctx, "CSR-nnz", A, x, sp.opts(algo=2, items_per_lane=8))
        if dt <= 3:
            assert np.array_equal(y, yp), f"noalloc dt{dt} n{A.nrows}: differs from the CSR-nnz plan"


def test_noalloc_limits(sp, ctx):
    # beyond the fixed scratch: rejected before anything is launched
    big = sp.NOALLOC_MAX_ITEMS
    rc = sp.spmv_csr_noalloc(ctx, 4, 1 << 20, 1 << 20, big, 16, 16, 16, 16, 16)
    assert rc == sp.ERR_RANGE
    # no rows: nothing to do; bad dtype: argument error
    assert sp.spmv_csr_noalloc(ctx, 4, 0, 0, 0, 0, 0, 0, 0, 0) == 0
    assert sp.spmv_csr_noalloc(ctx, 9, 4, 4, 4, 16, 16, 16, 16, 16) == -1
    # no entries at all: every row is zero
    A = sp.Coo.from_arrays(1000, 1000, np.zeros(0, np.int64), np.zeros(0, np.int64), np.zeros(0, np.float32))
    C = A.to_csr()
    y = _noalloc(sp, ctx, C, np.ones(1000, np.float32), 4)
    assert np.array_equal(y, np.zeros(1000, np.float32))
