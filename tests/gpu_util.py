"""Helpers for the GPU parity tests: device buffers through the C ABI
(sp_dmalloc / sp_memcpy_*), one SpMV per call, and the parity check against
the oracle.  torch is deliberately not imported here: it bundles its own HIP
runtime, and one process must use a single runtime."""
import numpy as np

import oracle_lib as orc


def run(sp, ctx, kernel, A, x, o=None, k=1, plan=None):
    """y = A x (or Y = A X for k > 1, X of shape (ncols, k)) through sp_spmv/sp_spmm."""
    plan = plan or sp.Plan(ctx, kernel, A, o)
    xd = sp.DeviceArray.from_numpy(ctx, x)
    shape = (A.nrows,) if k == 1 else (A.nrows, k)
    yd = sp.DeviceArray(ctx, shape, x.dtype)
    yd.fill_bytes(0x5A)  # poison: every row must be written
    if k == 1:
        plan.spmv(xd.addr, yd.addr)
    else:
        plan.spmm(k, xd.addr, k, yd.addr, k)
    ctx.sync()
    return yd.numpy(), plan


def oracle_y(dt, A, x):
    """The reference spmv_host result (CSR order) for COO A."""
    C = A.to_csr()
    return orc.spmv_csr(dt, C.rowptr, C.colind, C.values, x), C


def assert_parity(dt, y, A, x, what=""):
    """Integers: bit-exact against the oracle.  Floats: within the
    gamma(len+2)*sum|a x| bound of the long-double reference."""
    C = A.to_csr()
    if dt <= 3:
        yo = orc.spmv_csr(dt, C.rowptr, C.colind, C.values, x)
        bad = np.nonzero(y != yo)[0]
        assert bad.size == 0, f"{what}: {bad.size} rows differ, first {bad[:5]} got {y[bad[:5]]} want {yo[bad[:5]]}"
    else:
        yr, b = orc.ref_csr(dt, C.rowptr, C.colind, C.values, x)
        lens = np.diff(C.rowptr.astype(np.int64))
        tol = orc.fp_tolerance(dt, lens, b)
        err = np.abs(y.astype(np.float64) - yr)
        bad = np.nonzero(~(err <= tol))[0]
        assert bad.size == 0, (f"{what}: {bad.size} rows outside tolerance, first {bad[:5]} "
                               f"err {err[bad[:5]]} tol {tol[bad[:5]]}")


def assert_parity_multi(dt, Y, A, X, what=""):
    for q in range(X.shape[1]):
        assert_parity(dt, np.ascontiguousarray(Y[:, q]), A, np.ascontiguousarray(X[:, q]), f"{what} rhs {q}")
