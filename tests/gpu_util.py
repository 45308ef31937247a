"""Helpers for the GPU parity tests: device buffers through torch, one SpMV
through the C ABI, and the parity check against the oracle."""
import numpy as np

import oracle_lib as orc

TORCH_DT = None


def torch_mod():
    import torch
    return torch


def dev(a):
    torch = torch_mod()
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def run(sp, ctx, kernel, A, x, o=None, k=1, plan=None):
    """y = A x (or Y = A X for k > 1, X of shape (ncols, k)) through sp_spmv/sp_spmm."""
    torch = torch_mod()
    plan = plan or sp.Plan(ctx, kernel, A, o)
    xd = dev(x)
    if k == 1:
        yd = torch.full((A.nrows,), 0x5A, dtype=xd.dtype, device="cuda")
    else:
        yd = torch.full((A.nrows, k), 0x5A, dtype=xd.dtype, device="cuda")
    torch.cuda.synchronize()
    if k == 1:
        plan.spmv(xd.data_ptr(), yd.data_ptr())
    else:
        plan.spmm(k, xd.data_ptr(), k, yd.data_ptr(), k)
    ctx.sync()
    return yd.cpu().numpy(), plan


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
