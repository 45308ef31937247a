"""Conjugate gradient through one plan (sp_cg) on SPD systems: the 2D
Laplacian (5-point stencil, diagonal 4, off-diagonals -1) and a 3D 27-point
stencil made diagonally dominant.  The solution is checked against the known
x_true of b = A x_true, computed on the CPU in float64."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _laplacian(sp, g, dt):
    return sp.gen_stencil5(g, g, dt, offdiag=sp.dist(sp.DIST_CONST, -1.0, -1.0), diag=4.0, seed=1)


@pytest.mark.parametrize("dt,rtol,kernel", [(5, 1e-10, "CSR-nnz"), (5, 1e-10, "BCSR-block"), (4, 1e-5, "CSR-row"),
                                            (5, 1e-10, "RBDCSR")])
def test_cg_laplacian(sp, ctx, dt, rtol, kernel):
    A = _laplacian(sp, 64, dt)
    rng = np.random.default_rng(3)
    xt = rng.standard_normal(A.nrows)
    b = (A.dense() @ xt).astype(sp.NP_DTYPE[dt])
    plan = sp.Plan(ctx, kernel, A)
    bd = sp.DeviceArray.from_numpy(ctx, b)
    xd = sp.DeviceArray.from_numpy(ctx, np.zeros(A.nrows, sp.NP_DTYPE[dt]))
    it, res = plan.cg(bd.addr, xd.addr, maxit=2000, rtol=rtol)
    x = xd.numpy().astype(np.float64)
    assert res <= rtol and 0 < it < 2000, (it, res)
    # forward error bounded by cond(A) * rtol; the 64x64 Laplacian has cond ~ 1.7e3
    err = np.linalg.norm(x - xt) / np.linalg.norm(xt)
    assert err <= 3e3 * rtol * 10, err
    # the residual the solver reports matches a CPU recomputation
    r = b.astype(np.float64) - A.dense() @ x
    assert np.linalg.norm(r) / np.linalg.norm(b) <= 4 * rtol + (1e-6 if dt == 4 else 1e-14)


def test_cg_is_deterministic_and_restartable(sp, ctx):
    A = _laplacian(sp, 40, 5)
    b = sp.gen_vector(5, A.nrows, seed=9)
    plan = sp.Plan(ctx, "CSR-nnz", A)
    bd = sp.DeviceArray.from_numpy(ctx, b)
    out = []
    for _ in range(2):
        xd = sp.DeviceArray.from_numpy(ctx, np.zeros(A.nrows))
        it, res = plan.cg(bd.addr, xd.addr, maxit=50, rtol=0.0)
        assert it == 50
        out.append(xd.numpy())
    np.testing.assert_array_equal(out[0], out[1])
    # continuing from the 50-iteration iterate converges
    it, res = plan.cg(bd.addr, xd.addr, maxit=2000, rtol=1e-10)
    assert res <= 1e-10


def test_cg_rejects_indefinite_and_integer(sp, ctx):
    A = sp.gen_stencil5(16, 16, 5, offdiag=sp.dist(sp.DIST_CONST, -1.0, -1.0), diag=-4.0, seed=1)
    plan = sp.Plan(ctx, "CSR-nnz", A)
    bd = sp.DeviceArray.from_numpy(ctx, np.ones(A.nrows))
    xd = sp.DeviceArray.from_numpy(ctx, np.zeros(A.nrows))
    with pytest.raises(sp._lib.SparsePError) as e:
        plan.cg(bd.addr, xd.addr, maxit=10, rtol=1e-8)
    assert e.value.rc == -7
    Ai = sp.gen_stencil5(16, 16, 2, seed=1)
    pi = sp.Plan(ctx, "CSR-nnz", Ai)
    with pytest.raises(sp._lib.SparsePError) as e:
        pi.cg(bd.addr, xd.addr)
    assert e.value.rc == -5
