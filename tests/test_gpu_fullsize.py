"""GPU parity at the configurations' full sizes, through the C ABI.

* R-MAT 2^20 x 2^20 with 16,777,216 unique entries (a, b, c = 0.57, 0.19,
  0.19; seed 2): every 1D kernel id (CSR, COO, BCSR 4x4, BCOO 4x4, row /
  entry / block balanced) in all six dtypes at its default configuration.
* The 64^3 x 4 DOF 27-point stencil, periodic (113,246,208 entries, seed 3):
  all 25 kernel ids in FP32 and FP64.
* The 256-problem batch (rows log-uniform in [1k, 200k], 2..64 entries per
  row, uniform / banded / R-MAT, 1, 4, 8 or 32 right-hand sides; seed 4) in
  FP32 and INT16 through one hipGraph replay of sp_batch_run.

Integers are bit-exact against the oracle's spmv_host (modular sums are
order- and partition-free); floats lie within gamma(len+2)*sum|a x| of the
long-double reference.  Plans are built on host threads (plan creation is
thread-safe), and the oracle runs on host threads too.
"""
import math
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import oracle_lib as orc

pytestmark = pytest.mark.gpu

POOL = 8


class Want:
    """The oracle's answer for one CSR matrix and x, computed once."""

    def __init__(self, dt, C, x):
        self.dt = dt
        if dt <= 3:
            self.y = orc.spmv_csr(dt, C.rowptr, C.colind, C.values, x)
        else:
            self.y, b = orc.ref_csr(dt, C.rowptr, C.colind, C.values, x)
            self.tol = orc.fp_tolerance(dt, np.diff(C.rowptr.astype(np.int64)), b)

    def check(self, y, what):
        if self.dt <= 3:
            bad = np.nonzero(y != self.y)[0]
            assert bad.size == 0, f"{what}: {bad.size} rows differ, first {bad[:5]}"
        else:
            err = np.abs(y.astype(np.float64) - self.y)
            bad = np.nonzero(~(err <= self.tol))[0]
            assert bad.size == 0, f"{what}: {bad.size} rows outside tolerance, first {bad[:5]} err {err[bad[:5]]}"


def _build(sp, ctx, A, names):
    with ThreadPoolExecutor(POOL) as ex:
        return list(ex.map(lambda k: sp.Plan(ctx, k, A, sp.opts()), names))


def _run_all(sp, ctx, A, x, want, names, tag):
    xd = sp.DeviceArray.from_numpy(ctx, x)
    yd = sp.DeviceArray(ctx, (A.nrows,), x.dtype)
    for name, plan in zip(names, _build(sp, ctx, A, names)):
        yd.fill_bytes(0x5A)  # poison: every row must be written
        plan.spmv(xd.addr, yd.addr)
        ctx.sync()
        print(f"{tag} {name}", flush=True)
        want.check(yd.numpy(), f"{tag} {name}")
        del plan


ONE_D = ["CSR-row", "CSR-nnz", "COO-row", "COO-nnz-rgrn", "COO-nnz", "BCSR-block", "BCSR-nnz", "BCOO-block",
         "BCOO-nnz"]


@pytest.mark.timeout(240)
@pytest.mark.parametrize("dt", [0, 1, 2, 3, 4, 5])
def test_rmat20_every_1d_kernel(sp, ctx, dt):
    vals = sp.dist(sp.DIST_UNIFORM, -1, 1) if dt >= 4 else sp.dist(sp.DIST_UNIFORM, -4, 4)
    A = sp.gen_rmat(20, 1 << 24, 0.57, 0.19, 0.19, dtype=dt, vals=vals, seed=2)
    assert A.nrows == 1 << 20 and A.nnz == 1 << 24
    x = sp.gen_vector(dt, A.ncols, d=vals, seed=22)
    want = Want(dt, A.to_csr(), x)
    _run_all(sp, ctx, A, x, want, ONE_D, f"R-MAT 2^20 dt{dt}")


@pytest.fixture(scope="module")
def stencil27(sp):
    cache = {}

    def get(dt):
        if dt not in cache:
            cache.clear()
            A = sp.gen_stencil27(64, 4, True, dtype=dt, seed=3)
            x = sp.gen_vector(dt, A.ncols, d=sp.dist(sp.DIST_NORMAL, 0, 1), seed=4)
            cache[dt] = (A, x, Want(dt, A.to_csr(), x))
        return cache[dt]
    return get


@pytest.mark.timeout(240)
@pytest.mark.parametrize("dt", [4, 5])
@pytest.mark.parametrize("group", ["1D", "2D"])
def test_stencil27_all_25_kernels(sp, ctx, stencil27, dt, group):
    A, x, want = stencil27(dt)
    assert A.nrows == 1 << 20 and A.nnz == 113246208
    names = sp.KERNELS[:9] if group == "1D" else sp.KERNELS[9:]
    _run_all(sp, ctx, A, x, want, names, f"stencil27 64^3x4 dt{dt}")


def _batch_specs(n=256, seed=4):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        rows = int(math.exp(rng.uniform(math.log(1000), math.log(200000))))
        out.append((i % 3, rows, [1, 4, 8, 32][i % 4], 4000 + i))
    return out


@pytest.mark.timeout(280)
@pytest.mark.parametrize("dt", [4, 1])
def test_batch256_one_graph(sp, ctx, dt):
    specs = _batch_specs()

    def gen(spec):
        kind, rows, k, seed = spec
        if kind == 0:
            return sp.gen_uniform(rows, rows, 2, 64, dtype=dt, seed=seed)
        if kind == 1:
            return sp.gen_banded(rows, 32, 2, 64, dtype=dt, seed=seed)
        scale = max(10, int(round(math.log2(rows))))
        return sp.gen_rmat(scale, (1 << scale) * 33, dtype=dt, seed=seed)

    with ThreadPoolExecutor(16) as ex:
        mats = list(ex.map(gen, specs))
    ks = [s[2] for s in specs]
    print(f"batch dt{dt}: {sum(A.nnz for A in mats)} entries generated", flush=True)
    with ThreadPoolExecutor(POOL) as ex:
        plans = list(ex.map(lambda A: sp.Plan(ctx, "CSR-nnz", A, sp.opts()), mats))
    Xs, Xd, Yd = [], [], []
    for i, (A, k) in enumerate(zip(mats, ks)):
        rng = np.random.default_rng(i)
        X = (rng.standard_normal((A.ncols, k)).astype(np.float32) if dt == 4
             else rng.integers(-4, 5, (A.ncols, k)).astype(np.int16))
        Xs.append(X)
        Xd.append(sp.DeviceArray.from_numpy(ctx, X))
        y = sp.DeviceArray(ctx, (A.nrows, k), X.dtype)
        y.fill_bytes(0x5A)
        Yd.append(y)
    b = sp.Batch(ctx, plans, ks, graph=True)
    b.run([x.addr for x in Xd], [y.addr for y in Yd])
    ctx.sync()
    print(f"batch dt{dt}: {b.info()['groups']} groups ran", flush=True)

    def check(i):
        A, X, Y = mats[i], Xs[i], Yd[i].numpy()
        C = A.to_csr()
        for q in range(X.shape[1]):
            Want(dt, C, np.ascontiguousarray(X[:, q])).check(np.ascontiguousarray(Y[:, q]), f"problem {i} rhs {q}")

    with ThreadPoolExecutor(16) as ex:
        list(ex.map(check, range(len(mats))))
