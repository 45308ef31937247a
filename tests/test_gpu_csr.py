"""GPU parity of the CSR variants (CSR-row, CSR-nnz) against the oracle.

Integer dtypes must be bit-exact with the reference spmv_host
(spmv/1D/CSR-row/host/app.c:66-77); float dtypes must lie within the
gamma(len+2)*sum|a_ij x_j| rounding bound of a long-double reference.
"""
import numpy as np
import pytest

from gpu_util import assert_parity, assert_parity_multi, run

pytestmark = pytest.mark.gpu

DTYPES = [0, 1, 2, 3, 4, 5]
ALGOS = [("vec1", dict(algo=1, lanes_per_row=1)), ("vec4", dict(algo=1, lanes_per_row=4)),
         ("vec16", dict(algo=1, lanes_per_row=16)), ("vec64", dict(algo=1, lanes_per_row=64)),
         ("seg4", dict(algo=2, items_per_lane=4)), ("seg8", dict(algo=2, items_per_lane=8)),
         ("seg16", dict(algo=2, items_per_lane=16)), ("sorted4", dict(algo=3, items_per_lane=4)),
         ("sorted16", dict(algo=3, items_per_lane=16)), ("hot", dict(algo=4)),
         ("chunk4", dict(algo=1, lanes_per_row=4, items_per_lane=4)),
         ("chunk32", dict(algo=1, lanes_per_row=32, items_per_lane=4))]


def test_stencil5_int32_bit_exact(sp, ctx):
    """The 256x256 5-point stencil in INT32, every CSR algorithm."""
    A = sp.gen_stencil5(256, 256, sp.I32, seed=1)
    assert A.nnz == 326656
    x = sp.gen_vector(sp.I32, A.ncols, sp.dist(0, -4, 4), seed=11)
    for kern in ("CSR-row", "CSR-nnz"):
        for name, kw in ALGOS:
            y, _ = run(sp, ctx, kern, A, x, sp.opts(**kw))
            assert_parity(2, y, A, x, f"{kern}/{name}")


@pytest.mark.parametrize("dt", DTYPES)
def test_rmat_all_dtypes(sp, ctx, dt):
    A = sp.gen_rmat(12, 60000, dtype=dt, seed=5)
    x = sp.gen_vector(dt, A.ncols, seed=6)
    for kern in ("CSR-row", "CSR-nnz"):
        for name, kw in ALGOS:
            y, _ = run(sp, ctx, kern, A, x, sp.opts(**kw))
            assert_parity(dt, y, A, x, f"{kern}/{name}/dt{dt}")


@pytest.mark.parametrize("dt", [0, 2, 4, 5])
def test_uniform_and_banded(sp, ctx, dt):
    for A in (sp.gen_uniform(5000, 7000, 0, 40, dtype=dt, seed=8), sp.gen_banded(6000, 32, 2, 64, dtype=dt, seed=9)):
        x = sp.gen_vector(dt, A.ncols, seed=10)
        for name, kw in ALGOS:
            y, _ = run(sp, ctx, "CSR-nnz", A, x, sp.opts(**kw))
            assert_parity(dt, y, A, x, name)


def _coo(sp, nrows, ncols, rows, cols, dt, seed=3):
    rng = np.random.default_rng(seed)
    vals = rng.integers(-4, 5, size=len(rows)).astype(sp.NP_DTYPE[dt])
    return sp.Coo.from_arrays(nrows, ncols, np.asarray(rows), np.asarray(cols), vals, sort=True)


@pytest.mark.parametrize("dt", [1, 2, 4])
def test_edge_cases(sp, ctx, dt):
    cases = []
    # no entries at all
    cases.append(_coo(sp, 100, 50, [], [], dt))
    # one long row spanning many tiles, surrounded by empty rows
    n = 9000
    cases.append(_coo(sp, 10, n, [5] * n, list(range(n)), dt))
    # long first and last row, empty rows between, trailing empty rows
    r = [0] * 5000 + [7] * 3 + [2999] * 6000
    c = list(range(5000)) + [1, 2, 3] + list(range(6000))
    cases.append(_coo(sp, 3500, 6000, r, c, dt))
    # single row / single column
    cases.append(_coo(sp, 1, 1, [0], [0], dt))
    cases.append(_coo(sp, 3000, 1, list(range(0, 3000, 2)), [0] * 1500, dt))
    for i, A in enumerate(cases):
        x = np.random.default_rng(i).integers(-4, 5, size=A.ncols).astype(A.values.dtype)
        for kern in ("CSR-row", "CSR-nnz"):
            for name, kw in ALGOS:
                y, _ = run(sp, ctx, kern, A, x, sp.opts(**kw))
                assert_parity(dt, y, A, x, f"case{i}/{kern}/{name}")


@pytest.mark.parametrize("dt", [2, 4, 5])
@pytest.mark.parametrize("k", [2, 3, 4, 8, 11, 16, 32, 45])
def test_spmm(sp, ctx, dt, k):
    A = sp.gen_rmat(11, 20000, dtype=dt, seed=12)
    X = np.stack([sp.gen_vector(dt, A.ncols, seed=20 + q) for q in range(k)], axis=1)
    for name, kw in [("vec4", dict(algo=1, lanes_per_row=4)), ("seg4", dict(algo=2, items_per_lane=4)),
                     ("seg8", dict(algo=2, items_per_lane=8)), ("sorted8", dict(algo=3, items_per_lane=8)),
                     ("hot", dict(algo=4))]:
        Y, _ = run(sp, ctx, "CSR-nnz", A, X, sp.opts(**kw), k=k)
        assert_parity_multi(dt, Y, A, X, name)


def test_repeatable(sp, ctx):
    """Deterministic: two runs give identical bits (float too)."""
    A = sp.gen_rmat(14, 200000, dtype=sp.F32, seed=13)
    x = sp.gen_vector(sp.F32, A.ncols, seed=14)
    y1, plan = run(sp, ctx, "CSR-nnz", A, x, sp.opts(algo=2))
    y2, _ = run(sp, ctx, "CSR-nnz", A, x, plan=plan)
    assert np.array_equal(y1.view(np.uint32), y2.view(np.uint32))


@pytest.mark.parametrize("dt", [2, 4])
def test_rmat_full_size(sp, ctx, dt):
    """R-MAT 2^20 x 2^20, 2^24 unique entries (the bench workload)."""
    A = sp.gen_rmat(20, 1 << 24, dtype=dt, seed=2)
    x = sp.gen_vector(dt, A.ncols, seed=22)
    y, plan = run(sp, ctx, "CSR-nnz", A, x, sp.opts(algo=2, items_per_lane=8))
    assert_parity(dt, y, A, x, "rmat-full")


@pytest.mark.parametrize("dt", [4, 5])
@pytest.mark.parametrize("ipt", [4, 8, 16])
def test_sorted_tiles_same_bits(sp, ctx, dt, ipt):
    """Column-ordered tiles put the products back in row order before the
    segmented sums, so every float result is bit-identical to the row-ordered
    merge kernel with the same tile size (one and several right-hand sides)."""
    A = sp.gen_rmat(16, 900000, dtype=dt, seed=31)
    x = sp.gen_vector(dt, A.ncols, seed=32)
    y0, _ = run(sp, ctx, "CSR-nnz", A, x, sp.opts(algo=2, items_per_lane=ipt))
    y1, _ = run(sp, ctx, "CSR-nnz", A, x, sp.opts(algo=3, items_per_lane=ipt))
    assert_parity(dt, y1, A, x, "sorted")
    assert np.array_equal(y0, y1)
    X = np.stack([sp.gen_vector(dt, A.ncols, seed=40 + q) for q in range(5)], axis=1)
    Y0, _ = run(sp, ctx, "CSR-nnz", A, X, sp.opts(algo=2, items_per_lane=ipt), k=5)
    Y1, _ = run(sp, ctx, "CSR-nnz", A, X, sp.opts(algo=3, items_per_lane=ipt), k=5)
    assert np.array_equal(Y0, Y1)


def test_sorted_tiles_column_limit(sp, ctx):
    """col << log2(tile) must fit 32 bits: 2^20 columns take 16 items per lane,
    one more column is refused; 8 items per lane take up to 2^21."""
    for ncols, ipt, ok in ((1 << 20, 16, True), ((1 << 20) + 1, 16, False), ((1 << 20) + 1, 8, True)):
        A = _coo(sp, 4, ncols, [0, 1, 3, 3], [0, ncols - 1, 5, ncols - 1], sp.I32)
        x = (np.arange(ncols) % 7 - 3).astype(np.int32)
        if ok:
            y, _ = run(sp, ctx, "CSR-nnz", A, x, sp.opts(algo=3, items_per_lane=ipt))
            assert_parity(sp.I32, y, A, x, f"ncols {ncols}")
        else:
            with pytest.raises(sp._lib.SparsePError):
                sp.Plan(ctx, "CSR-nnz", A, sp.opts(algo=3, items_per_lane=ipt))


def test_rmat_full_size_sorted(sp, ctx):
    """The bench matrix with column-ordered 16-item tiles (2^20 columns: the largest that fits)."""
    A = sp.gen_rmat(20, 1 << 24, dtype=sp.F32, seed=2)
    x = sp.gen_vector(sp.F32, A.ncols, seed=22)
    y, _ = run(sp, ctx, "CSR-nnz", A, x, sp.opts(algo=3, items_per_lane=16))
    assert_parity(sp.F32, y, A, x, "rmat-full-sorted")


@pytest.mark.parametrize("dt", [0, 2, 3, 4, 5])
def test_hot_staging(sp, ctx, dt):
    """Hot-x tiles: more frequent columns than LDS slots (the excess stays in
    global memory), every column hot (a narrow matrix), and no column used
    twice (nothing staged)."""
    cases = [sp.gen_rmat(17, 1500000, dtype=dt, seed=41),          # 128K columns, power law
             sp.gen_uniform(40000, 3000, 10, 60, dtype=dt, seed=42),  # 3000 columns, all hot
             _coo(sp, 5000, 5000, list(range(5000)), list(range(5000))[::-1], dt)]  # permutation
    for i, A in enumerate(cases):
        x = sp.gen_vector(dt, A.ncols, seed=43 + i)
        y, plan = run(sp, ctx, "CSR-nnz", A, x, sp.opts(algo=4))
        assert plan.info()["algo"] == 4
        assert_parity(dt, y, A, x, f"hot case{i}")
        y2, _ = run(sp, ctx, "CSR-nnz", A, x, plan=plan)  # the staged buffer is rewritten per call
        assert np.array_equal(y, y2)


def test_hot_column_limit(sp, ctx):
    A = _coo(sp, 4, (1 << 20) + 1, [0, 1, 3, 3], [0, 1 << 20, 5, 1 << 20], sp.I32)
    with pytest.raises(sp._lib.SparsePError):
        sp.Plan(ctx, "CSR-nnz", A, sp.opts(algo=4))


def test_rmat_full_size_hot(sp, ctx):
    """The bench matrix with hot-x tiles (28,672 staged FP32 columns)."""
    A = sp.gen_rmat(20, 1 << 24, dtype=sp.F32, seed=2)
    x = sp.gen_vector(sp.F32, A.ncols, seed=22)
    y, _ = run(sp, ctx, "CSR-nnz", A, x, sp.opts(algo=4))
    assert_parity(sp.F32, y, A, x, "rmat-full-hot")


def test_mem_cap_reports_nomem(sp, ctx):
    """sp_opts.mem_cap_bytes, the GPU counterpart of the reference's MRAM capacity
    assert (spmv/1D/CSR-row/host/app.c:214-215): a plan that needs more device
    memory than the cap fails with SP_ERR_NOMEM instead of aborting; within the
    cap it builds, and its device bytes stay under the cap."""
    A = sp.gen_rmat(12, 40000, dtype=4, seed=5)
    for kern in ("CSR-nnz", "COO-nnz", "BCSR-block", "RBDCSR"):
        with pytest.raises(sp._lib.SparsePError) as e:
            sp.Plan(ctx, kern, A, sp.opts(mem_cap_bytes=4096))
        assert e.value.rc == -2, kern
        cap = 64 << 20
        p = sp.Plan(ctx, kern, A, sp.opts(mem_cap_bytes=cap))
        assert 0 < p.info()["device_bytes"] <= cap, kern


def test_bandwidth_probes(sp, ctx):
    """sp_time_copy and sp_time_read time every pass and give rates below the
    published 8 TB/s of HBM; bad arguments are refused."""
    import ctypes
    n = 256 << 20
    for name, fn, moved in (("copy", ctx.time_copy, 2 * n), ("read", ctx.time_read, n)):
        ms = fn(n, 4)
        assert len(ms) == 4 and all(t > 0 for t in ms), name
        assert moved / (min(ms) * 1e-3) < 8.5e12, name
    out = (ctypes.c_float * 1)()
    assert sp.lib().sp_time_read(ctx.ptr, 8, 1, out) == -1  # SP_ERR_ARG
    assert sp.lib().sp_time_read(ctx.ptr, n, 0, out) == -1


@pytest.mark.parametrize("dt", DTYPES)
def test_quad_rows(sp, ctx, dt):
    """Rows made of aligned runs of four columns (a 27-point stencil with 4
    unknowns per point; a matrix of random aligned quads with empty rows and
    rows of 4 to 64 entries): the chunked row kernel reads x four values at a
    time.  Every lane count, both kernel ids; x at an address that is not a
    multiple of four values takes the gathering path.  Integers bit-exact."""
    rng = np.random.default_rng(40 + dt)
    nrows, ncols = 3000, 2400
    rows, cols = [], []
    for r in range(nrows):
        if r % 11 == 0:
            continue
        q = rng.choice(ncols // 4, size=int(rng.integers(1, 17)), replace=False)
        for c0 in np.sort(q):
            rows += [r] * 4
            cols += [4 * int(c0) + i for i in range(4)]
    mats = [sp.gen_stencil27(5, 4, True, dt, seed=41), _coo(sp, nrows, ncols, rows, cols, dt, seed=42)]
    for A in mats:
        x = sp.gen_vector(dt, A.ncols, seed=43)
        for kern in ("CSR-row", "CSR-nnz"):
            for L in (4, 8, 16, 32, 64):
                y, _ = run(sp, ctx, kern, A, x, sp.opts(algo=1, lanes_per_row=L, items_per_lane=4))
                assert_parity(dt, y, A, x, f"{kern}/quads/L{L}/n{A.nrows}")
        # x one value past an aligned address
        plan = sp.Plan(ctx, "CSR-nnz", A, sp.opts(algo=1, lanes_per_row=16, items_per_lane=4))
        xe = np.concatenate([np.zeros(1, x.dtype), x])
        xd = sp.DeviceArray.from_numpy(ctx, xe)
        yd = sp.DeviceArray(ctx, (A.nrows,), x.dtype)
        plan.spmv(xd.addr + x.itemsize, yd.addr)
        ctx.sync()
        assert_parity(dt, yd.numpy(), A, x, f"quads/x offset/n{A.nrows}")
