"""GPU parity of the nine 1D variants (CSR-row .. BCOO-nnz) against the
oracle, over dtypes, matrix families, the reference partition knobs
(parts = ex NR_DPUS, sub-parts = ex NR_TASKLETS, tasklet balance) and block
sizes.  Integers bit-exact; floats within the gamma(len+2)*sum|a x| bound.
"""
import numpy as np
import pytest

from gpu_util import assert_parity, assert_parity_multi, run

pytestmark = pytest.mark.gpu

K1D = ["CSR-row", "CSR-nnz", "COO-row", "COO-nnz-rgrn", "COO-nnz", "BCSR-block", "BCSR-nnz", "BCOO-block",
       "BCOO-nnz"]


def matrices(sp, dt):
    return {
        "stencil5": sp.gen_stencil5(48, 40, dt, sp.dist(0, -8, 8), 4.0, seed=1),
        "rmat": sp.gen_rmat(12, 50000, dtype=dt, seed=5),
        "uniform": sp.gen_uniform(3000, 4100, 0, 30, dtype=dt, seed=8),
        "stencil27": sp.gen_stencil27(6, 4, True, dt, seed=3),
        "banded": sp.gen_banded(2500, 32, 2, 64, dtype=dt, seed=9),
    }


OPTS = [
    ("default", {}),
    ("p16s4", dict(nparts=16, nsub=4)),
    ("p64s16-rowbal", dict(nparts=64, nsub=16, sub_balance=0)),
    ("p3s1", dict(nparts=3, nsub=1)),
]


@pytest.mark.parametrize("dt", [0, 1, 2, 3, 4, 5])
def test_all_1d_variants(sp, ctx, dt):
    for mname, A in matrices(sp, dt).items():
        x = sp.gen_vector(dt, A.ncols, seed=17)
        for kern in K1D:
            for oname, kw in OPTS:
                y, _ = run(sp, ctx, kern, A, x, sp.opts(**kw))
                assert_parity(dt, y, A, x, f"{mname}/{kern}/{oname}")


@pytest.mark.parametrize("rc", [(1, 1), (2, 2), (4, 4), (8, 8), (2, 4), (4, 2)])
@pytest.mark.parametrize("dt", [1, 2, 3, 4, 5])
def test_block_sizes(sp, ctx, rc, dt):
    """Default lanes pick the coalesced BCSR kernel where a block row holds whole
    16-byte chunks (e.g. FP64 2x2/4x4, FP32 4x4/8x8, INT16 8x8); explicit lane
    counts pick the register-blocked kernel."""
    r, c = rc
    for A in (sp.gen_rmat(11, 20000, dtype=dt, seed=12), sp.gen_stencil27(5, 4, False, dt, seed=4),
              sp.gen_uniform(1001, 999, 1, 20, dtype=dt, seed=2)):
        x = sp.gen_vector(dt, A.ncols, seed=3)
        for kern in ("BCSR-block", "BCSR-nnz", "BCOO-block", "BCOO-nnz"):
            for kw in ({}, dict(nparts=8, nsub=4), dict(lanes_per_row=1), dict(lanes_per_row=64)):
                if kern.startswith("BCOO") and "lanes_per_row" in kw:
                    continue
                y, _ = run(sp, ctx, kern, A, x, sp.opts(row_block_size=r, col_block_size=c, **kw))
                assert_parity(dt, y, A, x, f"{kern} {r}x{c} {kw}")


@pytest.mark.parametrize("rc", [(1, 1), (2, 2), (2, 4), (4, 2), (4, 4), (8, 8)])
@pytest.mark.parametrize("k", [16, 32, 37])
@pytest.mark.parametrize("dt", [3, 4, 5])
def test_block_spmm_wide(sp, ctx, rc, k, dt):
    """Right-hand-side chunks of 16/32: register-blocked (4-byte values, R*K <= 64),
    lanes across the right-hand sides (8-byte values, R*C <= 16), chunks of 8 for
    8x8 blocks.  INT64 bit-exact."""
    r, c = rc
    A = sp.gen_stencil27(6, 2, False, dt, seed=5)
    X = np.stack([sp.gen_vector(dt, A.ncols, seed=60 + q) for q in range(k)], axis=1)
    for kw in ({}, dict(lanes_per_row=1), dict(lanes_per_row=64)):
        Y, _ = run(sp, ctx, "BCSR-nnz", A, X, sp.opts(row_block_size=r, col_block_size=c, **kw), k=k)
        assert_parity_multi(dt, Y, A, X, f"{r}x{c} {kw}")


@pytest.mark.parametrize("dt", [2, 4])
@pytest.mark.parametrize("k", [3, 8, 19, 32])
def test_spmm_all_kinds(sp, ctx, dt, k):
    A = sp.gen_rmat(11, 25000, dtype=dt, seed=21)
    X = np.stack([sp.gen_vector(dt, A.ncols, seed=40 + q) for q in range(k)], axis=1)
    for kern in K1D:
        Y, _ = run(sp, ctx, kern, A, X, sp.opts(), k=k)
        assert_parity_multi(dt, Y, A, X, kern)


@pytest.mark.parametrize("dt", [1, 4])
def test_edge_cases_all_kinds(sp, ctx, dt):
    rng = np.random.default_rng(0)
    cases = []
    n = 9000
    cases.append((10, n, [5] * n, list(range(n))))                       # one long row
    cases.append((100, 50, [], []))                                       # no entries
    r = [0] * 5000 + [7] * 3 + [2999] * 6000                              # long rows, gaps
    c = list(range(5000)) + [1, 2, 3] + list(range(6000))
    cases.append((3001, 6000, r, c))
    cases.append((1, 1, [0], [0]))
    cases.append((7, 13, [6, 6, 6], [0, 5, 12]))                          # rows not multiple of blocks
    for i, (nr, nc, rr, cc) in enumerate(cases):
        vals = rng.integers(-4, 5, size=len(rr)).astype(sp.NP_DTYPE[dt])
        A = sp.Coo.from_arrays(nr, nc, np.asarray(rr), np.asarray(cc), vals, sort=True)
        x = rng.integers(-4, 5, size=nc).astype(sp.NP_DTYPE[dt])
        for kern in K1D:
            for kw in ({}, dict(nparts=4, nsub=3)):
                y, _ = run(sp, ctx, kern, A, x, sp.opts(**kw))
                assert_parity(dt, y, A, x, f"case{i}/{kern}/{kw}")


def test_unsorted_coo_input(sp, ctx):
    """COO plans sort a copy when the input is not in (row, col) order."""
    rng = np.random.default_rng(3)
    A0 = sp.gen_rmat(10, 5000, dtype=sp.I32, seed=3)
    perm = rng.permutation(A0.nnz)
    A = sp.Coo.from_arrays(A0.nrows, A0.ncols, A0.rowind[perm], A0.colind[perm], A0.values[perm])
    x = sp.gen_vector(sp.I32, A.ncols, seed=4)
    for kern in K1D:
        y, _ = run(sp, ctx, kern, A, x)
        assert_parity(2, y, A, x, kern)


@pytest.mark.parametrize("dt", [2, 4, 5])
def test_coo_column_order_same_bits(sp, ctx, dt):
    """COO tiles in column order (the default where the columns fit) give the
    same bits as the row-ordered tiles (algo=2), over the partition knobs and
    for several right-hand sides."""
    A = sp.gen_rmat(14, 150000, dtype=dt, seed=51)
    x = sp.gen_vector(dt, A.ncols, seed=52)
    X = np.stack([sp.gen_vector(dt, A.ncols, seed=53 + q) for q in range(3)], axis=1)
    for kern in ("COO-row", "COO-nnz-rgrn", "COO-nnz"):
        for oname, kw in OPTS:
            for ipt in (4, 16):
                y0, _ = run(sp, ctx, kern, A, x, sp.opts(algo=2, items_per_lane=ipt, **kw))
                y1, plan = run(sp, ctx, kern, A, x, sp.opts(items_per_lane=ipt, **kw))
                assert plan.info()["algo"] == 3
                assert_parity(dt, y1, A, x, f"{kern}/{oname}/ipt{ipt}")
                assert np.array_equal(y0, y1), f"{kern}/{oname}/ipt{ipt}"
        Y0, _ = run(sp, ctx, kern, A, X, sp.opts(algo=2), k=3)
        Y1, _ = run(sp, ctx, kern, A, X, sp.opts(), k=3)
        assert np.array_equal(Y0, Y1)


def _coo_rows(sp, nrows, ncols, rows, cols, dt, seed=3):
    rng = np.random.default_rng(seed)
    vals = rng.integers(-4, 5, size=len(rows)).astype(sp.NP_DTYPE[dt])
    return sp.Coo.from_arrays(nrows, ncols, np.asarray(rows), np.asarray(cols), vals, sort=True)


@pytest.mark.parametrize("rc", [(1, 1), (2, 2), (4, 4), (8, 8), (2, 4), (4, 2)])
@pytest.mark.parametrize("dt", [0, 2, 3, 4, 5])
def test_bcsr_segmented(sp, ctx, rc, dt):
    """Entry-balanced BCSR (merge-path tiles over block-row ends and blocks,
    algo=2, and the automatic choice for skewed block rows): power-law and
    uniform matrices, a hub row spanning many tiles between empty block rows,
    trailing empty rows, and several right-hand sides."""
    r, c = rc
    n = 9000
    hub = _coo_rows(sp, 4000, n, [37] * n + [3999] * 5, list(range(n)) + [1, 2, 3, 4, 5], dt)
    mats = [sp.gen_rmat(12, 60000, dtype=dt, seed=71), sp.gen_uniform(2001, 1999, 0, 30, dtype=dt, seed=72), hub,
            _coo_rows(sp, 300, 50, [], [], dt)]
    for i, A in enumerate(mats):
        x = sp.gen_vector(dt, A.ncols, seed=73 + i)
        for kern in ("BCSR-block", "BCSR-nnz"):
            y, plan = run(sp, ctx, kern, A, x, sp.opts(algo=2, row_block_size=r, col_block_size=c))
            assert plan.info()["algo"] == 2
            assert_parity(dt, y, A, x, f"mat{i}/{kern}/{r}x{c}")
        if i == 0:
            X = np.stack([sp.gen_vector(dt, A.ncols, seed=80 + q) for q in range(3)], axis=1)
            Y, _ = run(sp, ctx, "BCSR-nnz", A, X, sp.opts(algo=2, row_block_size=r, col_block_size=c), k=3)
            assert_parity_multi(dt, Y, A, X, f"spmm {r}x{c}")
    # the hub matrix is skewed: the default picks the entry-balanced kernel
    _, plan = run(sp, ctx, "BCSR-nnz", hub, sp.gen_vector(dt, hub.ncols, seed=9), sp.opts(row_block_size=r, col_block_size=c))
    assert plan.info()["algo"] == 2


@pytest.mark.parametrize("dt", [0, 1, 2, 3, 4, 5])
def test_atomic_merges(sp, ctx, dt):
    """algo=5: COO and BCOO plans whose rows crossing tiles are added to y with
    device atomics (the reference's lock-based cg/fg merges) after a pre-zero
    of those rows.  y starts poisoned, so every row must be written.  Integers
    bit-exact (and equal to the carry-merged plans); floats within the bound.
    Covers a hub row over many tiles between empty rows, partition knobs,
    4096-entry COO tiles and several right-hand sides."""
    n = 9000
    hub = _coo_rows(sp, 4000, n, [37] * n + [3999] * 5, list(range(n)) + [1, 2, 3, 4, 5], dt)
    mats = [sp.gen_rmat(12, 60000, dtype=dt, seed=91), sp.gen_uniform(2001, 1999, 0, 30, dtype=dt, seed=92), hub,
            _coo_rows(sp, 300, 50, [], [], dt), _coo_rows(sp, 5, 5, [4], [2], dt)]
    for i, A in enumerate(mats):
        x = sp.gen_vector(dt, A.ncols, seed=93 + i)
        for kern in ("COO-row", "COO-nnz-rgrn", "COO-nnz"):
            for kw in ({}, dict(nparts=16, nsub=4), dict(items_per_lane=16), dict(items_per_lane=8, nparts=3)):
                y, plan = run(sp, ctx, kern, A, x, sp.opts(algo=5, **kw))
                assert plan.info()["algo"] == 5
                assert_parity(dt, y, A, x, f"mat{i}/{kern}/{kw}")
                if dt <= 3:
                    y0, _ = run(sp, ctx, kern, A, x, sp.opts(**kw))
                    assert np.array_equal(y, y0), f"mat{i}/{kern}/{kw}"
        for kern in ("BCOO-block", "BCOO-nnz"):
            for r, c in ((4, 4), (2, 2), (8, 8)):
                for kw in ({}, dict(nparts=8, nsub=4)):
                    y, plan = run(sp, ctx, kern, A, x, sp.opts(algo=5, row_block_size=r, col_block_size=c, **kw))
                    assert plan.info()["algo"] == 5
                    assert_parity(dt, y, A, x, f"mat{i}/{kern}/{r}x{c}/{kw}")
        if i == 0:
            X = np.stack([sp.gen_vector(dt, A.ncols, seed=99 + q) for q in range(3)], axis=1)
            for kern in ("COO-nnz", "BCOO-nnz"):
                Y, _ = run(sp, ctx, kern, A, X, sp.opts(algo=5), k=3)
                assert_parity_multi(dt, Y, A, X, f"spmm {kern}")


def test_atomic_merges_rejected(sp, ctx):
    """algo=5 exists for the COO and BCOO ids (1D and 2D); the BCSR ids take it
    as their default plan (no block row is shared between workgroups); the CSR
    ids refuse it; unknown algorithms are argument errors; 2D atomic plans do
    not stage x slices."""
    A = sp.gen_rmat(10, 5000, dtype=sp.F32, seed=3)
    for kern in ("CSR-row", "CSR-nnz", "DCSR", "RBDCSR", "BDCSR"):
        with pytest.raises(sp._lib.SparsePError) as e:
            sp.Plan(ctx, kern, A, sp.opts(algo=5))
        assert e.value.rc == -5, kern
    A2 = sp.gen_rmat(10, 5000, dtype=sp.I32, seed=3)
    x = sp.gen_vector(sp.I32, A2.ncols, seed=4)
    for kern in ("BCSR-block", "BCSR-nnz", "DBCSR", "RBDBCSR-block", "RBDBCSR-nnz", "BDBCSR-block", "BDBCSR-nnz"):
        y5, p5 = run(sp, ctx, kern, A2, x, sp.opts(algo=5))
        y0, p0 = run(sp, ctx, kern, A2, x, sp.opts())
        assert p5.info()["algo"] == p0.info()["algo"], kern
        assert np.array_equal(y5, y0), kern
        assert_parity(sp.I32, y5, A2, x, kern)
    with pytest.raises(sp._lib.SparsePError) as e:
        sp.Plan(ctx, "COO-nnz", A, sp.opts(algo=6))
    assert e.value.rc == -1
    with pytest.raises(sp._lib.SparsePError) as e:
        sp.Plan(ctx, "DCOO", A, sp.opts(algo=5, vert=4, lds_slice=4096))
    assert e.value.rc == -5


@pytest.mark.parametrize("dt", [2, 4])
def test_coo_strided_output(sp, ctx, dt):
    """COO SpMM into Y with a row stride wider than k: every row's first k
    columns are written (empty rows zeroed by the fix-up) and the columns
    past k keep what the caller had there."""
    A = sp.gen_rmat(12, 30000, dtype=dt, seed=61)
    k, ld = 3, 5
    X = np.stack([sp.gen_vector(dt, A.ncols, seed=62 + q) for q in range(k)], axis=1)
    for kern in ("COO-row", "COO-nnz"):
        plan = sp.Plan(ctx, kern, A)
        xd = sp.DeviceArray.from_numpy(ctx, np.ascontiguousarray(X))
        yd = sp.DeviceArray(ctx, (A.nrows, ld), X.dtype)
        yd.fill_bytes(0x5A)
        plan.spmm(k, xd.addr, k, yd.addr, ld)
        ctx.sync()
        Y = yd.numpy()
        assert_parity_multi(dt, np.ascontiguousarray(Y[:, :k]), A, X, kern)
        pad = np.ascontiguousarray(Y[:, k:]).view(np.uint8)
        assert np.array_equal(pad, np.full(pad.shape, 0x5A, np.uint8)), kern


@pytest.mark.parametrize("dt", [4, 5])
@pytest.mark.parametrize("k", [8, 16, 24, 32, 35, 40, 64])
def test_bcsr4_spmm_matrix_cores(sp, ctx, dt, k):
    """4x4 BCSR blocks times 8/16/32-column chunks (matrix cores; FP32
    32-column chunks on the vector ALUs): dense 4x4 blocks (4 unknowns per
    grid point), a truncated last block row and block column (rows and
    columns not multiples of 4), empty block rows, 1D plans of both
    balances, 2D tilings, every chunk size, and x / y rows of 35 values
    (not whole 16-byte pieces)."""
    cases = [sp.gen_stencil27(7, 4, True, dt, seed=7),
             sp.gen_uniform(1001, 998, 0, 30, dtype=dt, seed=8)]
    for A in cases:
        X = np.stack([sp.gen_vector(dt, A.ncols, seed=80 + q) for q in range(k)], axis=1)
        for kern in ("BCSR-block", "BCSR-nnz", "DBCSR", "RBDBCSR-nnz", "BDBCSR-block"):
            kw = dict(vert=3) if kern[0] in "DRB" and kern[1] != "C" else {}
            Y, _ = run(sp, ctx, kern, A, X, sp.opts(row_block_size=4, col_block_size=4, **kw), k=k)
            assert_parity_multi(dt, Y, A, X, f"{kern} k{k} n{A.nrows}")


@pytest.mark.parametrize("dt", [0, 1, 2, 3, 4, 5])
def test_coo_quad_rows(sp, ctx, dt):
    """COO of rows made of aligned runs of four columns (a 27-point stencil
    with 4 unknowns per point; random aligned quads with empty rows): the
    automatic plan takes row-ordered chunk tiles reading x four values per
    load, for every tile size and partition knob; integers bit-exact with the
    oracle and equal to the column-ordered plan; x at an address that is not a
    multiple of four values takes the gathering path."""
    rng = np.random.default_rng(60 + dt)
    nrows, ncols = 2500, 2000
    rows, cols = [], []
    for r in range(nrows):
        if r % 7 == 0:
            continue
        q = rng.choice(ncols // 4, size=int(rng.integers(1, 12)), replace=False)
        for c0 in np.sort(q):
            rows += [r] * 4
            cols += [4 * int(c0) + i for i in range(4)]
    vals = rng.integers(-4, 5, size=len(rows)).astype(sp.NP_DTYPE[dt])
    Q = sp.Coo.from_arrays(nrows, ncols, np.asarray(rows), np.asarray(cols), vals, sort=True)
    for A in (sp.gen_stencil27(5, 4, True, dt, seed=61), Q):
        x = sp.gen_vector(dt, A.ncols, seed=62)
        for kern in ("COO-row", "COO-nnz-rgrn", "COO-nnz"):
            for kw in (dict(), dict(nparts=16), dict(nparts=64, nsub=3)):
                for ipt in (4, 8, 16):
                    y, plan = run(sp, ctx, kern, A, x, sp.opts(items_per_lane=ipt, **kw))
                    assert plan.info()["algo"] == 2, (kern, kw, ipt)
                    assert_parity(dt, y, A, x, f"{kern}/quads/{kw}/ipt{ipt}")
                    if dt <= 3:
                        y3, _ = run(sp, ctx, kern, A, x, sp.opts(algo=3, items_per_lane=ipt, **kw))
                        assert np.array_equal(y, y3), (kern, kw, ipt)
        plan = sp.Plan(ctx, "COO-nnz", A, sp.opts())
        xe = np.concatenate([np.zeros(1, x.dtype), x])
        xd = sp.DeviceArray.from_numpy(ctx, xe)
        yd = sp.DeviceArray(ctx, (A.nrows,), x.dtype)
        plan.spmv(xd.addr + x.itemsize, yd.addr)
        ctx.sync()
        assert_parity(dt, yd.numpy(), A, x, f"COO quads/x offset/n{A.nrows}")
