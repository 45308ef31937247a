"""GPU parity of the sixteen 2D variants (DCSR .. BDBCOO-nnz): tiles write
per-stripe partial rows that a reduction adds in stripe order.  Integers
bit-exact; floats within the gamma(len+2)*sum|a x| bound."""
import numpy as np
import pytest

from gpu_util import assert_parity, assert_parity_multi, run

pytestmark = pytest.mark.gpu

K2D = ["DCSR", "DCOO", "DBCSR", "DBCOO", "RBDCSR", "RBDCOO", "RBDBCSR-block", "RBDBCSR-nnz", "RBDBCOO-block",
       "RBDBCOO-nnz", "BDCSR", "BDCOO", "BDBCSR-block", "BDBCSR-nnz", "BDBCOO-block", "BDBCOO-nnz"]


@pytest.mark.parametrize("dt", [0, 2, 4, 5])
def test_all_2d_variants(sp, ctx, dt):
    mats = {
        "rmat": sp.gen_rmat(12, 40000, dtype=dt, seed=5),
        "stencil27": sp.gen_stencil27(6, 4, True, dt, seed=3),
        "uniform": sp.gen_uniform(2000, 3001, 0, 25, dtype=dt, seed=8),
    }
    for mname, A in mats.items():
        x = sp.gen_vector(dt, A.ncols, seed=17)
        for kern in K2D:
            for kw in (dict(vert=1), dict(vert=2), dict(vert=3, nparts=24), dict(vert=8, nparts=64, nsub=4)):
                y, _ = run(sp, ctx, kern, A, x, sp.opts(**kw))
                assert_parity(dt, y, A, x, f"{mname}/{kern}/{kw}")


@pytest.mark.parametrize("dt", [2, 4])
def test_2d_csr_algorithms_and_blocks(sp, ctx, dt):
    A = sp.gen_rmat(11, 25000, dtype=dt, seed=9)
    x = sp.gen_vector(dt, A.ncols, seed=2)
    for kern in ("DCSR", "RBDCSR", "BDCSR"):
        for kw in (dict(algo=1, lanes_per_row=4), dict(algo=2, items_per_lane=4), dict(algo=2, items_per_lane=16),
                   dict(algo=1, items_per_lane=4), dict(algo=1, items_per_lane=4, lanes_per_row=16)):
            y, _ = run(sp, ctx, kern, A, x, sp.opts(vert=4, **kw))
            assert_parity(dt, y, A, x, f"{kern}/{kw}")
    # 2D CSR tiles in the chunked row kernel (each tile's entries 4-aligned in the arrays),
    # also with 3 right-hand sides
    B = sp.gen_stencil27(8, 4, True, dt, seed=6)
    xb = sp.gen_vector(dt, B.ncols, seed=3)
    Xb = np.stack([sp.gen_vector(dt, B.ncols, seed=30 + q) for q in range(3)], axis=1)
    for kern in ("DCSR", "RBDCSR", "BDCSR"):
        for kw in (dict(vert=3), dict(vert=2, nparts=8, nsub=3)):
            y, plan = run(sp, ctx, kern, B, xb, sp.opts(algo=1, items_per_lane=4, **kw))
            assert plan.info()["algo"] == 1 and plan.info()["items_per_lane"] == 4
            assert_parity(dt, y, B, xb, f"stencil {kern}/{kw}")
        Y, _ = run(sp, ctx, kern, B, Xb, sp.opts(algo=1, items_per_lane=4, vert=3), k=3)
        assert_parity_multi(dt, Y, B, Xb, f"stencil spmm {kern}")
    for kern in ("DBCSR", "RBDBCSR-nnz", "BDBCOO-block"):
        for r, c in ((2, 2), (8, 8), (2, 4)):
            y, _ = run(sp, ctx, kern, A, x, sp.opts(vert=3, row_block_size=r, col_block_size=c))
            assert_parity(dt, y, A, x, f"{kern}/{r}x{c}")


@pytest.mark.parametrize("k", [2, 5])
def test_2d_spmm(sp, ctx, k):
    # the stencil's tiles are made of aligned four-column runs (padded chunk layouts)
    for A in (sp.gen_rmat(10, 9000, dtype=4, seed=4), sp.gen_stencil27(6, 4, True, 4, seed=5)):
        X = np.stack([sp.gen_vector(4, A.ncols, seed=60 + q) for q in range(k)], axis=1)
        for kern in K2D:
            Y, _ = run(sp, ctx, kern, A, X, sp.opts(vert=3), k=k)
            assert_parity_multi(4, Y, A, X, f"{kern}/n{A.nrows}")


def test_2d_edge_cases(sp, ctx):
    rng = np.random.default_rng(1)
    # more stripes than columns with entries, tiles past the last row, empty matrix
    cases = [(37, 5, [0, 3, 36, 36], [4, 0, 1, 4]), (10, 100, [], []), (1000, 7, list(range(0, 1000, 3)), [6] * 334)]
    for nr, nc, r, c in cases:
        v = rng.integers(-4, 5, len(r)).astype(np.int32)
        A = sp.Coo.from_arrays(nr, nc, np.asarray(r), np.asarray(c), v, sort=True)
        x = rng.integers(-4, 5, nc).astype(np.int32)
        for kern in K2D:
            for kw in (dict(vert=4), dict(vert=16, nparts=64)):
                y, _ = run(sp, ctx, kern, A, x, sp.opts(**kw))
                assert_parity(2, y, A, x, f"{nr}x{nc}/{kern}/{kw}")


@pytest.mark.parametrize("dt", [0, 2, 4, 5])
def test_lds_x_slices(sp, ctx, dt):
    """x slices staged in LDS (lds_slice) for the 2D CSR and COO tilings give
    the same bits as gathers from global memory, with one and several
    right-hand sides (several run without the slice), and a stripe wider than
    the slice is refused."""
    A = sp.gen_rmat(14, 200000, dtype=dt, seed=61)       # 16384 columns
    x = sp.gen_vector(dt, A.ncols, seed=62)
    X = np.stack([sp.gen_vector(dt, A.ncols, seed=63 + q) for q in range(3)], axis=1)
    cap = 32768 // np.dtype(sp.NP_DTYPE[dt]).itemsize
    vert = max(2, -(-A.ncols // cap))
    vert = 1 << (vert - 1).bit_length()
    for kern in ("DCSR", "DCOO", "RBDCSR", "RBDCOO", "BDCSR", "BDCOO"):
        for kw in (dict(), dict(items_per_lane=4), dict(nparts=64, nsub=4)):
            base = dict(vert=2 * vert, **kw)   # even stripes at most cap / 2 wide; BD stripes may vary
            y0, _ = run(sp, ctx, kern, A, x, sp.opts(**base))
            try:
                y1, _ = run(sp, ctx, kern, A, x, sp.opts(lds_slice=cap, **base))
            except sp._lib.SparsePError as e:
                assert kern.startswith("BD") and e.rc == -5, (kern, kw)  # a variable stripe too wide
                continue
            assert_parity(dt, y1, A, x, f"{kern}/{kw}/slice")
            assert np.array_equal(y0, y1), f"{kern}/{kw}"
            Y1, _ = run(sp, ctx, kern, A, X, sp.opts(lds_slice=cap, **base), k=3)
            assert_parity_multi(dt, Y1, A, X, f"{kern}/{kw}/slice spmm")
    with pytest.raises(sp._lib.SparsePError):
        sp.Plan(ctx, "RBDCSR", A, sp.opts(vert=2, lds_slice=A.ncols // 4))


@pytest.mark.parametrize("dt", [2, 4, 5])
def test_2d_column_order_same_bits(sp, ctx, dt):
    """2D CSR (merge path) and COO tiles in column order (the default where the
    columns fit) give the same bits as row-ordered tiles (algo=2)."""
    A = sp.gen_rmat(15, 400000, dtype=dt, seed=81)
    x = sp.gen_vector(dt, A.ncols, seed=82)
    for kern in ("DCSR", "RBDCSR", "BDCSR", "DCOO", "RBDCOO", "BDCOO"):
        for kw in (dict(vert=2), dict(vert=4, nparts=32, nsub=4), dict(vert=3, items_per_lane=16)):
            extra = dict(algo=3) if "CSR" in kern else {}   # CSR: column-ordered merge tiles
            y0, _ = run(sp, ctx, kern, A, x, sp.opts(algo=2, **kw))
            y1, plan = run(sp, ctx, kern, A, x, sp.opts(**extra, **kw))
            assert plan.info()["algo"] == 3, (kern, kw)
            assert_parity(dt, y1, A, x, f"{kern}/{kw}")
            assert np.array_equal(y0, y1), f"{kern}/{kw}"


@pytest.mark.parametrize("dt", [0, 1, 2, 3, 4, 5])
def test_2d_atomic_merges(sp, ctx, dt):
    """algo=5 on the 2D COO/BCOO tilings: rows shared between a stripe's tiles
    are added into the zeroed partial buffers with atomics (the 1- and 2-byte
    integers by compare-and-swap on the enclosing word, 8-byte integers by a
    64-bit add).  Integers bit-exact and equal to the carry-merged plans,
    floats within the bound; also three right-hand sides."""
    mats = [sp.gen_rmat(12, 40000, dtype=dt, seed=15), sp.gen_uniform(2000, 3001, 0, 25, dtype=dt, seed=18)]
    kerns = ["DCOO", "DBCOO", "RBDCOO", "RBDBCOO-block", "RBDBCOO-nnz", "BDCOO", "BDBCOO-block", "BDBCOO-nnz"]
    for i, A in enumerate(mats):
        x = sp.gen_vector(dt, A.ncols, seed=27 + i)
        for kern in kerns:
            for kw in (dict(vert=1), dict(vert=4, nparts=16), dict(vert=8, nparts=64, nsub=4)):
                y, plan = run(sp, ctx, kern, A, x, sp.opts(algo=5, **kw))
                assert plan.info()["algo"] == 5
                assert_parity(dt, y, A, x, f"mat{i}/{kern}/{kw}")
                if dt <= 3:
                    y0, _ = run(sp, ctx, kern, A, x, sp.opts(**kw))
                    assert np.array_equal(y, y0), f"mat{i}/{kern}/{kw}"
    A = mats[0]
    X = np.stack([sp.gen_vector(dt, A.ncols, seed=40 + q) for q in range(3)], axis=1)
    for kern in ("DCOO", "BDBCOO-nnz"):
        Y, _ = run(sp, ctx, kern, A, X, sp.opts(algo=5, vert=4), k=3)
        assert_parity_multi(dt, Y, A, X, f"spmm {kern}")
