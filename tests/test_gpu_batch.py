"""GPU parity of the batched SpMV/SpMM path: many uneven problems with
1/4/8/32 right-hand sides in persistent launches captured in a hipGraph,
checked per problem against the oracle, over several graph replays with
new inputs."""
import numpy as np
import pytest

from gpu_util import assert_parity_multi

pytestmark = pytest.mark.gpu


def problems(sp, dt, n=12, seed=4):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        rows = int(np.exp(rng.uniform(np.log(500), np.log(20000))))
        kind = i % 3
        if kind == 0:
            A = sp.gen_uniform(rows, rows, 2, 64, dtype=dt, seed=seed + i)
        elif kind == 1:
            A = sp.gen_banded(rows, 32, 2, 64, dtype=dt, seed=seed + i)
        else:
            scale = max(9, int(np.ceil(np.log2(rows))))
            A = sp.gen_rmat(scale, rows * 16, dtype=dt, seed=seed + i)
        out.append((A, [1, 4, 8, 32][i % 4]))
    return out


@pytest.mark.parametrize("dt", [1, 4])
@pytest.mark.parametrize("graph", [True, False])
def test_batch_matches_oracle(sp, ctx, dt, graph):
    probs = problems(sp, dt)
    kinds = ["CSR-nnz", "CSR-nnz", "COO-nnz", "CSR-row"]
    plans = [sp.Plan(ctx, kinds[i % 4], A, sp.opts(algo=2) if kinds[i % 4] == "CSR-nnz" else sp.opts())
             for i, (A, k) in enumerate(probs)]
    b = sp.Batch(ctx, plans, [k for _, k in probs], graph=graph)
    assert b.info()["groups"] >= 1
    for rep in range(3):  # graph replays with new inputs each time
        Xs, Xd, Yd = [], [], []
        for i, (A, k) in enumerate(probs):
            X = np.stack([sp.gen_vector(dt, A.ncols, seed=100 * rep + 7 * i + q) for q in range(k)], axis=1)
            Xs.append(X)
            Xd.append(sp.DeviceArray.from_numpy(ctx, X))
            y = sp.DeviceArray(ctx, (A.nrows, k), X.dtype)
            y.fill_bytes(0x5A)
            Yd.append(y)
        b.run([x.addr for x in Xd], [y.addr for y in Yd])
        ctx.sync()
        for i, (A, k) in enumerate(probs):
            assert_parity_multi(dt, Yd[i].numpy(), A, Xs[i], f"rep{rep} prob{i}")


def test_batch_rejects_mixed_dtypes(sp, ctx):
    A = sp.gen_uniform(100, 100, 1, 5, dtype=4, seed=1)
    B = sp.gen_uniform(100, 100, 1, 5, dtype=5, seed=1)
    with pytest.raises(sp._lib.SparsePError):
        sp.Batch(ctx, [sp.Plan(ctx, "CSR-nnz", A), sp.Plan(ctx, "CSR-nnz", B)], [1, 1])


@pytest.mark.parametrize("dt", [2, 4])
def test_batch_hot_plans(sp, ctx, dt):
    """Hot-x CSR plans in a batch (one plan in two slots: each slot stages its
    own x), mixed with merge-path plans, over graph replays."""
    probs = problems(sp, dt, n=6, seed=9)
    plans = [sp.Plan(ctx, "CSR-nnz", A, sp.opts(algo=4 if i % 2 == 0 else 2)) for i, (A, k) in enumerate(probs)]
    plans.append(plans[0])
    probs.append((probs[0][0], 3))
    b = sp.Batch(ctx, plans, [k for _, k in probs], graph=True)
    for rep in range(3):
        Xs, Xd, Yd = [], [], []
        for i, (A, k) in enumerate(probs):
            X = np.stack([sp.gen_vector(dt, A.ncols, seed=300 * rep + 11 * i + q) for q in range(k)], axis=1)
            Xs.append(X)
            Xd.append(sp.DeviceArray.from_numpy(ctx, X))
            y = sp.DeviceArray(ctx, (A.nrows, k), X.dtype)
            y.fill_bytes(0x5A)
            Yd.append(y)
        b.run([x.addr for x in Xd], [y.addr for y in Yd])
        ctx.sync()
        for i, (A, k) in enumerate(probs):
            assert_parity_multi(dt, Yd[i].numpy(), A, Xs[i], f"rep{rep} prob{i}")


@pytest.mark.parametrize("dt", [2, 4, 5])
def test_batch_block_plans(sp, ctx, dt):
    """Block formats in a batch: entry-balanced BCSR, row-granular BCSR and
    BCOO plans (4x4 and 2x2 blocks) over several right-hand sides."""
    probs = problems(sp, dt, n=6, seed=13)
    opts = [dict(algo=2), dict(), dict(algo=2, row_block_size=2, col_block_size=2)]
    kinds = ["BCSR-nnz", "BCSR-block", "BCOO-nnz"]
    plans = [sp.Plan(ctx, kinds[i % 3], A, sp.opts(**opts[(i // 3) % 3])) for i, (A, k) in enumerate(probs)]
    b = sp.Batch(ctx, plans, [k for _, k in probs], graph=True)
    for rep in range(2):
        Xs, Xd, Yd = [], [], []
        for i, (A, k) in enumerate(probs):
            X = np.stack([sp.gen_vector(dt, A.ncols, seed=500 * rep + 13 * i + q) for q in range(k)], axis=1)
            Xs.append(X)
            Xd.append(sp.DeviceArray.from_numpy(ctx, X))
            y = sp.DeviceArray(ctx, (A.nrows, k), X.dtype)
            y.fill_bytes(0x5A)
            Yd.append(y)
        b.run([x.addr for x in Xd], [y.addr for y in Yd])
        ctx.sync()
        for i, (A, k) in enumerate(probs):
            assert_parity_multi(dt, Yd[i].numpy(), A, Xs[i], f"rep{rep} prob{i}")
