"""Autotuner: every candidate point of all 25 variants builds and times on
the GPU; the best point is no slower than any other; the timing database
round-trips."""
import json

import pytest

pytestmark = pytest.mark.gpu


def test_exhaustive_tuner(sp, ctx, tmp_path):
    A = sp.gen_rmat(12, 40000, dtype=sp.F32, seed=3)
    pts = sp.tune(ctx, A, iters=3)
    ok = [p for p in pts if p.status == 0]
    assert {p.kernel_id for p in ok} == set(range(25)), "every variant must have a working point"
    assert all(p.ms > 0 for p in ok)
    b = sp.tune_best(pts)
    assert all(pts[b].ms <= p.ms for p in ok)
    for kid in range(25):
        bk = sp.tune_best(pts, kid)
        assert pts[bk].kernel_id == kid
    path = str(tmp_path / "db.jsonl")
    sp.lib().sp_tune_db_append(path.encode(), b"rmat12", pts, len(pts))
    rows = [json.loads(l) for l in open(path)]
    assert len(rows) == len(pts) and rows[0]["matrix"] == "rmat12"


def test_tuner_for_several_right_hand_sides(sp, ctx):
    """sp_tune_run_spmm times every CSR point with sp_spmm for k columns; the
    best point's plan computes Y = A X within the oracle's bound, and a
    row-granular point (one pass over the matrix for up to 32 columns) is
    timed for the regular matrix."""
    import numpy as np
    from gpu_util import assert_parity_multi, run
    A = sp.gen_uniform(20000, 20000, 8, 40, dtype=sp.F32, seed=9)
    k = 32
    pts = sp.tune(ctx, A, sp.tune_space(A, sp.KID["CSR-nnz"]), iters=3, k=k)
    ok = [p for p in pts if p.status == 0]
    assert ok and all(p.ms > 0 for p in ok)
    assert any(p.opts.algo == sp.ALGO_VECTOR for p in ok)
    b = sp.tune_best(pts)
    X = np.stack([sp.gen_vector(sp.F32, A.ncols, seed=70 + q) for q in range(k)], axis=1)
    Y, _ = run(sp, ctx, "CSR-nnz", A, X, sp.point_opts(pts[b]), k=k)
    assert_parity_multi(sp.F32, Y, A, X, "tuned for k=32")
