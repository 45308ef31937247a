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
