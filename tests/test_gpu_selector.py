"""The learned variant selector on the GPU: the committed model
(sparsep_amd/data/selector_f32.bin) scores the held-out part of the committed
timing corpus (sparsep_amd/data/selector_corpus.jsonl.gz, labelled by the
exhaustive autotuner) with the HIP MLP kernels, and its picks are compared
with the exhaustive best and with the best single fixed variant (the method
without the network)."""
import gzip
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
DATA = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "sparsep_amd", "data")
MODEL = os.path.join(DATA, "selector_f32.bin")
CORPUS = os.path.join(DATA, "selector_corpus.jsonl.gz")


if not (os.path.exists(MODEL) and os.path.exists(CORPUS)):
    pytest.skip("selector model / corpus not trained yet (python -m sparsep_amd.selector tune|train)",
                allow_module_level=True)


@pytest.fixture(scope="module")
def heldout(sp, ctx):
    from sparsep_amd import selector as S
    with gzip.open(CORPUS, "rt") as f:
        recs = sorted((json.loads(line) for line in f if line.strip()), key=lambda r: r["idx"])
    X, T, tdef, best = S.labels(recs)
    tr, ho = S.split(len(recs), 0.2)
    model = S.read_model(MODEL)
    pred = S.model_scores(ctx, model, X[ho])
    return S, recs, model, pred, tdef, best, tr, ho


def test_learned_pick_beats_best_fixed_variant(sp, ctx, heldout):
    S, recs, model, pred, tdef, best, tr, ho = heldout
    learned = S.evaluate(pred, tdef[ho], best[ho])
    k, _ = S.best_fixed_variant(tdef[tr], best[tr])
    fixed = np.where(np.isfinite(tdef[ho, k]), best[ho] / tdef[ho, k], 1e-9)
    fixed_geo = float(np.exp(np.mean(np.log(fixed))))
    print(f"held-out {len(ho)}: learned geomean {learned['geomean_ratio']:.3f} "
          f"(within 10%: {learned['within_10pct']:.2f}), best fixed variant {sp.KERNELS[k]} {fixed_geo:.3f}")
    assert learned["failed_picks"] == 0
    assert learned["geomean_ratio"] > fixed_geo
    assert learned["geomean_ratio"] >= 0.90


def test_c_selector_matches_batch_scores(sp, ctx, heldout):
    S, recs, model, pred, tdef, best, tr, ho = heldout
    sel = sp.Selector(ctx, path=MODEL)
    for j, i in enumerate(ho[:6]):
        A = S.build_matrix(recs[i]["spec"])
        pt, scores, fell = sel.select(A, fallback_k=0)
        assert not fell
        np.testing.assert_allclose(scores, pred[j], rtol=1e-5, atol=1e-5)
        assert pt.kernel_id == int(np.argmin(scores))


def test_fallback_search_is_never_worse(sp, ctx, heldout):
    S, recs, model, pred, tdef, best, tr, ho = heldout
    plain = S.evaluate(pred, tdef[ho], best[ho])
    fb = S.evaluate(pred, tdef[ho], best[ho], [recs[i] for i in ho], margin=model["margin"], fallback_k=3)
    assert fb["geomean_ratio"] >= plain["geomean_ratio"] - 1e-9
    assert 0.0 < fb["fallback_frac"] <= 0.5


def test_select_with_live_fallback(sp, ctx):
    from sparsep_amd import selector as S
    sel = sp.Selector(ctx, path=MODEL)
    A = S.build_matrix({"family": "banded", "seed": 9, "nrows": 20000, "bw": 32, "klo": 4, "khi": 12})
    pt, scores, fell = sel.select(A, fallback_k=25)
    assert 0 <= pt.kernel_id < 25 and np.all(np.isfinite(scores))
    if fell:
        assert pt.status == 0 and pt.ms > 0
    plan = sp.Plan(ctx, pt.kernel_id, A, sp.point_opts(pt))
    assert plan.info()["kernel_id"] == pt.kernel_id
