"""Learned variant selector pipeline: the synthetic training corpus, the
exhaustive-autotuner timing run that labels it, training of the MLP with the
library's HIP kernels, and held-out evaluation against the exhaustive best.

Labels: for matrix i and kernel id k, t[i,k] is the median SpMV time of k's
default configuration (the first tuner point of k) and best[i] the fastest
time over every tuner point of every id; the MLP regresses
log2(t[i,k] / best[i]) (clipped to LABEL_CAP; ids whose plan could not be
built get LABEL_CAP), so the lowest output is the predicted fastest id.
No counterpart in the reference.
"""
import gzip
import json
import math
import time

import numpy as np

from . import api

LABEL_CAP = 6.0
FAMILIES = ("uniform", "rmat", "banded", "blockdiag", "stencil2d", "stencil3d")


# --------------------------------------------------------------------------
# corpus
# --------------------------------------------------------------------------
def corpus_specs(count, seed=5, rows=(10_000, 2_000_000), nnz=(50_000, 40_000_000)):
    """Deterministic list of `count` matrix specs, families taken in turn.
    Rows are log-uniform in `rows`; each spec is re-drawn until its expected
    nnz falls in `nnz`."""
    r = np.random.default_rng(seed)
    specs = []
    lo, hi = math.log(rows[0]), math.log(rows[1])
    i = 0
    while len(specs) < count:
        fam = FAMILIES[i % len(FAMILIES)]
        i += 1
        n = int(math.exp(r.uniform(lo, hi)))
        s = {"family": fam, "seed": int(seed * 1_000_003 + i)}
        if fam == "uniform":
            k = int(r.integers(2, 48))
            s.update(nrows=n, klo=max(1, k // 2), khi=k + k // 2)
            e = n * k
        elif fam == "rmat":
            scale = max(10, int(round(math.log2(n))))
            a = float(r.uniform(0.45, 0.7))
            deg = int(r.integers(4, 32))
            s.update(scale=scale, a=a, b=(1 - a) * 0.35, c=(1 - a) * 0.35, nnz=(1 << scale) * deg)
            e = (1 << scale) * deg
        elif fam == "banded":
            bw = int(r.choice([8, 32, 128, 512]))
            k = int(r.integers(2, min(2 * bw, 48)))
            s.update(nrows=n, bw=bw, klo=max(1, k // 2), khi=k + k // 2)
            e = n * k
        elif fam == "blockdiag":
            k = int(r.integers(2, 48))
            s.update(nrows=n, nblocks=int(r.choice([2, 4, 8])), klo=max(1, k // 2), khi=k + k // 2)
            e = n * k
        elif fam == "stencil2d":
            gx = max(8, int(math.sqrt(n) * r.uniform(0.5, 2.0)))
            gy = max(8, n // gx)
            s.update(gx=gx, gy=gy)
            e = 5 * gx * gy
        else:
            dof = int(r.choice([1, 2, 4]))
            g = max(4, int(round((n / dof) ** (1 / 3))))
            s.update(g=g, dof=dof)
            e = 27 * dof * dof * g ** 3  # g^3 dof rows of 27 dof entries
        if nnz[0] <= e <= nnz[1]:
            specs.append(s)
    return specs


def build_matrix(s, dtype=api.F32):
    fam = s["family"]
    if fam == "uniform":
        return api.gen_uniform(s["nrows"], s["nrows"], s["klo"], s["khi"], dtype=dtype, seed=s["seed"])
    if fam == "rmat":
        return api.gen_rmat(s["scale"], s["nnz"], s["a"], s["b"], s["c"], dtype=dtype, seed=s["seed"])
    if fam == "banded":
        return api.gen_banded(s["nrows"], s["bw"], s["klo"], s["khi"], dtype=dtype, seed=s["seed"])
    if fam == "blockdiag":
        return api.gen_blockdiag(s["nrows"], s["nblocks"], s["klo"], s["khi"], dtype=dtype, seed=s["seed"])
    if fam == "stencil2d":
        return api.gen_stencil5(s["gx"], s["gy"], dtype, offdiag=api.dist(api.DIST_UNIFORM, -1, 1),
                                diag=4.0, seed=s["seed"])
    return api.gen_stencil27(s["g"], s["dof"], True, dtype, seed=s["seed"])


# --------------------------------------------------------------------------
# labelling run (exhaustive autotuner)
# --------------------------------------------------------------------------
def tune_corpus(ctx, specs, out_path, iters=5, flush_bytes=512 << 20, start=0, log=print):
    """Tunes every spec and appends one JSON line per matrix to out_path:
    features, every tuner point (kernel id, point index within the id, ms,
    build_ms, status) and the wall time of the exhaustive run."""
    with open(out_path, "a") as f:
        for i in range(start, len(specs)):
            s = specs[i]
            A = build_matrix(s)
            feat = api.features(A)
            pts = api.tune_space(A)
            t0 = time.perf_counter()
            api.tune(ctx, A, pts, iters=iters, flush_bytes=flush_bytes)
            wall = time.perf_counter() - t0
            seen = {}
            rows = []
            for p in pts:
                j = seen.get(p.kernel_id, 0)
                seen[p.kernel_id] = j + 1
                rows.append([p.kernel_id, j, round(float(p.ms), 6), round(float(p.build_ms), 3), int(p.status)])
            rec = {"idx": i, "spec": s, "nrows": A.nrows, "nnz": A.nnz, "features": [float(v) for v in feat],
                   "points": rows, "tune_s": wall, "iters": iters}
            f.write(json.dumps(rec) + "\n")
            f.flush()
            ok = [r[2] for r in rows if r[4] == 0 and r[2] > 0]
            log(f"[tune] {i + 1}/{len(specs)} {s['family']} n={A.nrows} nnz={A.nnz} "
                f"best={min(ok) if ok else -1:.4f}ms wall={wall:.1f}s")


def load_records(path):
    """Timing records of a corpus file (JSON lines, optionally gzip'd)."""
    opener = gzip.open if str(path).endswith(".gz") else open
    with opener(path, "rt") as f:
        return [json.loads(line) for line in f if line.strip()]


def labels(records, nk=25):
    """(X features, T labels, t_default ms [inf if failed], t_best ms,
    per-point table) from timing records."""
    X = np.array([r["features"] for r in records], np.float32)
    n = len(records)
    tdef = np.full((n, nk), np.inf)
    best = np.full(n, np.inf)
    for i, r in enumerate(records):
        for k, j, ms, _bms, st in r["points"]:
            if st == 0 and ms > 0:
                best[i] = min(best[i], ms)
                if j == 0:
                    tdef[i, k] = ms
    T = np.where(np.isfinite(tdef), np.log2(tdef / best[:, None]), LABEL_CAP)
    T = np.minimum(T, LABEL_CAP).astype(np.float32)
    return X, T, tdef, best


def point_cost_ms(rec, ids=None):
    """Sequential cost of tuning the points of `ids` (all when None): build
    time plus the warm-up and timed SpMVs."""
    it = rec.get("iters", 5) + 1
    c = 0.0
    for k, _j, ms, bms, st in rec["points"]:
        if ids is None or k in ids:
            c += bms + (it * ms if st == 0 and ms > 0 else 0.0)
    return c


# --------------------------------------------------------------------------
# training and evaluation
# --------------------------------------------------------------------------
def read_model(path):
    """Fields of a model file written by sp_selector_save: sizes[4],
    margin, feature mean[64], feature std[64], parameters."""
    with open(path, "rb") as f:
        raw = f.read()
    if raw[:8] != b"SPSEL001":
        raise ValueError(f"{path}: not a selector model file")
    sizes = tuple(int(v) for v in np.frombuffer(raw[8:24], np.int32))
    margin = float(np.frombuffer(raw[24:28], np.float32)[0])
    mean = np.frombuffer(raw[28:284], np.float32).copy()
    std = np.frombuffer(raw[284:540], np.float32).copy()
    params = np.frombuffer(raw[540:], np.float32).copy()
    return {"sizes": sizes, "margin": margin, "mean": mean, "std": std, "params": params}


def model_scores(ctx, model, X):
    """Scores of feature rows X with the model's MLP (HIP kernels)."""
    m = api.Mlp(ctx, model["sizes"])
    m.set_params(model["params"])
    return m.predict((X - model["mean"]) / np.maximum(model["std"], 1e-6))


def standardise(X):
    mean = X.mean(0)
    std = X.std(0)
    std = np.where(std > 1e-6, std, 1.0)
    return mean.astype(np.float32), std.astype(np.float32)


def train(ctx, X, T, hidden=(128, 128), epochs=400, batch=64, lr=1e-3, seed=5):
    mean, std = standardise(X)
    Z = (X - mean) / std
    mlp = api.Mlp(ctx, (X.shape[1], hidden[0], hidden[1], T.shape[1]), seed=seed)
    loss = mlp.train(Z, T, epochs=epochs, batch=min(batch, len(Z)), lr=lr, seed=seed)
    return mlp, mean, std, loss


def evaluate(pred, tdef, best, records=None, margin=None, fallback_k=3):
    """Held-out quality of picks from scores `pred`: geometric-mean and
    aggregate (sum best / sum pick) throughput ratios; with a margin, picks
    whose top-2 gap is below it use the best point of the fallback_k
    lowest-scored ids, and the fallback tuning cost is returned as well."""
    n = len(pred)
    order = np.argsort(pred, axis=1, kind="stable")
    pick_t = tdef[np.arange(n), order[:, 0]].copy()
    fell = np.zeros(n, bool)
    fb_cost = np.zeros(n)
    if margin is not None and records is not None:
        gap = pred[np.arange(n), order[:, 1]] - pred[np.arange(n), order[:, 0]]
        for i in np.nonzero(gap < margin)[0]:
            ids = set(int(k) for k in order[i, :fallback_k])
            ok = [ms for k, _j, ms, _b, st in records[i]["points"] if k in ids and st == 0 and ms > 0]
            if ok:
                pick_t[i] = min(ok)
            fell[i] = True
            fb_cost[i] = point_cost_ms(records[i], ids)
    ratio = np.where(np.isfinite(pick_t), best / pick_t, 0.0)
    fin = np.isfinite(pick_t)
    return {
        "geomean_ratio": float(np.exp(np.mean(np.log(np.maximum(ratio, 1e-9))))),
        "aggregate_ratio": float(best[fin].sum() / pick_t[fin].sum()) if fin.any() else 0.0,
        "within_10pct": float(np.mean(ratio >= 1 / 1.1)),
        "failed_picks": int((~fin).sum()),
        "fallback_frac": float(fell.mean()),
        "fallback_cost_ms": fb_cost,
    }


def best_fixed_variant(tdef, best):
    """The single kernel id (default configuration) with the best geometric
    mean ratio over a set: the no-model baseline."""
    r = np.where(np.isfinite(tdef), best[:, None] / tdef, 1e-9)
    g = np.exp(np.mean(np.log(r), axis=0))
    return int(np.argmax(g)), g


# --------------------------------------------------------------------------
# command line: python -m sparsep_amd.selector {tune,train} ...
# --------------------------------------------------------------------------
def split(n, holdout):
    """Deterministic train / held-out split: every k-th record is held out."""
    k = max(2, int(round(1 / holdout)))
    idx = np.arange(n)
    return idx[idx % k != k - 1], idx[idx % k == k - 1]


def _selection_overhead(ctx, sel, records, idx, reps=3):
    """Live features + inference time (sp_select without fallback) per
    held-out matrix, in ms; matrix generation is not timed."""
    out = []
    for i in idx:
        A = build_matrix(records[i]["spec"])
        ts = []
        for _ in range(reps):
            t0 = time.perf_counter()
            sel.select(A, fallback_k=0)
            ts.append((time.perf_counter() - t0) * 1e3)
        out.append(min(ts))
    return np.array(out)


def main(argv=None):
    import argparse
    ap = argparse.ArgumentParser(prog="python -m sparsep_amd.selector")
    sub = ap.add_subparsers(dest="cmd", required=True)
    t = sub.add_parser("tune", help="label a synthetic corpus with the exhaustive autotuner")
    t.add_argument("--count", type=int, default=2500)
    t.add_argument("--seed", type=int, default=5)
    t.add_argument("--rows", type=int, nargs=2, default=[10_000, 2_000_000])
    t.add_argument("--nnz", type=int, nargs=2, default=[50_000, 40_000_000])
    t.add_argument("--iters", type=int, default=5)
    t.add_argument("--flush-mb", type=int, default=512)
    t.add_argument("--start", type=int, default=0)
    t.add_argument("--stop", type=int, default=None)
    t.add_argument("--out", required=True)
    r = sub.add_parser("train", help="train the selector MLP and evaluate it on held-out matrices")
    r.add_argument("--records", nargs="+", required=True)
    r.add_argument("--model", required=True)
    r.add_argument("--report", required=True)
    r.add_argument("--holdout", type=float, default=0.2)
    r.add_argument("--epochs", type=int, default=400)
    r.add_argument("--lr", type=float, default=1e-3)
    r.add_argument("--fallback-k", type=int, default=3)
    r.add_argument("--overhead-samples", type=int, default=50)
    a = ap.parse_args(argv)

    ctx = api.Context(0)
    if a.cmd == "tune":
        specs = corpus_specs(a.count, a.seed, tuple(a.rows), tuple(a.nnz))
        if a.stop is not None:
            specs = specs[:a.stop]
        tune_corpus(ctx, specs, a.out, iters=a.iters, flush_bytes=a.flush_mb << 20, start=a.start,
                    log=lambda m: print(m, flush=True))
        return 0

    recs = []
    for p in a.records:
        recs += load_records(p)
    recs.sort(key=lambda r: r["idx"])
    X, T, tdef, best = labels(recs)
    tr, ho = split(len(recs), a.holdout)
    t0 = time.perf_counter()
    mlp, mean, std, loss = train(ctx, X[tr], T[tr], epochs=a.epochs, lr=a.lr)
    train_s = time.perf_counter() - t0
    pred_tr = mlp.predict((X[tr] - mean) / std)
    pred_ho = mlp.predict((X[ho] - mean) / std)
    # margin: the 10th percentile of the top-2 score gap on the training set
    s_tr = np.sort(pred_tr, axis=1)
    margin = float(np.quantile(s_tr[:, 1] - s_tr[:, 0], 0.10))
    sel = api.Selector(ctx, params=mlp.get_params(), mean=mean, std=std, margin=margin)
    sel.save(a.model)
    ho_recs = [recs[i] for i in ho]
    plain = evaluate(pred_ho, tdef[ho], best[ho])
    fb = evaluate(pred_ho, tdef[ho], best[ho], ho_recs, margin=margin, fallback_k=a.fallback_k)
    fixed_k, fixed_g_tr = best_fixed_variant(tdef[tr], best[tr])
    fixed = np.where(np.isfinite(tdef[ho, fixed_k]), best[ho] / tdef[ho, fixed_k], 0.0)
    tune_ms = np.array([point_cost_ms(r) for r in ho_recs])
    tune_wall_ms = np.array([r["tune_s"] * 1e3 for r in ho_recs])
    nov = min(a.overhead_samples, len(ho))
    ov_ms = _selection_overhead(ctx, sel, recs, ho[:nov])
    rep = {
        "corpus": {"matrices": len(recs), "train": int(len(tr)), "held_out": int(len(ho)),
                   "families": {f: sum(r["spec"]["family"] == f for r in recs) for f in FAMILIES},
                   "rows": [int(min(r["nrows"] for r in recs)), int(max(r["nrows"] for r in recs))],
                   "nnz": [int(min(r["nnz"] for r in recs)), int(max(r["nnz"] for r in recs))]},
        "model": {"sizes": [int(X.shape[1]), 128, 128, int(T.shape[1])], "params": int(mlp.nparams),
                  "epochs": a.epochs, "final_train_loss": loss, "train_s": train_s, "margin": margin},
        "held_out": {
            "learned": {k: v for k, v in plain.items() if k != "fallback_cost_ms"},
            "learned_with_fallback": {k: v for k, v in fb.items() if k != "fallback_cost_ms"},
            "no_model_best_fixed_variant": {
                "kernel": api.KERNELS[fixed_k],
                "geomean_ratio": float(np.exp(np.mean(np.log(np.maximum(fixed, 1e-9))))),
                "aggregate_ratio": float(best[ho][fixed > 0].sum() / tdef[ho, fixed_k][fixed > 0].sum()),
                "within_10pct": float(np.mean(fixed >= 1 / 1.1))},
            "overhead": {
                "select_ms_median": float(np.median(ov_ms)),
                "exhaustive_cost_ms_median": float(np.median(tune_ms[:nov])),
                "exhaustive_wall_ms_median": float(np.median(tune_wall_ms[:nov])),
                "select_over_exhaustive_median": float(np.median(ov_ms / tune_ms[:nov])),
                "select_over_exhaustive_wall_median": float(np.median(ov_ms / tune_wall_ms[:nov])),
                "with_fallback_over_exhaustive_mean": float(np.mean((fb["fallback_cost_ms"][:nov] + ov_ms)
                                                                    / tune_ms[:nov])),
            },
        },
    }
    with open(a.report, "w") as f:
        json.dump(rep, f, indent=1)
    print(json.dumps(rep["held_out"], indent=1), flush=True)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
