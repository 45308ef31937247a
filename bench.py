#!/usr/bin/env python3
"""SpMV bench on one MI355X.

Workload: the seeded R-MAT matrix of 2^20 x 2^20 with 2^24 unique entries
(quadrants a=0.57, b=0.19, c=0.19; values and x uniform in [-1,1]; seed 2),
FP32, CSR with the entry-balanced (merge-path) kernel.  One step = one
SpMV y = A x with A, x, y resident in HBM.  Before every step a 1 GiB
buffer is overwritten so that L2 and the 256 MB MALL hold nothing of A or x
(the 147 MB working set would otherwise be served from MALL); the flush is
outside the timed events.

value = algorithmic bytes of one SpMV (nnz*(4+4) + (n+1)*4 + (m+n)*4,
x and y counted once) / mean SpMV time measured with HIP events on the
library's stream.  roofline.achieved uses the duration of the dominant
(matrix-streaming) kernel alone.  cpu_baseline times the oracle's C
restatement of SparseP's CSR spmv_host (single thread, reference order) on
the same matrix and vector, on this host.
"""
import argparse
import json
import os
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "SpMV GB/s & fraction of HBM peak; learned-pick vs exhaustive-best ratio"
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--flush-mb", type=int, default=1024)
    ap.add_argument("--cpu-budget", type=float, default=15.0, help="seconds of CPU baseline timing")
    ap.add_argument("--scale", type=int, default=20)
    ap.add_argument("--nnz-log2", type=int, default=24)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py measures one GPU only: this build targets a single MI355X and has no "
              f"multi-GPU path (asked for --gpus {args.gpus}).", file=sys.stderr)
        sys.exit(2)

    import numpy as np

    import sparsep_amd.api as sp

    sys.path.insert(0, os.path.join(ROOT, "tests"))

    ctx = sp.Context(0)
    info = ctx.info()
    log(f"[bench] device {info['arch']} CUs={info['num_cus']} HBM={info['hbm_bytes'] / 2**30:.0f} GiB")

    t0 = time.time()
    dt = sp.F32
    A = sp.gen_rmat(args.scale, 1 << args.nnz_log2, 0.57, 0.19, 0.19, dtype=dt, seed=2)
    x = sp.gen_vector(dt, A.ncols, seed=22)
    log(f"[bench] R-MAT 2^{args.scale}: nnz={A.nnz} generated in {time.time() - t0:.1f}s")

    plan = sp.Plan(ctx, "CSR-nnz", A, sp.opts(algo=sp.ALGO_SEGMENTED, items_per_lane=8))
    pinfo = plan.info()
    algo_bytes = pinfo["algo_bytes"]
    xd = sp.DeviceArray.from_numpy(ctx, x)
    yd = sp.DeviceArray(ctx, A.nrows, np.float32)
    flush = args.flush_mb << 20
    ctx.sync()

    # warmup (untimed)
    if args.warmup > 0:
        plan.time_spmv(xd.addr, yd.addr, args.warmup, flush)
    ctx.sync()
    t1 = time.perf_counter()
    ms, ms_dom = plan.time_spmv_split(xd.addr, yd.addr, args.steps, flush)
    ctx.sync()
    wall = time.perf_counter() - t1
    mean_ms = statistics.mean(ms)
    mean_dom = statistics.mean(ms_dom)
    value = algo_bytes / (mean_ms * 1e-3) / 1e9
    achieved = algo_bytes / (mean_dom * 1e-3) / 1e9
    log(f"[bench] SpMV mean {mean_ms * 1e3:.2f} us (median {statistics.median(ms) * 1e3:.2f}), "
        f"dominant kernel {mean_dom * 1e3:.2f} us, {value:.0f} GB/s, wall {wall:.2f}s")

    # warm-cache (MALL-resident) figure, reported separately
    ms_warm = plan.time_spmv(xd.addr, yd.addr, 20, 0)
    warm_gbs = algo_bytes / (statistics.median(ms_warm) * 1e-3) / 1e9

    # stream-copy probe: 1 GiB read + 1 GiB write
    cms = ctx.time_copy(1 << 30, 10)
    copy_gbs = 2 * (1 << 30) / (statistics.median(cms[2:]) * 1e-3) / 1e9

    # parity of the bench output against the oracle (float tolerance)
    import oracle_lib as orc

    y = yd.numpy()
    C = A.to_csr()
    yref, absb = orc.ref_csr(dt, C.rowptr, C.colind, C.values, x)
    tol = orc.fp_tolerance(dt, np.diff(C.rowptr.astype(np.int64)), absb)
    parity = bool(np.all(np.abs(y.astype(np.float64) - yref) <= tol))
    log(f"[bench] parity vs oracle: {parity}")

    cpu = None
    if not args.no_cpu:
        cms_cpu, iters, _ = orc.time_csr(dt, C.rowptr, C.colind, C.values, x, iters=30, warmup=2,
                                        omp_threads=0, budget_s=args.cpu_budget)
        cpu_gbs = algo_bytes / (cms_cpu * 1e-3) / 1e9
        cpu = {"value": round(cpu_gbs, 3), "unit": "GB/s", "cores": 1, "kind": "port",
               "sample": f"full matrix, median of {iters} runs of the C restatement of SparseP CSR "
                         f"spmv_host (spmv/1D/CSR-row/host/app.c:66-77), 1 thread, {cms_cpu:.1f} ms/run",
               "host_cpus": orc.num_cpus()}
        log(f"[bench] cpu baseline {cms_cpu:.1f} ms/run ({cpu_gbs:.2f} GB/s, {iters} runs)")

    out = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(mean_ms, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded R-MAT, project PRNG)",
        "config": {"workload": f"R-MAT 2^{args.scale} x 2^{args.scale}, {A.nnz} unique nnz, "
                               f"a=0.57 b=0.19 c=0.19, FP32 CSR SpMV (CSR-nnz, merge-path kernel), "
                               f"MALL flushed before each step",
                   "nrows": A.nrows, "nnz": A.nnz, "algo_bytes": algo_bytes,
                   "gflops": round(2 * A.nnz / (mean_ms * 1e-3) / 1e9, 2)},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None},
        "cpu_baseline": cpu,
        "extra": {"frac_of_peak_whole_spmv": round(value / HBM_PEAK_GBS, 4),
                  "measured_copy_gbs": round(copy_gbs, 1),
                  "frac_of_measured_copy": round(value / copy_gbs, 4),
                  "warm_cache_gbs": round(warm_gbs, 1),
                  "kernels_per_spmv": pinfo["launches"], "parity_vs_oracle": parity},
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
