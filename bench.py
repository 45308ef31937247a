#!/usr/bin/env python3
"""SpMV bench on one MI355X.

Default workload (--workload rmat): the seeded R-MAT matrix of 2^20 x 2^20
with 2^24 unique entries (quadrants a=0.57, b=0.19, c=0.19; values and x
uniform in [-1,1]; seed 2), FP32, CSR with the entry-balanced (merge-path)
kernel.  One step = one SpMV y = A x with A, x, y resident in HBM.  Before
every step a 1 GiB buffer is overwritten so that L2 and the 256 MB MALL hold
nothing of A or x (the 147 MB working set would otherwise be served from
MALL); the flush is outside the timed events.

value = algorithmic bytes of one SpMV (nnz*(4+4) + (n+1)*4 + (m+n)*4,
x and y counted once) / mean SpMV time measured with HIP events on the
library's stream.  roofline.achieved uses the duration of the dominant
(matrix-streaming) kernel alone.  cpu_baseline times the oracle's C
restatement of SparseP's CSR spmv_host (single thread, reference order) on
the same matrix and vector, on this host.

Other workloads print one JSON line each in the same format:
  stencil5   256x256 5-point Laplacian, INT32, CSR-row, bit-exact check
  formats    the R-MAT matrix in every 1D format and every dtype
  stencil27  64^3 x 4-DOF 27-point stencil, 2D kernels, FP32 and FP64
  batched    256 uneven SpMV/SpMM problems, one persistent hipGraph launch
             against one launch per problem, FP32 and INT16
  selector   learned variant choice against a live exhaustive search on
             held-out corpus matrices
"""
import argparse
import json
import math
import os
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "SpMV GB/s & fraction of HBM peak; learned-pick vs exhaustive-best ratio"
HBM_PEAK_GBS = 8000.0
DT_NAME = {0: "i8", 1: "i16", 2: "i32", 3: "i64", 4: "f32", 5: "f64"}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def line(args, value, ms, dtype, config, unit="GB/s", roofline=None, cpu=None, extra=None, data=None,
         higher=True):
    return {
        "metric": METRIC, "value": round(value, 4), "unit": unit, "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms, 6), "higher_is_better": higher, "scaling": "weak",
        "vs_baseline": None, "dtype": dtype, "data": data or "synthetic (seeded generators, project PRNG)",
        "config": config, "roofline": roofline, "cpu_baseline": cpu, "extra": extra or {},
    }


def measure(sp, plan, xd, yd, args, flush):
    """Untimed warm-up, then `steps` timed SpMVs with the dominant kernel split."""
    if args.warmup > 0:
        plan.time_spmv(xd.addr, yd.addr, args.warmup, flush)
    ms, md = plan.time_spmv_split(xd.addr, yd.addr, args.steps, flush)
    return statistics.mean(ms), statistics.mean(md)


def parity_ok(orc, dt, A, x, y):
    import numpy as np
    C = A.to_csr()
    if dt <= 3:
        return bool(np.array_equal(y, orc.spmv_csr(dt, C.rowptr, C.colind, C.values, x)))
    yref, absb = orc.ref_csr(dt, C.rowptr, C.colind, C.values, x)
    tol = orc.fp_tolerance(dt, np.diff(C.rowptr.astype(np.int64)), absb)
    return bool(np.all(np.abs(y.astype(np.float64) - yref) <= tol))


# The reference CPU path of each kernel id: which spmv_host restatement it is
# timed with (spmv/1D/CSR-row/host/app.c:66-77, COO-row:70-74, BCSR-block:70-85,
# BCOO-block:72-86, the 2D tile-major copies such as spmv/2D/DCSR/host/app.c:84-103)
# and, for the 2D ids, the OpenMP vertical merge (spmv/2D/DCSR/host/app.c:340-351).
_CPU_SRC = {0: "spmv/1D/CSR-row/host/app.c:66-77", 1: "spmv/1D/COO-row/host/app.c:70-74",
            2: "spmv/1D/BCSR-block/host/app.c:70-85", 3: "spmv/1D/BCOO-block/host/app.c:72-86"}
_TILED = {9: (0, 0), 10: (0, 1), 11: (0, 2), 12: (0, 3), 13: (1, 0), 14: (1, 1), 15: (1, 2), 16: (1, 2),
          17: (1, 3), 18: (1, 3), 19: (2, 0), 20: (2, 1), 21: (2, 2), 22: (2, 2), 23: (2, 3), 24: (2, 3)}


def cpu_baseline_for(sp, orc, kid, A, x, info, budget_s):
    """Times the reference CPU path of kernel id `kid` on the same matrix and x:
    1 thread, the restated spmv_host of its format; 2D ids also report the host
    merge of the partial vectors on all cores.  Bounded by budget_s seconds."""
    dt = A.dtype
    r, c = info["row_block_size"] or 4, info["col_block_size"] or 4
    merge = None
    if kid <= 1:
        C = A.to_csr()
        ms, done, _ = orc.time_csr(dt, C.rowptr, C.colind, C.values, x, iters=1000, warmup=1, omp_threads=0,
                                   budget_s=budget_s)
        src = _CPU_SRC[0]
    elif kid <= 4:
        ms, done, _ = orc.time_coo(dt, A.nrows, A.rowind, A.colind, A.values, x, warmup=1, budget_s=budget_s)
        src = _CPU_SRC[1]
    elif kid <= 8:
        B = A.to_csr().to_bcsr(r, c)
        if kid <= 6:
            ms, done, _ = orc.time_bcsr(dt, A.nrows, A.ncols, r, c, B.browptr, B.bcolind, B.bval, x, warmup=1,
                                        budget_s=budget_s)
            src = _CPU_SRC[2]
        else:
            O = B.to_bcoo()
            ms, done, _ = orc.time_bcoo(dt, A.nrows, A.ncols, r, c, O.browind, O.bcolind, O.bval, x, warmup=1,
                                        budget_s=budget_s)
            src = _CPU_SRC[3]
    else:
        tiling, fmt = _TILED[kid]
        vert = max(1, info["vert"])
        horz = max(1, info["nparts"] // vert)
        T = sp.coo2tiled(A, tiling, fmt, horz, vert, r, c)
        ms, done, _ = orc.time_tiled(dt, T, A.nrows, A.ncols, x, warmup=1, budget_s=budget_s)
        src = f"the 2D tile-major spmv_host of spmv/2D/{sp.KERNELS[kid]}/host/app.c"
        mms, mdone = orc.time_merge(dt, vert, A.nrows, orc.num_cpus(), warmup=1, budget_s=min(2.0, budget_s))
        merge = {"ms": round(mms, 3), "threads": orc.num_cpus(), "parts": vert,
                 "source": "spmv/2D/DCSR/host/app.c:340-351 (OpenMP vertical merge)"}
    cpu = {"value": round(info["algo_bytes"] / (ms * 1e-3) / 1e9, 3), "unit": "GB/s", "cores": 1, "kind": "port",
           "sample": f"full matrix, median of {done} runs of the C restatement of {src}, 1 thread, {ms:.1f} ms/run",
           "host_cpus": orc.num_cpus()}
    if merge:
        cpu["host_merge"] = merge
    return cpu


# --------------------------------------------------------------------------
def wl_rmat(sp, ctx, args, orc):
    import numpy as np
    t0 = time.time()
    dt = sp.F32
    A = sp.gen_rmat(args.scale, 1 << args.nnz_log2, 0.57, 0.19, 0.19, dtype=dt, seed=2)
    x = sp.gen_vector(dt, A.ncols, seed=22)
    log(f"[bench] R-MAT 2^{args.scale}: nnz={A.nnz} generated in {time.time() - t0:.1f}s")
    if args.algo < 0:
        # the library's own autotuner picks the CSR-nnz configuration (outside the timed region)
        pts = sp.tune(ctx, A, sp.tune_space(A, sp.KID["CSR-nnz"]), iters=3, flush_bytes=args.flush_mb << 20)
        best = sp.tune_best(pts)
        o = sp.point_opts(pts[best]) if best >= 0 else sp.opts(algo=sp.ALGO_SEGMENTED, items_per_lane=8)
        chosen = "chosen by the autotuner"
    else:
        o = sp.opts(algo=args.algo, items_per_lane=args.ipt)
        chosen = "fixed by --algo/--ipt"
    plan = sp.Plan(ctx, "CSR-nnz", A, o)
    pinfo = plan.info()
    algo_name = {sp.ALGO_SEGMENTED: f"merge-path, {pinfo['items_per_lane']} items per lane",
                 sp.ALGO_SEGMENTED_SORTED: f"merge-path with column-ordered tiles, {pinfo['items_per_lane']} "
                                           "items per lane",
                 sp.ALGO_SEGMENTED_HOT: "merge-path with column-ordered tiles and the most used x entries "
                                        "staged in LDS"}.get(pinfo["algo"],
                                                                 f"vector, {pinfo['lanes_per_row']} lanes per row")
    log(f"[bench] CSR-nnz configuration ({chosen}): {algo_name}")
    algo_bytes = pinfo["algo_bytes"]
    xd = sp.DeviceArray.from_numpy(ctx, x)
    yd = sp.DeviceArray(ctx, A.nrows, np.float32)
    flush = args.flush_mb << 20
    ctx.sync()
    t1 = time.perf_counter()
    mean_ms, mean_dom = measure(sp, plan, xd, yd, args, flush)
    wall = time.perf_counter() - t1
    value = algo_bytes / (mean_ms * 1e-3) / 1e9
    achieved = algo_bytes / (mean_dom * 1e-3) / 1e9
    log(f"[bench] SpMV mean {mean_ms * 1e3:.2f} us, dominant kernel {mean_dom * 1e3:.2f} us, "
        f"{value:.0f} GB/s, wall {wall:.2f}s")
    ms_warm = plan.time_spmv(xd.addr, yd.addr, 20, 0)
    warm_gbs = algo_bytes / (statistics.median(ms_warm) * 1e-3) / 1e9
    cms = ctx.time_copy(1 << 30, 10)
    copy_gbs = 2 * (1 << 30) / (statistics.median(cms[2:]) * 1e-3) / 1e9
    rms = ctx.time_read(1 << 30, 10)
    read_gbs = (1 << 30) / (statistics.median(rms[2:]) * 1e-3) / 1e9
    log(f"[bench] measured ceilings: copy {copy_gbs:.0f} GB/s, read-only stream {read_gbs:.0f} GB/s")
    parity = parity_ok(orc, dt, A, x, yd.numpy())
    log(f"[bench] parity vs oracle: {parity}")
    cpu = None
    if not args.no_cpu:
        C = A.to_csr()
        # about args.cpu_budget seconds of single-thread CPU work over the full matrix
        cms_cpu, iters, _ = orc.time_csr(dt, C.rowptr, C.colind, C.values, x, iters=5000, warmup=2,
                                        omp_threads=0, budget_s=args.cpu_budget)
        cpu_gbs = algo_bytes / (cms_cpu * 1e-3) / 1e9
        cpu = {"value": round(cpu_gbs, 3), "unit": "GB/s", "cores": 1, "kind": "port",
               "sample": f"full matrix, median of {iters} runs of the C restatement of SparseP CSR "
                         f"spmv_host (spmv/1D/CSR-row/host/app.c:66-77), 1 thread, {cms_cpu:.1f} ms/run",
               "host_cpus": orc.num_cpus()}
        log(f"[bench] cpu baseline {cms_cpu:.1f} ms/run ({cpu_gbs:.2f} GB/s, {iters} runs)")
    cfg = {"workload": f"R-MAT 2^{args.scale} x 2^{args.scale}, {A.nnz} unique nnz, a=0.57 b=0.19 c=0.19, "
                       f"FP32 CSR SpMV (CSR-nnz, {algo_name}, {chosen}), "
                       f"MALL flushed before each step",
           "nrows": A.nrows, "nnz": A.nnz, "algo_bytes": algo_bytes,
           "gflops": round(2 * A.nnz / (mean_ms * 1e-3) / 1e9, 2)}
    # HBM bytes per launch of this kernel configuration from the committed --pmc runs
    # (FETCH_SIZE / WRITE_SIZE scaled by a 1 GiB copy kernel of the same runs)
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "csr_merge_rmat20_f32_traffic.json")
    tkey = {sp.ALGO_SEGMENTED: "merge", sp.ALGO_SEGMENTED_SORTED: "sorted",
            sp.ALGO_SEGMENTED_HOT: "hot"}.get(pinfo["algo"])
    if os.path.exists(tpath) and tkey and args.scale == 20 and args.nnz_log2 == 24:
        with open(tpath) as f:
            t = json.load(f)["per_config"].get(f"{tkey}{pinfo['items_per_lane']}")
        traffic = t["traffic_bytes"] if t else None
    roof = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic}
    report = os.path.join(ROOT, "sparsep_amd", "data", "selector_report.json")
    sel = None
    if os.path.exists(report):
        with open(report) as f:
            ho = json.load(f)["held_out"]
        sel = {"source": "sparsep_amd/data/selector_report.json (held-out corpus, measured when trained)",
               "learned_pick_over_exhaustive_best_geomean": ho["learned"]["geomean_ratio"],
               "with_fallback_geomean": ho["learned_with_fallback"]["geomean_ratio"],
               "selection_overhead_median": ho["overhead"]["select_over_exhaustive_median"]}
    extra = {"selector": sel,
             "frac_of_peak_whole_spmv": round(value / HBM_PEAK_GBS, 4), "measured_copy_gbs": round(copy_gbs, 1),
             "frac_of_measured_copy": round(value / copy_gbs, 4), "measured_read_gbs": round(read_gbs, 1),
             "frac_of_measured_read": round(value / read_gbs, 4), "warm_cache_gbs": round(warm_gbs, 1),
             "kernels_per_spmv": pinfo["launches"], "parity_vs_oracle": parity}
    return [line(args, value, mean_ms, "f32", cfg, roofline=roof, cpu=cpu, extra=extra,
                 data="synthetic (seeded R-MAT, project PRNG)")]


def wl_stencil5(sp, ctx, args, orc):
    import numpy as np
    dt = sp.I32
    A = sp.gen_stencil5(256, 256, dt, offdiag=sp.dist(sp.DIST_UNIFORM, -8, 8), diag=4.0, seed=1)
    x = sp.gen_vector(dt, A.ncols, d=sp.dist(sp.DIST_UNIFORM, -4, 4), seed=1)
    plan = sp.Plan(ctx, "CSR-row", A)
    algo = plan.info()["algo_bytes"]
    xd = sp.DeviceArray.from_numpy(ctx, x)
    yd = sp.DeviceArray(ctx, A.nrows, np.int32)
    mean_ms, mean_dom = measure(sp, plan, xd, yd, args, 0)
    exact = parity_ok(orc, dt, A, x, yd.numpy())
    cpu = None
    if not args.no_cpu:
        C = A.to_csr()
        cms, iters, _ = orc.time_csr(dt, C.rowptr, C.colind, C.values, x, iters=200, warmup=5, omp_threads=0,
                                    budget_s=5.0)
        cpu = {"value": round(cms * 1e3, 2), "unit": "us", "cores": 1, "kind": "port",
               "sample": f"full matrix, median of {iters} runs of the C restatement of CSR spmv_host"}
    cfg = {"workload": "5-point Laplacian 256x256 (65,536 rows, 326,656 nnz), INT32 CSR-row SpMV, "
                       "offdiag U{-8..8}, diag 4, x U{-4..4}, seed 1; cache-resident (no flush)",
           "nrows": A.nrows, "nnz": A.nnz, "algo_bytes": algo}
    return [line(args, mean_ms * 1e3, mean_ms, "i32", cfg, unit="us per SpMV", cpu=cpu, higher=False,
                 extra={"bit_exact_vs_oracle": exact, "dominant_kernel_us": round(mean_dom * 1e3, 2),
                        "gbs_cache_resident": round(algo / (mean_ms * 1e-3) / 1e9, 1)})]


def _rmat_all(sp, dt):
    vals = sp.dist(sp.DIST_UNIFORM, -1, 1) if dt >= 4 else sp.dist(sp.DIST_UNIFORM, -4, 4)
    return sp.gen_rmat(20, 1 << 24, 0.57, 0.19, 0.19, dtype=dt, vals=vals, seed=2), vals


def wl_formats(sp, ctx, args, orc):
    import numpy as np
    out = []
    # (kernel, algo): default plans, then the atomic merges (algo 5) of COO-nnz and BCOO-nnz
    kinds = [(k, 0) for k in ("CSR-row", "CSR-nnz", "COO-row", "COO-nnz-rgrn", "COO-nnz", "BCSR-block", "BCSR-nnz",
                              "BCOO-block", "BCOO-nnz")] + [("COO-nnz", 5), ("BCOO-nnz", 5)]
    for dt in args.dtypes:
        A, vals = _rmat_all(sp, dt)
        x = sp.gen_vector(dt, A.ncols, d=vals, seed=22)
        xd = sp.DeviceArray.from_numpy(ctx, x)
        yd = sp.DeviceArray(ctx, A.nrows, x.dtype)
        for kbase, algo in kinds:
            kname = kbase + (" atomic" if algo else "")
            plan = sp.Plan(ctx, kbase, A, sp.opts(algo=algo))
            info = plan.info()
            mean_ms, mean_dom = measure(sp, plan, xd, yd, args, args.flush_mb << 20)
            ok = parity_ok(orc, dt, A, x, yd.numpy())
            gbs = info["algo_bytes"] / (mean_dom * 1e-3) / 1e9
            cpu = None if args.no_cpu else cpu_baseline_for(sp, orc, sp.KID[kbase], A, x, info, args.cpu_budget_each)
            log(f"[bench] formats {DT_NAME[dt]} {kname:13s} {mean_ms * 1e3:8.2f} us  {gbs:7.0f} GB/s  parity {ok}"
                + (f"  cpu {cpu['value']:.2f} GB/s" if cpu else ""))
            cfg = {"workload": f"R-MAT 2^20, 2^24 nnz, {kname}, MALL flushed", "kernel": kbase, "algo": info["algo"],
                   "nnz": A.nnz, "algo_bytes": info["algo_bytes"]}
            out.append(line(args, info["algo_bytes"] / (mean_ms * 1e-3) / 1e9, mean_ms, DT_NAME[dt], cfg,
                            roofline={"bound": "hbm", "achieved": round(gbs, 2), "peak": HBM_PEAK_GBS,
                                      "unit": "GB/s", "frac": round(gbs / HBM_PEAK_GBS, 4), "traffic": None},
                            cpu=cpu, extra={"parity_vs_oracle": ok, "launches": info["launches"]}))
            del plan
    return out


def wl_stencil27(sp, ctx, args, orc):
    out = []
    # (kernel, block edge); the 4 x 4 blocks match the stencil's 4 unknowns per grid point
    kinds = [("CSR-nnz", 2), ("BCSR-block", 2), ("BCSR-block", 4), ("BCOO-block", 4), ("DCSR", 2), ("DCOO", 2),
             ("RBDCSR", 2), ("RBDBCSR-block", 2), ("BDCSR", 2), ("BDBCOO-block", 2), ("BDBCOO-block", 4)]
    for dt in [d for d in args.dtypes if d in (sp.F32, sp.F64)]:
        A = sp.gen_stencil27(64, 4, True, dt, seed=3)
        x = sp.gen_vector(dt, A.ncols, d=sp.dist(sp.DIST_NORMAL, 0, 1), seed=33)
        xd = sp.DeviceArray.from_numpy(ctx, x)
        yd = sp.DeviceArray(ctx, A.nrows, x.dtype)
        for kname, be in kinds:
            t0 = time.time()
            kw = dict(row_block_size=be, col_block_size=be)
            if kname[0] in "DRB" and kname[1] != "C":
                kw["vert"] = args.vert
            plan = sp.Plan(ctx, kname, A, sp.opts(**kw))
            info = plan.info()
            build = time.time() - t0
            mean_ms, mean_dom = measure(sp, plan, xd, yd, args, args.flush_mb << 20)
            ok = parity_ok(orc, dt, A, x, yd.numpy())
            gbs = info["algo_bytes"] / (mean_dom * 1e-3) / 1e9
            cpu = None if args.no_cpu else cpu_baseline_for(sp, orc, sp.KID[kname], A, x, info, args.cpu_budget_each)
            tag = f"{kname} {be}x{be}" if "B" in kname.replace("BD", "") else kname
            log(f"[bench] stencil27 {DT_NAME[dt]} {tag:17s} {mean_ms * 1e3:8.2f} us {gbs:7.0f} GB/s "
                f"(dominant) parity {ok} build {build:.1f}s")
            cfg = {"workload": f"64^3 x 4 DOF 27-point stencil, periodic, {tag}", "kernel": kname, "nnz": A.nnz,
                   "block": [be, be], "algo_bytes": info["algo_bytes"], "vert": info.get("vert", 0)}
            out.append(line(args, info["algo_bytes"] / (mean_ms * 1e-3) / 1e9, mean_ms, DT_NAME[dt], cfg,
                            roofline={"bound": "hbm", "achieved": round(gbs, 2), "peak": HBM_PEAK_GBS,
                                      "unit": "GB/s", "frac": round(gbs / HBM_PEAK_GBS, 4), "traffic": None},
                            cpu=cpu,
                            extra={"parity_vs_oracle": ok, "launches": info["launches"], "build_s": round(build, 1)}))
            del plan
    return out


def wl_rmat2d(sp, ctx, args, orc):
    """The R-MAT FP32 matrix under the 2D tilings, over a range of vertical splits."""
    import numpy as np
    out = []
    A, vals = _rmat_all(sp, sp.F32)
    x = sp.gen_vector(sp.F32, A.ncols, d=vals, seed=22)
    xd = sp.DeviceArray.from_numpy(ctx, x)
    yd = sp.DeviceArray(ctx, A.nrows, np.float32)
    for kname in ("RBDCSR", "BDCSR", "BDCOO", "DCSR"):
        # 128 stripes of 8192 columns: the widest x slice that fits 32 KB of LDS, with and without it
        for vert, sl in ((2, 0), (4, 0), (8, 0), (16, 0), (32, 0), (128, 0), (128, 8192)):
            try:
                plan = sp.Plan(ctx, kname, A, sp.opts(vert=vert, lds_slice=sl))
            except sp._lib.SparsePError as e:
                log(f"[bench] rmat2d {kname:8s} vert {vert:3d} slice {sl}: not built ({e})")
                continue
            info = plan.info()
            mean_ms, mean_dom = measure(sp, plan, xd, yd, args, args.flush_mb << 20)
            ok = parity_ok(orc, sp.F32, A, x, yd.numpy())
            csr_bytes = 146800644
            log(f"[bench] rmat2d {kname:8s} vert {vert:3d} slice {sl:4d} {mean_ms * 1e3:8.2f} us (dominant "
                f"{mean_dom * 1e3:8.2f}) CSR-equivalent {csr_bytes / (mean_ms * 1e-3) / 1e9:6.0f} GB/s parity {ok}")
            cfg = {"workload": f"R-MAT 2^20 FP32, {kname} vert {vert}" + (f", x slices in LDS" if sl else ""),
                   "kernel": kname, "vert": vert, "lds_slice": sl, "algo_bytes": info["algo_bytes"]}
            out.append(line(args, info["algo_bytes"] / (mean_ms * 1e-3) / 1e9, mean_ms, "f32", cfg,
                            extra={"parity_vs_oracle": ok, "launches": info["launches"],
                                   "csr_equivalent_gbs": round(csr_bytes / (mean_ms * 1e-3) / 1e9, 1)}))
            del plan
    return out


def wl_phases(sp, ctx, args, orc):
    """SparseP's five-phase report (spmv/1D/CSR-row/host/app.c:291-300) for CSR-row on the R-MAT
    FP32 matrix, or on --mtx: host-to-device matrix + partition, x upload, kernel, y download,
    merge (on the GPU inside the kernel phase here).  PCIe-inclusive, host wall clock."""
    import numpy as np
    dt = sp.F32
    if args.mtx:
        A = sp.read_mtx(args.mtx, dt, sp.MTX_SORT)
        name = os.path.basename(args.mtx)
    else:
        A, vals = _rmat_all(sp, dt)
        name = "R-MAT 2^20, 2^24 nnz"
    x = sp.gen_vector(dt, A.ncols, seed=22)
    rows = []
    for rep in range(min(args.steps, 5)):
        t0 = time.perf_counter()
        plan = sp.Plan(ctx, "CSR-row", A)
        ctx.sync()
        t1 = time.perf_counter()
        xd = sp.DeviceArray.from_numpy(ctx, x)
        t2 = time.perf_counter()
        yd = sp.DeviceArray(ctx, A.nrows, np.float32)
        t2b = time.perf_counter()
        plan.spmv(xd.addr, yd.addr)
        ctx.sync()
        t3 = time.perf_counter()
        y = yd.numpy()
        t4 = time.perf_counter()
        rows.append((t1 - t0, t2 - t1, t3 - t2b, t4 - t3))
        del plan
    med = [statistics.median(r[i] for r in rows) * 1e3 for i in range(4)]
    ok = parity_ok(orc, dt, A, x, y)
    info_b = 146800644 if not args.mtx else None
    log(f"[bench] phases {name}: Load Matrix {med[0]:.2f} ms, Load Input Vector {med[1]:.3f} ms, "
        f"Kernel {med[2]:.3f} ms, Retrieve Output Vector {med[3]:.3f} ms, Merge 0 (on GPU), parity {ok}")
    total = sum(med)
    cfg = {"workload": f"five-phase end-to-end CSR-row SpMV, {name}, FP32, host buffers in and out "
                       f"(PCIe-inclusive, wall clock)", "nnz": A.nnz}
    return [line(args, total, total, "f32", cfg, unit="ms end to end", higher=False,
                 extra={"load_matrix_ms": round(med[0], 3), "load_input_vector_ms": round(med[1], 4),
                        "kernel_ms": round(med[2], 4), "retrieve_output_vector_ms": round(med[3], 4),
                        "merge_ms": 0.0, "parity_vs_oracle": ok,
                        "kernel_plus_transfers_gbs": round((info_b or 0) / ((med[1] + med[2] + med[3]) * 1e-3) / 1e9, 1)
                        if info_b else None})]


def wl_all25(sp, ctx, args, orc):
    """Every kernel id with its default configuration on the 27-point stencil (64^3 x 4 DOF) in
    the first of --dtypes that is FP32 or FP64 (default FP32)."""
    import numpy as np
    out = []
    dt = next((d for d in args.dtypes if d in (sp.F32, sp.F64)), sp.F32)
    A = sp.gen_stencil27(64, 4, True, dt, seed=3)
    x = sp.gen_vector(dt, A.ncols, d=sp.dist(sp.DIST_NORMAL, 0, 1), seed=33)
    xd = sp.DeviceArray.from_numpy(ctx, x)
    yd = sp.DeviceArray(ctx, A.nrows, x.dtype)
    csr_bytes = None
    for kid, kname in enumerate(sp.KERNELS):
        t0 = time.time()
        plan = sp.Plan(ctx, kid, A)
        info = plan.info()
        build = time.time() - t0
        mean_ms, mean_dom = measure(sp, plan, xd, yd, args, args.flush_mb << 20)
        ok = parity_ok(orc, dt, A, x, yd.numpy())
        cpu = None if args.no_cpu else cpu_baseline_for(sp, orc, kid, A, x, info, args.cpu_budget_each)
        if kid == 0:
            csr_bytes = info["algo_bytes"]
        gbs = info["algo_bytes"] / (mean_ms * 1e-3) / 1e9
        log(f"[bench] all25 {DT_NAME[dt]} {kname:14s} {mean_ms * 1e3:8.2f} us {gbs:7.0f} GB/s "
            f"(CSR-equivalent {csr_bytes / (mean_ms * 1e-3) / 1e9:6.0f}) parity {ok} build {build:.1f}s")
        cfg = {"workload": f"64^3 x 4 DOF 27-point stencil, {kname}, default configuration", "kernel": kname,
               "nnz": A.nnz, "algo_bytes": info["algo_bytes"], "vert": info["vert"]}
        out.append(line(args, gbs, mean_ms, DT_NAME[dt], cfg,
                        roofline={"bound": "hbm", "achieved": round(info["algo_bytes"] / (mean_dom * 1e-3) / 1e9, 2),
                                  "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                  "frac": round(info["algo_bytes"] / (mean_dom * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                                  "traffic": None},
                        cpu=cpu,
                        extra={"parity_vs_oracle": ok, "launches": info["launches"], "build_s": round(build, 1),
                               "csr_equivalent_gbs": round(csr_bytes / (mean_ms * 1e-3) / 1e9, 1)}))
        del plan
    return out


def wl_batched(sp, ctx, args, orc):
    import numpy as np
    out = []
    rng = np.random.default_rng(4)
    nprob = args.nprob
    specs = []
    for i in range(nprob):
        rows = int(math.exp(rng.uniform(math.log(1000), math.log(200000))))
        specs.append((i % 3, rows, [1, 4, 8, 32][i % 4], 4000 + i))
    for dt in (sp.F32, sp.I16):
        mats, plans, ks = [], [], []
        tune_s = 0.0
        for n_done, (kind, rows, k, seed) in enumerate(specs):
            if n_done % 32 == 0:
                log(f"[bench] batched {DT_NAME[dt]}: generating and tuning problem {n_done}/{nprob}")
            if kind == 0:
                A = sp.gen_uniform(rows, rows, 2, 64, dtype=dt, seed=seed)
            elif kind == 1:
                A = sp.gen_banded(rows, 32, 2, 64, dtype=dt, seed=seed)
            else:
                scale = max(10, int(round(math.log2(rows))))
                A = sp.gen_rmat(scale, (1 << scale) * 33, dtype=dt, seed=seed)
            t0 = time.perf_counter()
            pts = sp.tune_space(A, sp.KID["CSR-nnz"])
            sp.tune(ctx, A, pts, iters=3, k=k)  # timed with the problem's own number of columns
            best = sp.tune_best(pts)
            tune_s += time.perf_counter() - t0
            o = sp.point_opts(pts[best]) if best >= 0 else None
            mats.append(A)
            plans.append(sp.Plan(ctx, "CSR-nnz", A, o))
            ks.append(k)
        Xs, Xd, Yd = [], [], []
        for i, A in enumerate(mats):
            if dt >= 4:
                X = np.random.default_rng(i).standard_normal((A.ncols, ks[i])).astype(np.float32)
            else:
                X = np.random.default_rng(i).integers(-4, 5, (A.ncols, ks[i])).astype(np.int16)
            Xs.append(X)
            Xd.append(sp.DeviceArray.from_numpy(ctx, X))
            Yd.append(sp.DeviceArray(ctx, (A.nrows, ks[i]), X.dtype))
        xs, ys = [x.addr for x in Xd], [y.addr for y in Yd]
        b = sp.Batch(ctx, plans, ks, graph=True)
        binfo = b.info()
        flush = args.flush_mb << 20
        b.time(xs, ys, 3, flush)
        tb = statistics.mean(b.time(xs, ys, args.steps, flush))
        b.time_unbatched(xs, ys, ks, 2, flush)
        tu = statistics.mean(b.time_unbatched(xs, ys, ks, args.steps, flush))
        # parity of the batched result against the oracle
        b.run(xs, ys)
        ctx.sync()
        ok = True
        for i, A in enumerate(mats[:: max(1, nprob // 32)]):
            j = i * max(1, nprob // 32)
            Y = Yd[j].numpy()
            for q in range(ks[j]):
                ok &= parity_ok(orc, dt, A, np.ascontiguousarray(Xs[j][:, q]), np.ascontiguousarray(Y[:, q]))
        sv = 4 if dt == sp.F32 else 2
        algo = sum(p.info()["algo_bytes"] - (A.nrows + A.ncols) * sv + k * (A.nrows + A.ncols) * sv
                   for p, A, k in zip(plans, mats, ks))
        nnz = sum(A.nnz for A in mats)
        gbs_b, gbs_u = algo / (tb * 1e-3) / 1e9, algo / (tu * 1e-3) / 1e9
        log(f"[bench] batched {DT_NAME[dt]}: {nprob} problems, {nnz} nnz, graph {tb:.3f} ms ({gbs_b:.0f} GB/s), "
            f"per-problem launches {tu:.3f} ms ({gbs_u:.0f} GB/s), groups {binfo['groups']}, parity {ok}")
        cfg = {"workload": f"{nprob} problems (uniform / banded bw 32 / R-MAT), rows log-U[1k,200k], "
                           f"nnz/row 2..64, k in {{1,4,8,32}}, one persistent hipGraph launch per kind group",
               "problems": nprob, "nnz_total": nnz, "algo_bytes": algo, "groups": binfo["groups"]}
        out.append(line(args, gbs_b, tb, DT_NAME[dt], cfg,
                        roofline={"bound": "hbm", "achieved": round(gbs_b, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                  "frac": round(gbs_b / HBM_PEAK_GBS, 4), "traffic": None},
                        extra={"per_problem_launch_ms": round(tu, 4), "per_problem_gbs": round(gbs_u, 1),
                               "speedup_vs_per_problem": round(tu / tb, 3), "parity_vs_oracle": bool(ok),
                               "autotune_s": round(tune_s, 2)}))
    return out


def wl_selector(sp, ctx, args, orc):
    import numpy as np
    from sparsep_amd import selector as S
    model = os.path.join(ROOT, "sparsep_amd", "data", "selector_f32.bin")
    sel = sp.Selector(ctx, path=model)
    specs = S.corpus_specs(args.sel_count * 5, seed=args.sel_seed, rows=(10_000, 2_000_000), nnz=(50_000, 8_000_000))
    specs = specs[4::5][:args.sel_count]  # every fifth spec of an unseen seed
    ratios, retimed, sel_ms, tune_ms = [], [], [], []
    for s in specs:
        A = S.build_matrix(s)
        t0 = time.perf_counter()
        pt, scores, _ = sel.select(A, fallback_k=0)
        sel_ms.append((time.perf_counter() - t0) * 1e3)
        t0 = time.perf_counter()
        pts = sp.tune(ctx, A, iters=5, flush_bytes=512 << 20)
        tune_ms.append((time.perf_counter() - t0) * 1e3)
        ok = [p.ms for p in pts if p.status == 0 and p.ms > 0]
        # the pick runs its default configuration: the first point of its id
        first = next((p for p in pts if p.kernel_id == pt.kernel_id), None)
        good = first is not None and first.status == 0 and first.ms > 0
        ratios.append(min(ok) / first.ms if good else 0.0)
        # the minimum over ~80 noisy medians is biased low: re-time the exhaustive winner and the pick
        # with fresh plans and 20 flushed runs each
        bi = sp.tune_best(pts)
        x = sp.DeviceArray.from_numpy(ctx, sp.gen_vector(sp.F32, A.ncols, seed=5))
        y = sp.DeviceArray(ctx, A.nrows, np.float32)
        t_b = statistics.median(sp.Plan(ctx, pts[bi].kernel_id, A, sp.point_opts(pts[bi])).time_spmv(
            x.addr, y.addr, 20, 512 << 20))
        t_p = statistics.median(sp.Plan(ctx, pt.kernel_id, A).time_spmv(x.addr, y.addr, 20, 512 << 20))
        retimed.append(t_b / t_p)
        log(f"[bench] selector {s['family']} nnz={A.nnz} pick {sp.KERNELS[pt.kernel_id]} ratio {ratios[-1]:.3f} "
            f"re-timed {retimed[-1]:.3f}")
    r = np.array(ratios)
    geo = float(np.exp(np.mean(np.log(np.maximum(r, 1e-9)))))
    cfg = {"workload": f"learned variant selector on {len(specs)} unseen synthetic matrices (seed {args.sel_seed}), "
                       f"FP32, rows 10k-2M, nnz 50k-8M; pick = default configuration of the predicted id, "
                       f"best = live exhaustive autotuner", "matrices": len(specs)}
    return [line(args, geo, statistics.mean(sel_ms), "f32", cfg, unit="learned-pick / exhaustive-best throughput",
                 extra={"within_10pct": float(np.mean(r >= 1 / 1.1)), "min_ratio": float(r.min()),
                        "retimed_geomean": float(np.exp(np.mean(np.log(np.maximum(np.array(retimed), 1e-9))))),
                        "retimed_within_10pct": float(np.mean(np.array(retimed) >= 1 / 1.1)),
                        "select_ms_median": float(np.median(sel_ms)), "tune_ms_median": float(np.median(tune_ms)),
                        "overhead_median": float(np.median(np.array(sel_ms) / np.array(tune_ms)))})]


def wl_spmm(sp, ctx, args, orc):
    """Several right-hand sides: Y = A X on the 27-point stencil (64^3 x 4 DOF), CSR and BCSR 4 x 4,
    FP32 and FP64, k in {8, 32}.  Algorithmic bytes: the matrix once plus k(m+n)s."""
    import numpy as np
    out = []
    for dt in [d for d in args.dtypes if d in (sp.F32, sp.F64)]:
        A = sp.gen_stencil27(64, 4, True, dt, seed=3)
        s = 4 if dt == sp.F32 else 8
        for k in (8, 32):
            X = np.stack([sp.gen_vector(dt, A.ncols, d=sp.dist(sp.DIST_NORMAL, 0, 1), seed=70 + q)
                          for q in range(k)], axis=1)
            Xd = sp.DeviceArray.from_numpy(ctx, X)
            Yd = sp.DeviceArray(ctx, (A.nrows, k), X.dtype)
            # CSR: the automatic choice (column-ordered merge tiles, 8 columns per pass) and the
            # row-granular plan (lanes across up to 32 columns)
            for kname, be, algo in (("CSR-nnz", 2, 0), ("CSR-nnz", 2, 1), ("BCSR-block", 4, 0)):
                plan = sp.Plan(ctx, kname, A, sp.opts(row_block_size=be, col_block_size=be, algo=algo))
                info = plan.info()
                if args.warmup > 0:
                    plan.time_spmm(k, Xd.addr, k, Yd.addr, k, args.warmup, args.flush_mb << 20)
                ms = statistics.mean(plan.time_spmm(k, Xd.addr, k, Yd.addr, k, args.steps, args.flush_mb << 20))
                Y = Yd.numpy()
                ok = all(parity_ok(orc, dt, A, np.ascontiguousarray(X[:, q]), np.ascontiguousarray(Y[:, q]))
                         for q in (0, k - 1))
                nbytes = info["algo_bytes"] + (k - 1) * (A.nrows + A.ncols) * s
                gbs = nbytes / (ms * 1e-3) / 1e9
                tag = f"{kname} {be}x{be}" if kname.startswith("B") else kname + (" rows" if algo == 1 else "")
                log(f"[bench] spmm {DT_NAME[dt]} k={k:2d} {tag:14s} {ms * 1e3:8.1f} us {gbs:7.0f} GB/s "
                    f"{2 * A.nnz * k / (ms * 1e-3) / 1e12:6.2f} TFLOP/s parity(first,last column) {ok}")
                cfg = {"workload": f"64^3 x 4 DOF 27-point stencil, Y = A X with {k} right-hand sides, {tag}",
                       "kernel": kname, "k": k, "nnz": A.nnz, "algo_bytes": nbytes}
                out.append(line(args, gbs, ms, DT_NAME[dt], cfg,
                                roofline={"bound": "hbm", "achieved": round(gbs, 2), "peak": HBM_PEAK_GBS,
                                          "unit": "GB/s", "frac": round(gbs / HBM_PEAK_GBS, 4), "traffic": None},
                                extra={"parity_vs_oracle": ok,
                                       "tflops": round(2 * A.nnz * k / (ms * 1e-3) / 1e12, 3)}))
                del plan
            del Xd, Yd
    return out


WORKLOADS = {"rmat": wl_rmat, "stencil5": wl_stencil5, "formats": wl_formats, "stencil27": wl_stencil27,
             "rmat2d": wl_rmat2d, "phases": wl_phases, "all25": wl_all25,
             "batched": wl_batched, "selector": wl_selector, "spmm": wl_spmm}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--workload", default="rmat", choices=sorted(WORKLOADS))
    ap.add_argument("--flush-mb", type=int, default=1024)
    ap.add_argument("--cpu-budget", type=float, default=12.0, help="seconds of CPU baseline timing")
    ap.add_argument("--cpu-budget-each", type=float, default=3.0,
                    help="formats / stencil27 / all25: seconds of CPU baseline timing per variant")
    ap.add_argument("--scale", type=int, default=20)
    ap.add_argument("--nnz-log2", type=int, default=24)
    ap.add_argument("--vert", type=int, default=0, help="stencil27: vertical partitions of the 2D kernels")
    ap.add_argument("--nprob", type=int, default=256, help="batched: number of problems")
    ap.add_argument("--sel-count", type=int, default=40, help="selector: matrices evaluated")
    ap.add_argument("--sel-seed", type=int, default=55)
    ap.add_argument("--dtypes", default=None,
                    help="formats / stencil27: comma-separated dtypes (i8,i16,i32,i64,f32,f64); default all")
    ap.add_argument("--mtx", default=None, help="phases: a Matrix Market file instead of the R-MAT matrix")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--algo", type=int, default=-1,
                    help="rmat: fixed CSR algorithm (2 merge path, 3 column-ordered merge tiles) instead of the "
                         "autotuner's pick, for profiler runs")
    ap.add_argument("--ipt", type=int, default=16, help="rmat: items per lane with --algo")
    args = ap.parse_args()

    names = {v: k for k, v in DT_NAME.items()}
    args.dtypes = [names[d] for d in args.dtypes.split(",")] if args.dtypes else list(range(6))
    if args.gpus != 1:
        print(f"bench.py measures one GPU only: this build targets a single MI355X and has no "
              f"multi-GPU path (asked for --gpus {args.gpus}).", file=sys.stderr)
        sys.exit(2)

    import sparsep_amd.api as sp

    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as orc

    ctx = sp.Context(0)
    info = ctx.info()
    log(f"[bench] device {info['arch']} CUs={info['num_cus']} HBM={info['hbm_bytes'] / 2**30:.0f} GiB "
        f"workload {args.workload}")
    for rec in WORKLOADS[args.workload](sp, ctx, args, orc):
        print(json.dumps(rec), flush=True)


if __name__ == "__main__":
    main()
