"""ctypes access to the test oracle (oracle/liboracle.so) and to the
reference-built harness libraries (oracle/_ref/libref_<dir>_<TYPE>.so).

Only tests, the smoke check and the bench's cpu_baseline leg import this.
"""
import ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE = os.path.join(ROOT, "oracle", "liboracle.so")
REFDIR = os.path.join(ROOT, "oracle", "_ref")

TYPE_NAME = {0: "INT8", 1: "INT16", 2: "INT32", 3: "INT64", 4: "FP32", 5: "FP64"}
NP = {0: np.int8, 1: np.int16, 2: np.int32, 3: np.int64, 4: np.float32, 5: np.float64}

_orc = None


def lib():
    global _orc
    if _orc is None:
        if not os.path.exists(ORACLE):
            raise ImportError(f"{ORACLE} missing: run make -C oracle")
        _orc = ctypes.CDLL(ORACLE)
        _orc.orc_time_csr.restype = ctypes.c_double
    return _orc


def P(a):
    return ctypes.c_void_p(a.ctypes.data) if a is not None else None


def spmv_csr(dt, rowptr, colind, vals, x):
    n = len(rowptr) - 1
    y = np.zeros(n, NP[dt])
    lib().orc_spmv_csr(dt, ctypes.c_uint32(n), P(rowptr), P(colind), P(vals), P(x), P(y))
    return y


def spmv_coo(dt, nrows, rowind, colind, vals, x):
    y = np.zeros(nrows, NP[dt])
    lib().orc_spmv_coo(dt, ctypes.c_uint32(nrows), ctypes.c_uint32(len(vals)), P(rowind), P(colind),
                       P(vals), P(x), P(y))
    return y


def spmv_bcsr(dt, nrows, ncols, R, C, browptr, bcolind, bval, x):
    y = np.zeros(nrows, NP[dt])
    lib().orc_spmv_bcsr(dt, ctypes.c_uint32(nrows), ctypes.c_uint32(ncols), ctypes.c_uint32(R),
                        ctypes.c_uint32(C), ctypes.c_uint32(len(browptr) - 1), P(browptr), P(bcolind),
                        P(bval), P(x), P(y))
    return y


def spmv_bcoo(dt, nrows, ncols, R, C, browind, bcolind, bval, x):
    y = np.zeros(nrows, NP[dt])
    lib().orc_spmv_bcoo(dt, ctypes.c_uint32(nrows), ctypes.c_uint32(ncols), ctypes.c_uint32(R),
                        ctypes.c_uint32(C), ctypes.c_uint64(len(browind)), P(browind), P(bcolind),
                        P(bval), P(x), P(y))
    return y


def spmv_tiled(dt, fmt, nrows, ncols, row_off, rows, col_off, col_w, ptr_off, ent_off, ptr, ri, ci,
               R, C, vals, x):
    y = np.zeros(nrows, NP[dt])
    u32 = lambda a: np.ascontiguousarray(a, np.uint32)
    u64 = lambda a: np.ascontiguousarray(a, np.uint64)
    arrs = [u32(row_off), u32(rows), u32(col_off), u32(col_w), u64(ptr_off), u64(ent_off),
            u32(ptr) if ptr is not None else np.zeros(1, np.uint32),
            u32(ri) if ri is not None else np.zeros(1, np.uint32), u32(ci)]
    lib().orc_spmv_tiled(dt, fmt, ctypes.c_uint32(nrows), ctypes.c_uint32(ncols),
                         ctypes.c_uint32(len(rows)), *[P(a) for a in arrs], ctypes.c_uint32(R),
                         ctypes.c_uint32(C), P(vals), P(x), P(y))
    return y


def merge_parts(dt, parts, nthreads=1):
    """parts: (nparts, n) array, merged in place into parts[0]."""
    parts = np.ascontiguousarray(parts)
    lib().orc_merge_parts(dt, ctypes.c_uint32(parts.shape[0]), ctypes.c_uint64(parts.shape[1]),
                          ctypes.c_uint64(parts.shape[1]), P(parts), nthreads)
    return parts[0].copy()


def ref_csr(dt, rowptr, colind, vals, x):
    """long-double reference y and the per-row sum of |a*x| (error bound)."""
    n = len(rowptr) - 1
    y = np.zeros(n, np.float64)
    b = np.zeros(n, np.float64)
    lib().orc_ref_csr(dt, ctypes.c_uint32(n), P(rowptr), P(colind), P(vals), P(x), P(y), P(b))
    return y, b


def fp_tolerance(dt, rowlen, absb):
    """Bound gamma(len+2)*sum|a x| for any summation order (with or without
    fused multiply-add) in the dtype's precision, plus a tiny absolute slack
    for rows whose exact sum underflows."""
    u = 2.0 ** -24 if dt == 4 else 2.0 ** -53
    m = rowlen.astype(np.float64) + 2.0
    gamma = m * u / (1.0 - m * u)
    return gamma * absb + (1e-30 if dt == 4 else 1e-300)


def time_csr(dt, rowptr, colind, vals, x, iters, warmup, omp_threads=0, budget_s=20.0):
    n = len(rowptr) - 1
    y = np.zeros(n, NP[dt])
    done = ctypes.c_int(0)
    ms = lib().orc_time_csr(dt, ctypes.c_uint32(n), P(rowptr), P(colind), P(vals), P(x), P(y), iters,
                            warmup, omp_threads, ctypes.c_double(budget_s), ctypes.byref(done))
    return ms, done.value, y


def tile_args(T, nrows):
    """The tile descriptors orc_spmv_tiled takes, read from a built tiling."""
    nt, vert = T.ntiles, T.vert_partitions
    cuts = T.vert_tile_widths.astype(np.int64)
    t = np.arange(nt)
    c = t % vert
    if T.tiling == 0:
        row_off = (t // vert) * T.tile_height
    else:
        row_off = np.zeros(nt, np.int64)
    rows = np.full(nt, T.tile_height if T.tiling == 0 else nrows, np.int64)
    col_off, col_w = cuts[c], cuts[c + 1] - cuts[c]
    if T.format in (0, 1):
        ptr_off = t * (T.tile_height + 1)
        ent_off, ci, vals = T.nnz_offset, T.dcolind, T.dval
        ptr = T.drowptr if T.format == 0 else None
        ri = T.drowind if T.format == 1 else None
        R = C = 1
    else:
        ptr_off = t * (T.num_block_rows + 1)
        ent_off, ci, vals = T.block_offset, T.bcolind, T.bval
        ptr = T.browptr if T.format == 2 else None
        ri = T.browind if T.format == 3 else None
        R, C = T.row_block_size, T.col_block_size
    return dict(row_off=row_off, rows=rows, col_off=col_off, col_w=col_w, ptr_off=ptr_off, ent_off=ent_off,
                ptr=ptr, ri=ri, ci=ci, R=R, C=C, vals=vals)


def _timed(fn, *args, iters, warmup, budget_s):
    done = ctypes.c_int(0)
    ms = fn(*args, iters, warmup, ctypes.c_double(budget_s), ctypes.byref(done))
    return ms, done.value


def time_coo(dt, nrows, rowind, colind, vals, x, iters=1000, warmup=2, budget_s=10.0):
    """Median ms of the COO spmv_host restatement (1 thread), runs done, and
    the y of the last run."""
    y = np.zeros(nrows, NP[dt])
    lib().orc_time_coo.restype = ctypes.c_double
    ms, done = _timed(lib().orc_time_coo, dt, ctypes.c_uint32(nrows), ctypes.c_uint32(len(vals)), P(rowind),
                  P(colind), P(vals), P(x), P(y), iters=iters, warmup=warmup, budget_s=budget_s)
    return ms, done, y


def time_bcsr(dt, nrows, ncols, R, C, browptr, bcolind, bval, x, iters=1000, warmup=2, budget_s=10.0):
    y = np.zeros(nrows, NP[dt])
    lib().orc_time_bcsr.restype = ctypes.c_double
    ms, done = _timed(lib().orc_time_bcsr, dt, ctypes.c_uint32(nrows), ctypes.c_uint32(ncols), ctypes.c_uint32(R),
                  ctypes.c_uint32(C), ctypes.c_uint32(len(browptr) - 1), P(browptr), P(bcolind), P(bval), P(x),
                  P(y), iters=iters, warmup=warmup, budget_s=budget_s)
    return ms, done, y


def time_bcoo(dt, nrows, ncols, R, C, browind, bcolind, bval, x, iters=1000, warmup=2, budget_s=10.0):
    y = np.zeros(nrows, NP[dt])
    lib().orc_time_bcoo.restype = ctypes.c_double
    ms, done = _timed(lib().orc_time_bcoo, dt, ctypes.c_uint32(nrows), ctypes.c_uint32(ncols), ctypes.c_uint32(R),
                  ctypes.c_uint32(C), ctypes.c_uint64(len(browind)), P(browind), P(bcolind), P(bval), P(x),
                  P(y), iters=iters, warmup=warmup, budget_s=budget_s)
    return ms, done, y


def time_tiled(dt, T, nrows, ncols, x, iters=1000, warmup=2, budget_s=10.0):
    """The 2D tile-major spmv_host restatement on a built tiling (1 thread)."""
    a = tile_args(T, nrows)
    u32 = lambda v: np.ascontiguousarray(v, np.uint32)
    u64 = lambda v: np.ascontiguousarray(v, np.uint64)
    arrs = [u32(a["row_off"]), u32(a["rows"]), u32(a["col_off"]), u32(a["col_w"]), u64(a["ptr_off"]),
            u64(a["ent_off"]), u32(a["ptr"]) if a["ptr"] is not None else np.zeros(1, np.uint32),
            u32(a["ri"]) if a["ri"] is not None else np.zeros(1, np.uint32), u32(a["ci"])]
    y = np.zeros(nrows, NP[dt])
    lib().orc_time_tiled.restype = ctypes.c_double
    ms, done = _timed(lib().orc_time_tiled, dt, T.format, ctypes.c_uint32(nrows), ctypes.c_uint32(ncols),
                  ctypes.c_uint32(len(arrs[1])), *[P(v) for v in arrs], ctypes.c_uint32(a["R"]),
                  ctypes.c_uint32(a["C"]), P(a["vals"]), P(x), P(y), iters=iters, warmup=warmup, budget_s=budget_s)
    return ms, done, y


def time_merge(dt, nparts, n, nthreads, iters=1000, warmup=2, budget_s=5.0):
    """The 2D vertical merge of nparts partial vectors of n values (OpenMP)."""
    parts = np.zeros((nparts, n), NP[dt])
    lib().orc_time_merge.restype = ctypes.c_double
    return _timed(lib().orc_time_merge, dt, ctypes.c_uint32(nparts), ctypes.c_uint64(n), ctypes.c_uint64(n),
                  P(parts), nthreads, iters=iters, warmup=warmup, budget_s=budget_s)


def num_cpus():
    return lib().orc_num_cpus()


# ---------------------------------------------------------------------------
# reference harness libraries
# ---------------------------------------------------------------------------
_refs = {}


def ref(dirname, dt):
    """The reference's own support code for kernel directory `dirname`,
    compiled for value type dt, or None if it was not built."""
    key = (dirname, dt)
    if key not in _refs:
        path = os.path.join(REFDIR, f"libref_{dirname}_{TYPE_NAME[dt]}.so")
        _refs[key] = ctypes.CDLL(path) if os.path.exists(path) else None
    return _refs[key]


def have_refs():
    return os.path.isdir(REFDIR) and len(os.listdir(REFDIR)) > 0
