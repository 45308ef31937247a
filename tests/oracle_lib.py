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
