"""Python mirror of the C ABI (include/sparsep.h) used by the tests and the
bench.  Every call goes through libsparsep.so; numpy arrays are views of
library-owned memory kept alive by their owning wrapper object.
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import check, lib, sp_dist, sp_opts, sp_plan_info

I8, I16, I32, I64, F32, F64 = range(6)
NP_DTYPE = {I8: np.int8, I16: np.int16, I32: np.int32, I64: np.int64, F32: np.float32, F64: np.float64}
DTYPE_OF = {np.dtype(v): k for k, v in NP_DTYPE.items()}
DTYPE_NAME = {I8: "int8", I16: "int16", I32: "int32", I64: "int64", F32: "fp32", F64: "fp64"}

# sp_kernel_id
KERNELS = ["CSR-row", "CSR-nnz", "COO-row", "COO-nnz-rgrn", "COO-nnz", "BCSR-block", "BCSR-nnz",
           "BCOO-block", "BCOO-nnz", "DCSR", "DCOO", "DBCSR", "DBCOO", "RBDCSR", "RBDCOO",
           "RBDBCSR-block", "RBDBCSR-nnz", "RBDBCOO-block", "RBDBCOO-nnz", "BDCSR", "BDCOO",
           "BDBCSR-block", "BDBCSR-nnz", "BDBCOO-block", "BDBCOO-nnz"]
KID = {n: i for i, n in enumerate(KERNELS)}

ALGO_AUTO, ALGO_VECTOR, ALGO_SEGMENTED, ALGO_SEGMENTED_SORTED, ALGO_SEGMENTED_HOT, ALGO_SEGMENTED_ATOMIC = 0, 1, 2, 3, 4, 5
BAL_ROW, BAL_NNZ, BAL_BLOCK = 0, 1, 2
TILE_D, TILE_RB, TILE_BD = 0, 1, 2
FMT_CSR, FMT_COO, FMT_BCSR, FMT_BCOO = 0, 1, 2, 3
MTX_PAD, MTX_COMPAT, MTX_SORT = 1, 2, 4
DIST_UNIFORM, DIST_NORMAL, DIST_CONST = 0, 1, 2


def dist(kind=DIST_UNIFORM, lo=0.0, hi=1.0):
    return sp_dist(kind, float(lo), float(hi))


def _u32(a):
    a = np.ascontiguousarray(a, dtype=np.uint32)
    return a, a.ctypes.data_as(_lib.u32p)


def _view(ptr, n, dtype):
    if n == 0 or not ptr:
        return np.zeros(0, dtype=dtype)
    addr = ctypes.cast(ptr, ctypes.c_void_p).value
    buf = (ctypes.c_char * (int(n) * np.dtype(dtype).itemsize)).from_address(addr)
    return np.frombuffer(buf, dtype=dtype, count=int(n))


class Coo:
    """Owning wrapper of an sp_coo* (freed with sp_free_coo)."""

    def __init__(self, ptr):
        self.ptr = ptr
        s = ptr.contents
        self.nrows, self.ncols, self.nnz, self.dtype = s.nrows, s.ncols, s.nnz, s.dtype
        self.rowind = _view(s.rowind, s.nnz, np.uint32)
        self.colind = _view(s.colind, s.nnz, np.uint32)
        self.values = _view(s.values, s.nnz, NP_DTYPE[s.dtype])
        self.rows = _view(s.rows, s.nrows, np.uint32)

    def __del__(self):
        if getattr(self, "ptr", None):
            lib().sp_free_coo(self.ptr)
            self.ptr = None

    @staticmethod
    def from_arrays(nrows, ncols, rowind, colind, values, sort=False):
        values = np.ascontiguousarray(values)
        dt = DTYPE_OF[values.dtype]
        nnz = len(values)
        out = ctypes.POINTER(_lib.sp_coo)()
        check(lib().sp_coo_alloc(nrows, ncols, nnz, dt, ctypes.byref(out)), "sp_coo_alloc")
        A = Coo(out)
        A.rowind[:] = rowind
        A.colind[:] = colind
        A.values[:] = values
        A.rows[:] = np.bincount(np.asarray(rowind, dtype=np.int64), minlength=nrows)[:nrows] if nnz else 0
        if sort:
            check(lib().sp_sort_coo(A.ptr), "sp_sort_coo")
            A.__init__(A.ptr)
        return A

    def to_csr(self):
        out = ctypes.POINTER(_lib.sp_csr)()
        check(lib().sp_coo2csr(self.ptr, ctypes.byref(out)), "sp_coo2csr")
        return Csr(out)

    def dense(self):
        D = np.zeros((self.nrows, self.ncols), dtype=np.float64)
        np.add.at(D, (self.rowind.astype(np.int64), self.colind.astype(np.int64)), self.values.astype(np.float64))
        return D


class Csr:
    def __init__(self, ptr):
        self.ptr = ptr
        s = ptr.contents
        self.nrows, self.ncols, self.nnz, self.dtype = s.nrows, s.ncols, s.nnz, s.dtype
        self.rowptr = _view(s.rowptr, s.nrows + 1, np.uint32)
        self.colind = _view(s.colind, s.nnz, np.uint32)
        self.values = _view(s.values, s.nnz, NP_DTYPE[s.dtype])

    def __del__(self):
        if getattr(self, "ptr", None):
            lib().sp_free_csr(self.ptr)
            self.ptr = None

    def to_bcsr(self, r=4, c=4, count_values=False, sort=True):
        out = ctypes.POINTER(_lib.sp_bcsr)()
        if sort:
            check(lib().sp_build_bcsr(self.ptr, r, c, int(count_values), ctypes.byref(out)), "sp_build_bcsr")
        else:
            check(lib().sp_csr2bcsr(self.ptr, r, c, ctypes.byref(out)), "sp_csr2bcsr")
        return Bcsr(out)


class Bcsr:
    def __init__(self, ptr):
        self.ptr = ptr
        self.refresh()

    def refresh(self):
        s = self.ptr.contents
        self.r, self.c = s.row_block_size, s.col_block_size
        self.nrows, self.ncols, self.nnz, self.dtype = s.nrows, s.ncols, s.nnz, s.dtype
        self.num_block_rows, self.num_block_cols, self.num_blocks = s.num_block_rows, s.num_block_cols, s.num_blocks
        self.num_rows_left = s.num_rows_left
        self.browptr = _view(s.browptr, s.num_block_rows + 1, np.uint32)
        self.bcolind = _view(s.bcolind, s.num_blocks, np.uint32)
        self.bval = _view(s.bval, s.num_blocks * self.r * self.c, NP_DTYPE[s.dtype])
        self.nnz_per_block = _view(s.nnz_per_block, s.num_blocks, np.uint32)

    def __del__(self):
        if getattr(self, "ptr", None):
            lib().sp_free_bcsr(self.ptr)
            self.ptr = None

    def sort(self):
        check(lib().sp_sort_bcsr(self.ptr), "sp_sort_bcsr")

    def count_nnz_per_block(self, csr=None, count_values=True):
        check(lib().sp_count_nnz_per_block(self.ptr, csr.ptr if csr is not None else None,
                                           int(count_values)), "sp_count_nnz_per_block")

    def to_bcoo(self):
        out = ctypes.POINTER(_lib.sp_bcoo)()
        check(lib().sp_bcsr2bcoo(self.ptr, ctypes.byref(out)), "sp_bcsr2bcoo")
        return Bcoo(out)


class Bcoo:
    def __init__(self, ptr):
        self.ptr = ptr
        s = ptr.contents
        self.r, self.c = s.row_block_size, s.col_block_size
        self.nrows, self.ncols, self.nnz, self.dtype = s.nrows, s.ncols, s.nnz, s.dtype
        self.num_blocks = s.num_blocks
        self.browind = _view(s.browind, s.num_blocks, np.uint32)
        self.bcolind = _view(s.bcolind, s.num_blocks, np.uint32)
        self.bval = _view(s.bval, s.num_blocks * self.r * self.c, NP_DTYPE[s.dtype])
        self.nnz_per_block = _view(s.nnz_per_block, s.num_blocks, np.uint32)

    def __del__(self):
        if getattr(self, "ptr", None):
            lib().sp_free_bcoo(self.ptr)
            self.ptr = None


class Tiled:
    def __init__(self, ptr):
        self.ptr = ptr
        s = ptr.contents
        self.s = s
        for f, _ in _lib.sp_tiled._fields_:
            v = getattr(s, f)
            if isinstance(v, (int, float)):
                setattr(self, f, v)
        nt, v = s.ntiles, s.vert_partitions
        self.vert_tile_widths = _view(s.vert_tile_widths, v + 1, np.uint32)
        self.nnzs_per_partition = _view(s.nnzs_per_partition, nt, np.uint32)
        self.nnz_offset = _view(s.nnz_offset, nt + 1, np.uint64)
        vt = NP_DTYPE[s.dtype]
        if s.format in (FMT_CSR, FMT_COO):
            self.dcolind = _view(s.dcolind, s.nnz, np.uint32)
            self.dval = _view(s.dval, s.nnz, vt)
            if s.format == FMT_CSR:
                self.drowptr = _view(s.drowptr, nt * (s.tile_height + 1), np.uint32)
            else:
                self.drowind = _view(s.drowind, s.nnz, np.uint32)
        else:
            nb, rc = s.num_blocks, s.row_block_size * s.col_block_size
            self.blocks_per_partition = _view(s.blocks_per_partition, nt, np.uint32)
            self.block_offset = _view(s.block_offset, nt + 1, np.uint64)
            self.bcolind = _view(s.bcolind, nb, np.uint32)
            self.bval = _view(s.bval, nb * rc, vt)
            self.nnz_per_block = _view(s.nnz_per_block, nb, np.uint32)
            if s.format == FMT_BCSR:
                self.browptr = _view(s.browptr, nt * (s.num_block_rows + 1), np.uint32)
            else:
                self.browind = _view(s.browind, nb, np.uint32)

    def __del__(self):
        if getattr(self, "ptr", None):
            lib().sp_free_tiled(self.ptr)
            self.ptr = None


def coo2tiled(A, tiling, fmt, horz, vert, r=4, c=4, count_values=False):
    out = ctypes.POINTER(_lib.sp_tiled)()
    check(lib().sp_coo2tiled(A.ptr, tiling, fmt, horz, vert, r, c, int(count_values), ctypes.byref(out)),
          "sp_coo2tiled")
    return Tiled(out)


# ---------------------------------------------------------------------------
# generators / loader
# ---------------------------------------------------------------------------

def _gen(fn, *args):
    out = ctypes.POINTER(_lib.sp_coo)()
    check(fn(*args, ctypes.byref(out)), fn.__name__)
    return Coo(out)


def gen_stencil5(gx, gy, dtype, offdiag=None, diag=4.0, seed=1):
    offdiag = offdiag or dist(DIST_UNIFORM, -8, 8)
    return _gen(lib().sp_gen_stencil5, gx, gy, dtype, offdiag, float(diag), seed)


def gen_rmat(scale, nnz, a=0.57, b=0.19, c=0.19, dtype=F32, vals=None, seed=2):
    vals = vals or default_vals(dtype)
    return _gen(lib().sp_gen_rmat, scale, nnz, a, b, c, dtype, vals, seed)


def gen_stencil27(g, dof=4, periodic=True, dtype=F32, seed=3):
    return _gen(lib().sp_gen_stencil27, g, dof, int(periodic), dtype, seed)


def gen_uniform(nrows, ncols, klo, khi, dtype=F32, vals=None, seed=4):
    return _gen(lib().sp_gen_uniform, nrows, ncols, klo, khi, dtype, vals or default_vals(dtype), seed)


def gen_banded(nrows, bw, klo, khi, dtype=F32, vals=None, seed=4):
    return _gen(lib().sp_gen_banded, nrows, bw, klo, khi, dtype, vals or default_vals(dtype), seed)


def gen_blockdiag(nrows, nblocks, klo, khi, dtype=F32, vals=None, seed=4):
    return _gen(lib().sp_gen_blockdiag, nrows, nblocks, klo, khi, dtype, vals or default_vals(dtype), seed)


def read_mtx(path, dtype, flags=0):
    out = ctypes.POINTER(_lib.sp_coo)()
    check(lib().sp_read_mtx(path.encode(), dtype, flags, ctypes.byref(out)), "sp_read_mtx")
    return Coo(out)


def default_vals(dtype):
    """Config value ranges: ints uniform in [-4,4], floats uniform in [-1,1]."""
    return dist(DIST_UNIFORM, -4, 4) if dtype <= I64 else dist(DIST_UNIFORM, -1, 1)


def gen_vector(dtype, n, d=None, seed=7):
    d = d or default_vals(dtype)
    out = np.empty(n, dtype=NP_DTYPE[dtype])
    check(lib().sp_gen_vector(dtype, n, d, seed, out.ctypes.data_as(ctypes.c_void_p)), "sp_gen_vector")
    return out


def init_vector(dtype, n):
    """The reference's compat input vector x[i] = i % 4 + 1."""
    out = np.empty(n, dtype=NP_DTYPE[dtype])
    check(lib().sp_init_vector(dtype, n, out.ctypes.data_as(ctypes.c_void_p)), "sp_init_vector")
    return out


# ---------------------------------------------------------------------------
# partitioners
# ---------------------------------------------------------------------------

def partition_by_row(nrows, nparts):
    out = np.zeros(nparts + 1, np.uint32)
    check(lib().sp_partition_by_row(nrows, nparts, out.ctypes.data_as(_lib.u32p)))
    return out


def partition_by_nnz(rowptr, nparts):
    rp, p = _u32(rowptr)
    out = np.zeros(nparts + 1, np.uint32)
    check(lib().sp_partition_by_nnz(p, len(rp) - 1, nparts, out.ctypes.data_as(_lib.u32p)))
    return out


def partition_tsklt_by_row(rows, ntasklets, gran):
    out = np.zeros(ntasklets + 1, np.uint32)
    check(lib().sp_partition_tsklt_by_row(rows, ntasklets, gran, out.ctypes.data_as(_lib.u32p)))
    return out


def partition_tsklt_by_nnz(rowptr, row_begin, rows, nnz, ntasklets, gran):
    rp, p = _u32(rowptr)
    out = np.zeros(ntasklets + 1, np.uint32)
    check(lib().sp_partition_tsklt_by_nnz(p, row_begin, rows, nnz, ntasklets, gran,
                                          out.ctypes.data_as(_lib.u32p)))
    return out


def partition_coo_nnz(rows_nnz, nnz, nparts):
    rn, p = _u32(rows_nnz)
    outs = [np.zeros(nparts, np.uint32) for _ in range(5)]
    check(lib().sp_partition_coo_nnz(p, len(rn), nnz, nparts, *[o.ctypes.data_as(_lib.u32p) for o in outs]))
    return dict(zip(["prev_rows", "rows_per_part", "prev_nnz", "nnz_per_part", "merge"], outs))


def partition_even(count, ntasklets):
    s, c = np.zeros(ntasklets, np.uint32), np.zeros(ntasklets, np.uint32)
    check(lib().sp_partition_even(count, ntasklets, s.ctypes.data_as(_lib.u32p), c.ctypes.data_as(_lib.u32p)))
    return s, c


def partition_bcsr(browptr, npb, nnz, nparts, by_nnz):
    bp, pb = _u32(browptr)
    nn, pn = _u32(npb)
    bs, bl, ns = np.zeros(nparts + 1, np.uint32), np.zeros(nparts, np.uint32), np.zeros(nparts, np.uint32)
    check(lib().sp_partition_bcsr(pb, pn, len(bp) - 1, len(nn), nnz, nparts, int(by_nnz),
                                  bs.ctypes.data_as(_lib.u32p), bl.ctypes.data_as(_lib.u32p),
                                  ns.ctypes.data_as(_lib.u32p)))
    return bs, bl, ns


def partition_tsklt_bcsr(browptr, npb, brow_split, blocks, nnzs, part, ntasklets, by_nnz):
    arrs = [_u32(a) for a in (browptr, npb, brow_split, blocks, nnzs)]
    out = np.zeros(ntasklets + 1, np.uint32)
    check(lib().sp_partition_tsklt_bcsr(*[a[1] for a in arrs], part, ntasklets, int(by_nnz),
                                        out.ctypes.data_as(_lib.u32p)))
    return out


def partition_bcoo(npb, nnz, nparts, by_nnz):
    nn, pn = _u32(npb)
    bs, ns = np.zeros(nparts + 1, np.uint32), np.zeros(nparts, np.uint32)
    check(lib().sp_partition_bcoo(pn, len(nn), nnz, nparts, int(by_nnz), bs.ctypes.data_as(_lib.u32p),
                                  ns.ctypes.data_as(_lib.u32p)))
    return bs, ns


def partition_tsklt_bcoo(npb, block_split, nnzs, part, ntasklets, by_nnz):
    arrs = [_u32(a) for a in (npb, block_split, nnzs)]
    out = np.zeros(ntasklets + 1, np.uint32)
    check(lib().sp_partition_tsklt_bcoo(*[a[1] for a in arrs], part, ntasklets, int(by_nnz),
                                        out.ctypes.data_as(_lib.u32p)))
    return out


def partition_tiled(T, horz, by_nnz=False):
    out = np.zeros(T.vert_partitions * (horz + 1), np.uint32)
    check(lib().sp_partition_tiled(T.ptr, horz, int(by_nnz), out.ctypes.data_as(_lib.u32p)))
    return out


# ---------------------------------------------------------------------------
# device path
# ---------------------------------------------------------------------------

def opts(**kw):
    o = sp_opts()
    lib().sp_opts_default(ctypes.byref(o))
    for k, v in kw.items():
        setattr(o, k, v)
    return o


class Context:
    def __init__(self, device=0):
        self.ptr = ctypes.c_void_p()
        check(lib().sp_ctx_create(device, ctypes.byref(self.ptr)), "sp_ctx_create")

    def __del__(self):
        if getattr(self, "ptr", None):
            lib().sp_ctx_destroy(self.ptr)
            self.ptr = None

    def sync(self):
        check(lib().sp_ctx_sync(self.ptr), "sp_ctx_sync")

    def set_stream(self, stream_handle):
        check(lib().sp_ctx_set_stream(self.ptr, ctypes.c_void_p(stream_handle)))

    def info(self):
        cus, lds = ctypes.c_int(), ctypes.c_int()
        hbm = ctypes.c_uint64()
        arch = ctypes.create_string_buffer(64)
        check(lib().sp_ctx_info(self.ptr, ctypes.byref(cus), ctypes.byref(lds), ctypes.byref(hbm), arch, 64))
        return {"num_cus": cus.value, "lds_per_block": lds.value, "hbm_bytes": hbm.value,
                "arch": arch.value.decode()}

    def time_copy(self, nbytes, iters=10):
        ms = (ctypes.c_float * iters)()
        check(lib().sp_time_copy(self.ptr, nbytes, iters, ms), "sp_time_copy")
        return list(ms)

    def time_read(self, nbytes, iters=10):
        """ms per read-only pass over nbytes of device memory (sp_time_read)."""
        ms = (ctypes.c_float * iters)()
        check(lib().sp_time_read(self.ptr, nbytes, iters, ms), "sp_time_read")
        return list(ms)


def gpu_present():
    """True when a GPU kernel driver node exists (does not initialise HIP)."""
    import os
    return os.path.exists("/dev/kfd")


def spmv_csr_noalloc(ctx, dtype, nrows, ncols, nnz, rowptr_ptr, colind_ptr, values_ptr, x_ptr, y_ptr):
    """sp_spmv_csr_noalloc: y = A x over CSR arrays already in device memory,
    without a plan and without allocating (fixed scratch, nrows + nnz <=
    NOALLOC_MAX_ITEMS).  Returns the status code instead of raising, so
    callers can check SP_ERR_RANGE."""
    return lib().sp_spmv_csr_noalloc(ctx.ptr, dtype, nrows, ncols, nnz, ctypes.c_void_p(rowptr_ptr),
                                     ctypes.c_void_p(colind_ptr), ctypes.c_void_p(values_ptr),
                                     ctypes.c_void_p(x_ptr), ctypes.c_void_p(y_ptr))


NOALLOC_MAX_ITEMS = 134217728
ERR_RANGE = -8


class DeviceArray:
    """Device buffer owned through the C ABI (sp_dmalloc / sp_dfree) with
    synchronous host copies on the context stream."""

    def __init__(self, ctx, shape, dtype):
        self.ctx = ctx
        self.shape = tuple(shape) if isinstance(shape, (tuple, list)) else (int(shape),)
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        self.ptr = ctypes.c_void_p()
        check(lib().sp_dmalloc(ctx.ptr, max(self.nbytes, 1), ctypes.byref(self.ptr)), "sp_dmalloc")

    @staticmethod
    def from_numpy(ctx, a):
        a = np.ascontiguousarray(a)
        d = DeviceArray(ctx, a.shape, a.dtype)
        d.upload(a)
        return d

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.nbytes == self.nbytes
        if self.nbytes:
            check(lib().sp_memcpy_h2d(self.ctx.ptr, self.ptr, a.ctypes.data_as(ctypes.c_void_p), self.nbytes))

    def fill_bytes(self, byte):
        self.upload(np.full(self.nbytes, byte, np.uint8).view(self.dtype).reshape(self.shape))

    def numpy(self):
        out = np.empty(self.shape, self.dtype)
        if self.nbytes:
            check(lib().sp_memcpy_d2h(self.ctx.ptr, out.ctypes.data_as(ctypes.c_void_p), self.ptr, self.nbytes))
        return out

    @property
    def addr(self):
        return self.ptr.value

    def __del__(self):
        if getattr(self, "ptr", None) and self.ptr.value:
            lib().sp_dfree(self.ctx.ptr, self.ptr)
            self.ptr = None


class Plan:
    """A device plan for one kernel variant (sp_plan_create)."""

    def __init__(self, ctx, kernel, A, o=None, csr=False):
        self.ctx = ctx
        kid = KID[kernel] if isinstance(kernel, str) else int(kernel)
        self.ptr = ctypes.c_void_p()
        o = o if o is not None else opts()
        if csr:
            check(lib().sp_plan_create_csr(ctx.ptr, kid, A.ptr, ctypes.byref(o), ctypes.byref(self.ptr)),
                  f"sp_plan_create_csr({KERNELS[kid]})")
        else:
            check(lib().sp_plan_create(ctx.ptr, kid, A.ptr, ctypes.byref(o), ctypes.byref(self.ptr)),
                  f"sp_plan_create({KERNELS[kid]})")
        self.kid = kid

    def __del__(self):
        if getattr(self, "ptr", None):
            lib().sp_plan_destroy(self.ptr)
            self.ptr = None

    def info(self):
        i = sp_plan_info()
        check(lib().sp_plan_get_info(self.ptr, ctypes.byref(i)))
        return {f: getattr(i, f) for f, _ in sp_plan_info._fields_}

    def spmv(self, x_ptr, y_ptr):
        check(lib().sp_spmv(self.ptr, ctypes.c_void_p(x_ptr), ctypes.c_void_p(y_ptr)), "sp_spmv")

    def spmm(self, k, X_ptr, ldx, Y_ptr, ldy):
        check(lib().sp_spmm(self.ptr, k, ctypes.c_void_p(X_ptr), ldx, ctypes.c_void_p(Y_ptr), ldy), "sp_spmm")

    def time_spmv(self, x_ptr, y_ptr, iters, flush_bytes=0):
        ms = (ctypes.c_float * iters)()
        check(lib().sp_time_spmv(self.ptr, ctypes.c_void_p(x_ptr), ctypes.c_void_p(y_ptr), iters,
                                 flush_bytes, ms), "sp_time_spmv")
        return list(ms)

    def time_spmv_split(self, x_ptr, y_ptr, iters, flush_bytes=0):
        ms = (ctypes.c_float * iters)()
        md = (ctypes.c_float * iters)()
        check(lib().sp_time_spmv_split(self.ptr, ctypes.c_void_p(x_ptr), ctypes.c_void_p(y_ptr), iters,
                                       flush_bytes, ms, md), "sp_time_spmv_split")
        return list(ms), list(md)

    def cg(self, b_ptr, x_ptr, maxit=1000, rtol=1e-8):
        """Conjugate gradient with this plan's matrix (sp_cg); x holds the start on entry."""
        it, res = ctypes.c_int(), ctypes.c_double()
        check(lib().sp_cg(self.ptr, ctypes.c_void_p(b_ptr), ctypes.c_void_p(x_ptr), maxit, rtol, ctypes.byref(it),
                          ctypes.byref(res)), "sp_cg")
        return it.value, res.value

    def time_spmm(self, k, X_ptr, ldx, Y_ptr, ldy, iters, flush_bytes=0):
        ms = (ctypes.c_float * iters)()
        check(lib().sp_time_spmm(self.ptr, k, ctypes.c_void_p(X_ptr), ldx, ctypes.c_void_p(Y_ptr), ldy,
                                 iters, flush_bytes, ms), "sp_time_spmm")
        return list(ms)


class Batch:
    """Batched SpMV/SpMM over several plans (sp_batch_create)."""

    GRAPH = 1

    def __init__(self, ctx, plans, k, graph=True):
        self.ctx = ctx
        self.plans = list(plans)
        n = len(self.plans)
        arr = (ctypes.c_void_p * n)(*[p.ptr.value for p in self.plans])
        ks = (ctypes.c_int * n)(*[int(v) for v in k])
        self.ptr = ctypes.c_void_p()
        check(lib().sp_batch_create(ctx.ptr, n, arr, ks, self.GRAPH if graph else 0, ctypes.byref(self.ptr)),
              "sp_batch_create")

    def __del__(self):
        if getattr(self, "ptr", None):
            lib().sp_batch_destroy(self.ptr)
            self.ptr = None

    @staticmethod
    def _tab(ptrs):
        return (ctypes.c_void_p * len(ptrs))(*ptrs)

    def run(self, xs, ys):
        check(lib().sp_batch_run(self.ptr, self._tab(xs), self._tab(ys)), "sp_batch_run")

    def time(self, xs, ys, iters, flush_bytes=0):
        ms = (ctypes.c_float * iters)()
        check(lib().sp_batch_time(self.ptr, self._tab(xs), self._tab(ys), iters, flush_bytes, ms), "sp_batch_time")
        return list(ms)

    def time_unbatched(self, xs, ys, k, iters, flush_bytes=0):
        """The same problems as one sp_spmm launch sequence per problem."""
        n = len(self.plans)
        arr = (ctypes.c_void_p * n)(*[p.ptr.value for p in self.plans])
        ks = (ctypes.c_int * n)(*[int(v) for v in k])
        ms = (ctypes.c_float * iters)()
        check(lib().sp_time_plans(self.ctx.ptr, n, arr, self._tab(xs), self._tab(ys), ks, iters, flush_bytes, ms),
              "sp_time_plans")
        return list(ms)

    def info(self):
        ng, ab = ctypes.c_uint32(), ctypes.c_uint64()
        check(lib().sp_batch_info(self.ptr, ctypes.byref(ng), ctypes.byref(ab)))
        return {"groups": ng.value, "algo_bytes": ab.value}


def tune_space(A, kernel_id=-1):
    """Candidate (kernel id, options) points of the exhaustive autotuner."""
    n = ctypes.c_int()
    check(lib().sp_tune_space(A.ptr, kernel_id, None, 0, ctypes.byref(n)))
    pts = (_lib.sp_tune_point * n.value)()
    check(lib().sp_tune_space(A.ptr, kernel_id, pts, n.value, ctypes.byref(n)))
    return pts


def tune(ctx, A, pts=None, iters=5, flush_bytes=0, k=1):
    """Build and time every point (SpMV, or SpMM with k right-hand sides);
    returns the point array (ms, status filled)."""
    pts = pts if pts is not None else tune_space(A)
    if k == 1:
        check(lib().sp_tune_run(ctx.ptr, A.ptr, pts, len(pts), iters, flush_bytes), "sp_tune_run")
    else:
        check(lib().sp_tune_run_spmm(ctx.ptr, A.ptr, pts, len(pts), iters, flush_bytes, k), "sp_tune_run_spmm")
    return pts


def tune_best(pts, kernel_id=-1):
    return lib().sp_tune_best(pts, len(pts), kernel_id)


def point_opts(pt):
    """sp_opts of a tune point (a copy)."""
    o = sp_opts()
    ctypes.memmove(ctypes.byref(o), ctypes.byref(pt.opts), ctypes.sizeof(sp_opts))
    return o


# --------------------------------------------------------------------------
# learned variant selector: features and the MLP
# --------------------------------------------------------------------------
NFEAT = 64


def features(A):
    """64-float structural feature vector of a Coo (sp_features)."""
    f = np.zeros(NFEAT, np.float32)
    check(lib().sp_features(A.ptr, f.ctypes.data_as(_lib.fp)), "sp_features")
    return f


def _fptr(a):
    return None if a is None else a.ctypes.data_as(_lib.fp)


def _f32(a, shape=None):
    if a is None:
        return None
    a = np.ascontiguousarray(a, dtype=np.float32)
    if shape is not None:
        a = a.reshape(shape)
    return a


class Mlp:
    """in -> h1 -> h2 -> out MLP trained and run by the library's HIP kernels
    (BF16 weights, FP32 activations and accumulation)."""

    def __init__(self, ctx, sizes=(64, 128, 128, 25), seed=5):
        self.ctx = ctx
        self.sizes = tuple(int(s) for s in sizes)
        self.ptr = ctypes.c_void_p()
        check(lib().sp_mlp_create(ctx.ptr, *self.sizes, seed, ctypes.byref(self.ptr)), "sp_mlp_create")
        self.nparams = lib().sp_mlp_num_params(self.ptr)

    def __del__(self):
        if getattr(self, "ptr", None) and self.ptr.value:
            lib().sp_mlp_destroy(self.ptr)
            self.ptr = None

    def get_params(self):
        p = np.empty(self.nparams, np.float32)
        check(lib().sp_mlp_get_params(self.ptr, _fptr(p), self.nparams))
        return p

    def set_params(self, p):
        p = _f32(p)
        assert p.size == self.nparams
        check(lib().sp_mlp_set_params(self.ptr, _fptr(p), self.nparams))

    def predict(self, X):
        X = _f32(X, (-1, self.sizes[0]))
        Y = np.empty((X.shape[0], self.sizes[-1]), np.float32)
        check(lib().sp_mlp_predict(self.ptr, _fptr(X), X.shape[0], _fptr(Y)), "sp_mlp_predict")
        return Y

    def loss_grad(self, X, T, M=None):
        X = _f32(X, (-1, self.sizes[0]))
        T = _f32(T, (X.shape[0], self.sizes[-1]))
        M = _f32(M, (X.shape[0], self.sizes[-1]))
        loss = ctypes.c_float()
        g = np.empty(self.nparams, np.float32)
        check(lib().sp_mlp_loss_grad(self.ptr, _fptr(X), _fptr(T), _fptr(M), X.shape[0],
                                     ctypes.byref(loss), _fptr(g)), "sp_mlp_loss_grad")
        return loss.value, g

    def adam(self, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
        check(lib().sp_mlp_adam(self.ptr, lr, b1, b2, eps))

    def train(self, X, T, M=None, epochs=100, batch=64, lr=1e-3, seed=5):
        X = _f32(X, (-1, self.sizes[0]))
        T = _f32(T, (X.shape[0], self.sizes[-1]))
        M = _f32(M, (X.shape[0], self.sizes[-1]))
        loss = ctypes.c_float()
        check(lib().sp_mlp_train(self.ptr, _fptr(X), _fptr(T), _fptr(M), X.shape[0], epochs, batch,
                                 lr, seed, ctypes.byref(loss)), "sp_mlp_train")
        return loss.value


class Selector:
    """Learned variant selector (sp_selector): load a model file or build one
    from trained MLP parameters and the feature standardisation."""

    def __init__(self, ctx, path=None, params=None, mean=None, std=None, margin=0.0,
                 sizes=(NFEAT, 128, 128, 25)):
        self.ctx = ctx
        self.ptr = ctypes.c_void_p()
        if path is not None:
            check(lib().sp_selector_load(ctx.ptr, str(path).encode(), ctypes.byref(self.ptr)),
                  f"sp_selector_load({path})")
        else:
            sz = (ctypes.c_int * 4)(*sizes)
            params, mean, std = _f32(params), _f32(mean), _f32(std)
            check(lib().sp_selector_create(ctx.ptr, sz, _fptr(params), params.size, _fptr(mean), _fptr(std),
                                           margin, ctypes.byref(self.ptr)), "sp_selector_create")

    def __del__(self):
        if getattr(self, "ptr", None) and self.ptr.value:
            lib().sp_selector_destroy(self.ptr)
            self.ptr = None

    def save(self, path):
        check(lib().sp_selector_save(self.ptr, str(path).encode()), "sp_selector_save")

    def select(self, A, fallback_k=0, iters=5, flush_bytes=0):
        """Returns (tune point, scores[25], fell_back)."""
        pt = _lib.sp_tune_point()
        scores = np.zeros(len(KERNELS), np.float32)
        fb = ctypes.c_int()
        check(lib().sp_select(self.ptr, A.ptr, fallback_k, iters, flush_bytes, ctypes.byref(pt), _fptr(scores),
                              ctypes.byref(fb)), "sp_select")
        return pt, scores, bool(fb.value)
