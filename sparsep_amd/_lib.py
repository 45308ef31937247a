"""ctypes binding of libsparsep.so (the C ABI declared in include/sparsep.h).

The library is built in-tree by ``__graft_entry__.build()`` (make -C
sparsep_amd/csrc).  Loading fails loudly when it is missing: there is no
fallback path.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libsparsep.so")

u32p = ctypes.POINTER(ctypes.c_uint32)
u64p = ctypes.POINTER(ctypes.c_uint64)
fp = ctypes.POINTER(ctypes.c_float)


class sp_coo(ctypes.Structure):
    _fields_ = [("nrows", ctypes.c_uint32), ("ncols", ctypes.c_uint32), ("nnz", ctypes.c_uint32),
                ("dtype", ctypes.c_int32), ("rowind", u32p), ("colind", u32p),
                ("values", ctypes.c_void_p), ("rows", u32p)]


class sp_csr(ctypes.Structure):
    _fields_ = [("nrows", ctypes.c_uint32), ("ncols", ctypes.c_uint32), ("nnz", ctypes.c_uint32),
                ("dtype", ctypes.c_int32), ("rowptr", u32p), ("colind", u32p),
                ("values", ctypes.c_void_p)]


class sp_bcsr(ctypes.Structure):
    _fields_ = [("row_block_size", ctypes.c_uint32), ("col_block_size", ctypes.c_uint32),
                ("num_block_rows", ctypes.c_uint32), ("num_block_cols", ctypes.c_uint32),
                ("num_blocks", ctypes.c_uint64), ("num_rows_left", ctypes.c_uint32),
                ("nrows", ctypes.c_uint32), ("ncols", ctypes.c_uint32), ("nnz", ctypes.c_uint32),
                ("dtype", ctypes.c_int32), ("browptr", u32p), ("bcolind", u32p),
                ("bval", ctypes.c_void_p), ("nnz_per_block", u32p)]


class sp_bcoo(ctypes.Structure):
    _fields_ = [("row_block_size", ctypes.c_uint32), ("col_block_size", ctypes.c_uint32),
                ("num_block_rows", ctypes.c_uint32), ("num_block_cols", ctypes.c_uint32),
                ("num_blocks", ctypes.c_uint64), ("num_rows_left", ctypes.c_uint32),
                ("nrows", ctypes.c_uint32), ("ncols", ctypes.c_uint32), ("nnz", ctypes.c_uint32),
                ("dtype", ctypes.c_int32), ("browind", u32p), ("bcolind", u32p),
                ("bval", ctypes.c_void_p), ("nnz_per_block", u32p)]


class sp_tiled(ctypes.Structure):
    _fields_ = [("tiling", ctypes.c_int32), ("format", ctypes.c_int32), ("dtype", ctypes.c_int32),
                ("nrows", ctypes.c_uint32), ("ncols", ctypes.c_uint32), ("nnz", ctypes.c_uint32),
                ("horz_partitions", ctypes.c_uint32), ("vert_partitions", ctypes.c_uint32),
                ("ntiles", ctypes.c_uint32), ("tile_height", ctypes.c_uint32),
                ("tile_width", ctypes.c_uint32), ("vert_tile_widths", u32p),
                ("nnzs_per_partition", u32p), ("nnz_offset", u64p), ("drowptr", u32p),
                ("drowind", u32p), ("dcolind", u32p), ("dval", ctypes.c_void_p),
                ("row_block_size", ctypes.c_uint32), ("col_block_size", ctypes.c_uint32),
                ("num_block_rows", ctypes.c_uint32), ("num_block_cols", ctypes.c_uint32),
                ("num_blocks", ctypes.c_uint64), ("blocks_per_partition", u32p),
                ("block_offset", u64p), ("browptr", u32p), ("browind", u32p), ("bcolind", u32p),
                ("bval", ctypes.c_void_p), ("nnz_per_block", u32p)]


class sp_dist(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int32), ("lo", ctypes.c_double), ("hi", ctypes.c_double)]


class sp_opts(ctypes.Structure):
    _fields_ = [("nparts", ctypes.c_uint32), ("nsub", ctypes.c_uint32),
                ("sub_balance", ctypes.c_int32), ("algo", ctypes.c_int32),
                ("lanes_per_row", ctypes.c_uint32), ("items_per_lane", ctypes.c_uint32),
                ("vert", ctypes.c_uint32), ("row_block_size", ctypes.c_uint32),
                ("col_block_size", ctypes.c_uint32), ("lds_slice", ctypes.c_uint32),
                ("count_values", ctypes.c_int32), ("mem_cap_bytes", ctypes.c_uint64)]


class sp_tune_point(ctypes.Structure):
    _fields_ = [("kernel_id", ctypes.c_int32), ("opts", sp_opts), ("ms", ctypes.c_float),
                ("build_ms", ctypes.c_float), ("status", ctypes.c_int32), ("algo_bytes", ctypes.c_uint64)]


class sp_plan_info(ctypes.Structure):
    _fields_ = [("kernel_id", ctypes.c_int32), ("dtype", ctypes.c_int32), ("algo", ctypes.c_int32),
                ("nrows", ctypes.c_uint32), ("ncols", ctypes.c_uint32), ("nnz", ctypes.c_uint32),
                ("num_blocks", ctypes.c_uint64), ("algo_bytes", ctypes.c_uint64),
                ("device_bytes", ctypes.c_uint64), ("nparts", ctypes.c_uint32),
                ("nsub", ctypes.c_uint32), ("lanes_per_row", ctypes.c_uint32),
                ("items_per_lane", ctypes.c_uint32), ("vert", ctypes.c_uint32),
                ("ntiles", ctypes.c_uint32), ("row_block_size", ctypes.c_uint32),
                ("col_block_size", ctypes.c_uint32), ("launches", ctypes.c_uint32)]


_lib = None

I, U, P, V, D, S = (ctypes.c_int, ctypes.c_uint32, ctypes.c_void_p, None, ctypes.c_double,
                    ctypes.c_size_t)
U64 = ctypes.c_uint64
PP = ctypes.POINTER(ctypes.c_void_p)

# name -> (restype, argtypes)
SIGNATURES = {
    "sp_strerror": (ctypes.c_char_p, [I]),
    "sp_dtype_size": (S, [I]),
    "sp_read_mtx": (I, [ctypes.c_char_p, I, I, ctypes.POINTER(ctypes.POINTER(sp_coo))]),
    "sp_rng_seed": (V, [P, U64]),
    "sp_rng_next": (U64, [P]),
    "sp_rng_uniform": (D, [P]),
    "sp_rng_normal": (D, [P]),
    "sp_gen_vector": (I, [I, U64, sp_dist, U64, P]),
    "sp_init_vector": (I, [I, U64, P]),
    "sp_gen_stencil5": (I, [U, U, I, sp_dist, D, U64, ctypes.POINTER(ctypes.POINTER(sp_coo))]),
    "sp_gen_rmat": (I, [U, U64, D, D, D, I, sp_dist, U64, ctypes.POINTER(ctypes.POINTER(sp_coo))]),
    "sp_gen_stencil27": (I, [U, U, I, I, U64, ctypes.POINTER(ctypes.POINTER(sp_coo))]),
    "sp_gen_uniform": (I, [U, U, U, U, I, sp_dist, U64, ctypes.POINTER(ctypes.POINTER(sp_coo))]),
    "sp_gen_banded": (I, [U, U, U, U, I, sp_dist, U64, ctypes.POINTER(ctypes.POINTER(sp_coo))]),
    "sp_gen_blockdiag": (I, [U, U, U, U, I, sp_dist, U64, ctypes.POINTER(ctypes.POINTER(sp_coo))]),
    "sp_coo_alloc": (I, [U, U, U, I, ctypes.POINTER(ctypes.POINTER(sp_coo))]),
    "sp_free_coo": (V, [ctypes.POINTER(sp_coo)]),
    "sp_free_csr": (V, [ctypes.POINTER(sp_csr)]),
    "sp_free_bcsr": (V, [ctypes.POINTER(sp_bcsr)]),
    "sp_free_bcoo": (V, [ctypes.POINTER(sp_bcoo)]),
    "sp_free_tiled": (V, [ctypes.POINTER(sp_tiled)]),
    "sp_sort_coo": (I, [ctypes.POINTER(sp_coo)]),
    "sp_coo2csr": (I, [ctypes.POINTER(sp_coo), ctypes.POINTER(ctypes.POINTER(sp_csr))]),
    "sp_csr2bcsr": (I, [ctypes.POINTER(sp_csr), U, U, ctypes.POINTER(ctypes.POINTER(sp_bcsr))]),
    "sp_sort_bcsr": (I, [ctypes.POINTER(sp_bcsr)]),
    "sp_count_nnz_per_block": (I, [ctypes.POINTER(sp_bcsr), ctypes.POINTER(sp_csr), I]),
    "sp_bcsr2bcoo": (I, [ctypes.POINTER(sp_bcsr), ctypes.POINTER(ctypes.POINTER(sp_bcoo))]),
    "sp_build_bcsr": (I, [ctypes.POINTER(sp_csr), U, U, I, ctypes.POINTER(ctypes.POINTER(sp_bcsr))]),
    "sp_coo2tiled": (I, [ctypes.POINTER(sp_coo), I, I, U, U, U, U, I,
                         ctypes.POINTER(ctypes.POINTER(sp_tiled))]),
    "sp_partition_by_row": (I, [U, U, u32p]),
    "sp_partition_by_nnz": (I, [u32p, U, U, u32p]),
    "sp_partition_tsklt_by_row": (I, [U, U, U, u32p]),
    "sp_partition_tsklt_by_nnz": (I, [u32p, U, U, U, U, U, u32p]),
    "sp_partition_coo_nnz": (I, [u32p, U, U, U, u32p, u32p, u32p, u32p, u32p]),
    "sp_partition_even": (I, [U, U, u32p, u32p]),
    "sp_partition_bcsr": (I, [u32p, u32p, U, U64, U, U, I, u32p, u32p, u32p]),
    "sp_partition_tsklt_bcsr": (I, [u32p, u32p, u32p, u32p, u32p, U, U, I, u32p]),
    "sp_partition_bcoo": (I, [u32p, U64, U, U, I, u32p, u32p]),
    "sp_partition_tsklt_bcoo": (I, [u32p, u32p, u32p, U, U, I, u32p]),
    "sp_find_partitions": (U, [U, U]),
    "sp_partition_tiled": (I, [ctypes.POINTER(sp_tiled), U, I, u32p]),
    "sp_kernel_name": (ctypes.c_char_p, [I]),
    "sp_opts_default": (V, [ctypes.POINTER(sp_opts)]),
    "sp_ctx_create": (I, [I, PP]),
    "sp_ctx_destroy": (I, [P]),
    "sp_ctx_set_stream": (I, [P, P]),
    "sp_ctx_stream": (P, [P]),
    "sp_ctx_sync": (I, [P]),
    "sp_ctx_info": (I, [P, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int), u64p,
                        ctypes.c_char_p, S]),
    "sp_dmalloc": (I, [P, S, PP]),
    "sp_dfree": (I, [P, P]),
    "sp_memcpy_h2d": (I, [P, P, P, S]),
    "sp_memcpy_d2h": (I, [P, P, P, S]),
    "sp_plan_create": (I, [P, I, ctypes.POINTER(sp_coo), ctypes.POINTER(sp_opts), PP]),
    "sp_plan_create_csr": (I, [P, I, ctypes.POINTER(sp_csr), ctypes.POINTER(sp_opts), PP]),
    "sp_plan_destroy": (I, [P]),
    "sp_plan_get_info": (I, [P, ctypes.POINTER(sp_plan_info)]),
    "sp_spmv": (I, [P, P, P]),
    "sp_spmm": (I, [P, I, P, I, P, I]),
    "sp_spmv_csr_noalloc": (I, [P, I, U, U, U, P, P, P, P, P]),
    "sp_time_spmv": (I, [P, P, P, I, U64, ctypes.POINTER(ctypes.c_float)]),
    "sp_time_spmm": (I, [P, I, P, I, P, I, I, U64, ctypes.POINTER(ctypes.c_float)]),
    "sp_time_spmv_split": (I, [P, P, P, I, U64, ctypes.POINTER(ctypes.c_float),
                               ctypes.POINTER(ctypes.c_float)]),
    "sp_time_copy": (I, [P, U64, I, ctypes.POINTER(ctypes.c_float)]),
    "sp_time_read": (I, [P, U64, I, ctypes.POINTER(ctypes.c_float)]),
    "sp_tune_space": (I, [ctypes.POINTER(sp_coo), I, ctypes.POINTER(sp_tune_point), I, ctypes.POINTER(ctypes.c_int)]),
    "sp_tune_run": (I, [P, ctypes.POINTER(sp_coo), ctypes.POINTER(sp_tune_point), I, I, U64]),
    "sp_tune_run_spmm": (I, [P, ctypes.POINTER(sp_coo), ctypes.POINTER(sp_tune_point), I, I, U64, I]),
    "sp_tune_best": (I, [ctypes.POINTER(sp_tune_point), I, I]),
    "sp_tune_db_append": (I, [ctypes.c_char_p, ctypes.c_char_p, ctypes.POINTER(sp_tune_point), I]),
    "sp_batch_create": (I, [P, I, PP, ctypes.POINTER(ctypes.c_int), I, PP]),
    "sp_batch_run": (I, [P, PP, PP]),
    "sp_batch_time": (I, [P, PP, PP, I, U64, ctypes.POINTER(ctypes.c_float)]),
    "sp_batch_info": (I, [P, ctypes.POINTER(ctypes.c_uint32), u64p]),
    "sp_batch_destroy": (I, [P]),
    "sp_batch_run_once": (I, [P, I, PP, PP, PP, ctypes.POINTER(ctypes.c_int)]),
    "sp_time_plans": (I, [P, I, PP, PP, PP, ctypes.POINTER(ctypes.c_int), I, U64, fp]),
    "sp_cg": (I, [P, P, P, I, D, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_double)]),
    "sp_features": (I, [ctypes.POINTER(sp_coo), fp]),
    "sp_mlp_create": (I, [P, I, I, I, I, U64, PP]),
    "sp_mlp_destroy": (I, [P]),
    "sp_mlp_num_params": (S, [P]),
    "sp_mlp_set_params": (I, [P, fp, S]),
    "sp_mlp_get_params": (I, [P, fp, S]),
    "sp_mlp_predict": (I, [P, fp, I, fp]),
    "sp_mlp_loss_grad": (I, [P, fp, fp, fp, I, fp, fp]),
    "sp_mlp_adam": (I, [P, ctypes.c_float, ctypes.c_float, ctypes.c_float, ctypes.c_float]),
    "sp_mlp_train": (I, [P, fp, fp, fp, I, I, I, ctypes.c_float, U64, fp]),
    "sp_selector_create": (I, [P, ctypes.POINTER(ctypes.c_int), fp, S, fp, fp, ctypes.c_float, PP]),
    "sp_selector_load": (I, [P, ctypes.c_char_p, PP]),
    "sp_selector_save": (I, [P, ctypes.c_char_p]),
    "sp_selector_destroy": (I, [P]),
    "sp_select": (I, [P, ctypes.POINTER(sp_coo), I, I, U64, ctypes.POINTER(sp_tune_point), fp,
                      ctypes.POINTER(ctypes.c_int)]),
}


def lib():
    """Load libsparsep.so once; raise if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'`"
            " (make -C sparsep_amd/csrc). There is no fallback implementation.")
    L = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(L, name, None)
        if fn is None:
            continue
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


class SparsePError(RuntimeError):
    def __init__(self, rc, what=""):
        msg = lib().sp_strerror(rc).decode()
        super().__init__(f"{what}: {msg} (rc={rc})" if what else f"{msg} (rc={rc})")
        self.rc = rc


def check(rc, what=""):
    if rc != 0:
        raise SparsePError(rc, what)
    return rc
