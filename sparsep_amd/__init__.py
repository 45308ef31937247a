"""sparsep_amd: SpMV for AMD MI355X (gfx950) behind SparseP's host surface.

The product is the C ABI in include/sparsep.h, implemented by
sparsep_amd/libsparsep.so (C11 host surface + hand-written HIP kernels).
This package is the thin Python mirror of that ABI.
"""
from . import api  # noqa: F401
from ._lib import LIB_PATH, SparsePError, lib  # noqa: F401
