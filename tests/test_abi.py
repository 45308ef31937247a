"""The C-ABI library loads and exports every function include/sparsep.h declares."""
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    src = open(os.path.join(ROOT, "include", "sparsep.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    names = set(re.findall(r"\b(sp_[a-z0-9_]+)\s*\(", src))
    return sorted(names)


def test_library_exports_all_declared_symbols(sp):
    L = sp.lib()
    missing = [n for n in declared_functions() if not hasattr(L, n)]
    assert not missing, f"declared but not exported: {missing}"
    assert len(declared_functions()) > 50


def test_strerror_and_sizes(sp):
    L = sp.lib()
    assert L.sp_strerror(0) == b"ok"
    assert [L.sp_dtype_size(d) for d in range(6)] == [1, 2, 4, 8, 4, 8]
    assert L.sp_dtype_size(9) == 0


def test_kernel_names(sp):
    assert [sp.lib().sp_kernel_name(i).decode() for i in range(25)] == sp.KERNELS


def test_ctx_without_gpu_reports_error(sp):
    if sp.gpu_present():
        return
    import ctypes
    p = ctypes.c_void_p()
    rc = sp.lib().sp_ctx_create(0, ctypes.byref(p))
    assert rc in (-6, -4)


def test_c_caller_of_the_host_surface(tmp_path):
    """tests/c/host_surface_demo.c compiled with gcc against include/sparsep.h and libsparsep.so."""
    import shutil
    import subprocess
    if shutil.which("gcc") is None:
        pytest.skip("no C compiler")
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    exe = str(tmp_path / "demo")
    libdir = os.path.join(root, "sparsep_amd")
    subprocess.run(["gcc", "-O1", "-std=c11", "-I", os.path.join(root, "include"),
                    os.path.join(root, "tests", "c", "host_surface_demo.c"), "-o", exe, "-L", libdir, "-lsparsep",
                    f"-Wl,-rpath,{libdir}"], check=True)
    out = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    assert out.stdout.startswith("ok")


def test_init_vector_compat(sp):
    """init_vector (spmv/1D/CSR-row/host/app.c:57-61): x[i] = i % 4 + 1 in val_dt."""
    for dt in range(6):
        x = sp.init_vector(dt, 1001)
        assert x.dtype == sp.NP_DTYPE[dt]
        assert np.array_equal(x, (np.arange(1001) % 4 + 1).astype(sp.NP_DTYPE[dt]))
