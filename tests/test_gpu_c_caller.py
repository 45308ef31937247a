"""The C caller (tests/c/host_surface_demo.c) on the GPU: one CSR-row INT32
SpMV through the device API, checked in C against the wrapping loop."""
import os
import shutil
import subprocess

import pytest

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_c_caller_runs_spmv_on_gpu(tmp_path, ctx):
    if shutil.which("gcc") is None:
        pytest.skip("no C compiler")
    exe = str(tmp_path / "demo")
    libdir = os.path.join(ROOT, "sparsep_amd")
    subprocess.run(["gcc", "-O1", "-std=c11", "-I", os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "c", "host_surface_demo.c"), "-o", exe, "-L", libdir, "-lsparsep",
                    f"-Wl,-rpath,{libdir}"], check=True)
    out = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    assert out.stdout.strip() == "ok (gpu)"
