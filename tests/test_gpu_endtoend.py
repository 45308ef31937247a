"""End-to-end paths of the reference's host program on the GPU: a Matrix
Market file through the compat loader (readCOOMatrix,
spmv/1D/CSR-row/support/matrix.h:44-96) with the compat input vector
(init_vector, spmv/1D/CSR-row/host/app.c:57-61) into every kernel id, and the
five-phase report (spmv/1D/CSR-row/host/app.c:291-300) of bench.py's phases
workload.  Integers are compared bit-exact with the oracle, floats within
the rounding bound."""
import argparse
import os
import sys

import numpy as np
import pytest

from gpu_util import assert_parity, run

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden")


def write_mtx(path, nrows, ncols, rows, cols, vals):
    with open(path, "w") as f:
        f.write("%%MatrixMarket matrix coordinate real general\n% generated\n")
        f.write(f"{nrows} {ncols} {len(rows)}\n")
        for r, c, v in zip(rows, cols, vals):
            f.write(f"{r + 1} {c + 1} {v:.6f}\n")


@pytest.fixture(scope="module")
def mtx_file(tmp_path_factory):
    """A 3,000 x 2,500 power-law-ish matrix with empty rows and a dense row."""
    rng = np.random.default_rng(61)
    nrows, ncols = 3000, 2500
    deg = np.minimum(rng.zipf(1.8, nrows), 400)
    deg[::7] = 0
    deg[1234] = 2500
    rows = np.repeat(np.arange(nrows), deg)
    cols = np.concatenate([rng.choice(ncols, d, replace=False) for d in deg if d])
    vals = rng.uniform(-2, 2, rows.size)
    path = str(tmp_path_factory.mktemp("mtx") / "zipf.mtx")
    write_mtx(path, nrows, ncols, rows, cols, vals)
    return path


@pytest.mark.parametrize("dt", [2, 4])
@pytest.mark.parametrize("flags", [3, 7])  # compat values; padded dims; sorted or file order
def test_mtx_every_kernel(sp, ctx, mtx_file, dt, flags):
    """The file read the reference's way (values rand()%4+1, dims padded) and
    x[i] = i%4+1, through all 25 kernel ids with their default plans."""
    for path in (mtx_file, os.path.join(GOLD, "small.mtx")):
        A = sp.read_mtx(path, dt, flags)
        assert A.nrows % (8 // np.dtype(sp.NP_DTYPE[dt]).itemsize) == 0  # padded
        x = sp.init_vector(dt, A.ncols)
        assert np.array_equal(x[:8], (np.arange(8) % 4 + 1).astype(x.dtype))
        for kid, kname in enumerate(sp.KERNELS):
            y, _ = run(sp, ctx, kid, A, x)
            assert_parity(dt, y, A, x, f"{os.path.basename(path)}/{kname}/flags{flags}")


def test_mtx_file_values(sp, ctx, mtx_file):
    """Without the compat flag the file's own values are used (FP64)."""
    A = sp.read_mtx(mtx_file, 5, 4)
    x = sp.gen_vector(5, A.ncols, seed=62)
    for kname in ("CSR-nnz", "COO-nnz", "BCSR-nnz", "RBDCSR"):
        y, _ = run(sp, ctx, kname, A, x)
        assert_parity(5, y, A, x, kname)


def test_five_phase_report(sp, ctx, mtx_file):
    """bench.py --workload phases on a Matrix Market file: the five phases are
    reported (load matrix, load input vector, kernel, retrieve output, merge),
    each non-negative, and the result matches the oracle."""
    sys.path.insert(0, ROOT)
    import bench
    import oracle_lib as orc
    args = argparse.Namespace(mtx=mtx_file, steps=3, warmup=1)
    (rec,) = bench.wl_phases(sp, ctx, args, orc)
    ex = rec["extra"]
    for k in ("load_matrix_ms", "load_input_vector_ms", "kernel_ms", "retrieve_output_vector_ms", "merge_ms"):
        assert ex[k] >= 0.0, k
    assert ex["kernel_ms"] > 0.0 and ex["load_matrix_ms"] > 0.0
    assert ex["parity_vs_oracle"] is True
    assert rec["unit"] == "ms end to end" and rec["higher_is_better"] is False
    assert abs(rec["value"] - (ex["load_matrix_ms"] + ex["load_input_vector_ms"] + ex["kernel_ms"]
                               + ex["retrieve_output_vector_ms"])) < 1e-2
