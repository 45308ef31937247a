"""Replays the committed reference vectors (tests/golden/host_surface.*,
written by tests/golden/make_golden.py from the reference's own host support
code) against libsparsep.so: builders, 1D/tasklet/block partitioners, the
compat Matrix Market loader and the 2D tilings.  Needs no reference sources,
so it runs everywhere."""
import ctypes
import json
import os

import numpy as np
import pytest

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
with open(os.path.join(GOLD, "host_surface.json")) as _f:
    CASES = json.load(_f)
_ARR = None


def arrays(case):
    global _ARR
    if _ARR is None:
        _ARR = np.load(os.path.join(GOLD, "host_surface.npz"), allow_pickle=False)
    return {k: _ARR[f"{case['id']}/{k}"] for k in case["arrays"]}


def cases(kind):
    return [c for c in CASES if c["kind"] == kind]


def ids(kind):
    return [f"{kind}{c['id']}" for c in cases(kind)]


def test_fixture_covers_every_kind():
    kinds = {c["kind"] for c in CASES}
    assert kinds == {"coo2csr", "sort_coo", "bcsr", "bcoo", "part1d", "tsklt", "part_bcsr", "part_bcoo", "mtx",
                     "tiled_csr", "tiled_coo", "tiled_bcoo", "coo_split"}


@pytest.mark.parametrize("c", cases("coo2csr"), ids=ids("coo2csr"))
def test_coo2csr(sp, c):
    a = arrays(c)
    C = sp.Coo.from_arrays(c["nrows"], c["ncols"], a["r"], a["c"], a["v"]).to_csr()
    assert np.array_equal(C.rowptr, a["rowptr"])
    assert np.array_equal(C.colind, a["colind"]) and np.array_equal(C.values, a["values"])


@pytest.mark.parametrize("c", cases("sort_coo"), ids=ids("sort_coo"))
def test_sort_coo(sp, c):
    a = arrays(c)
    A = sp.Coo.from_arrays(c["nrows"], c["ncols"], a["r"], a["c"], a["v"], sort=True)
    assert np.array_equal(A.rowind, a["ro"]) and np.array_equal(A.colind, a["co"])
    assert np.array_equal(A.values, a["vo"]) and np.array_equal(A.rows, a["rows"])


@pytest.mark.parametrize("c", cases("bcsr"), ids=ids("bcsr"))
def test_bcsr_builder(sp, c):
    a = arrays(c)
    Cm = sp.Coo.from_arrays(c["nrows"], c["ncols"], a["r"], a["c"], a["v"]).to_csr()
    B = Cm.to_bcsr(c["R"], c["C"], count_values=True, sort=bool(c["sort"]))
    if not c["sort"]:
        B.count_nnz_per_block(count_values=True)
    assert (B.num_block_rows, B.num_block_cols, B.num_blocks) == (c["nbr"], c["nbc"], c["nb"])
    assert np.array_equal(B.browptr, a["browptr"]) and np.array_equal(B.bcolind, a["bcolind"])
    assert np.array_equal(B.bval, a["bval"]) and np.array_equal(B.nnz_per_block, a["npb"])


@pytest.mark.parametrize("c", cases("bcoo"), ids=ids("bcoo"))
def test_bcsr2bcoo(sp, c):
    a = arrays(c)
    O = sp.Coo.from_arrays(c["nrows"], c["ncols"], a["r"], a["c"], a["v"]).to_csr().to_bcsr(
        4, 4, count_values=True).to_bcoo()
    assert np.array_equal(O.browind, a["browind"]) and np.array_equal(O.bcolind, a["bcolind"])
    assert np.array_equal(O.bval, a["bval"]) and np.array_equal(O.nnz_per_block, a["npb"])


def test_1d_partitioners(sp):
    for c in cases("part1d"):
        a = arrays(c)
        mine = sp.partition_by_nnz(a["rowptr"], c["P"]) if c["by_nnz"] else sp.partition_by_row(c["n"], c["P"])
        assert np.array_equal(mine, a["split"]), c


def test_tasklet_partitioners(sp):
    for c in cases("tsklt"):
        a = arrays(c)
        assert np.array_equal(sp.partition_tsklt_by_row(c["rows"], c["T"], c["gran"]), a["by_row"]), c
        got = sp.partition_tsklt_by_nnz(a["rowptr"], c["r0"], c["rows"], c["nnz"], c["T"], c["gran"])
        assert np.array_equal(got, a["by_nnz"]), c


def test_bcsr_partitioners(sp):
    for c in cases("part_bcsr"):
        a = arrays(c)
        B = sp.Coo.from_arrays(c["nrows"], c["ncols"], a["r"], a["c"], a["v"]).to_csr().to_bcsr(
            4, 4, count_values=False)
        assert np.array_equal(B.browptr, a["browptr"])
        npb = a["npb"]
        bs, bl, ns = sp.partition_bcsr(B.browptr, npb, int(npb.sum()), c["P"], c["by_nnz"])
        assert np.array_equal(bs, a["block_split"]) and np.array_equal(bl, a["blocks"])
        assert np.array_equal(ns, a["nnzs"])
        T = c["T"]
        for d in range(c["P"]):
            t = sp.partition_tsklt_bcsr(B.browptr, npb, bs, bl, ns, d, T, c["tsk_nnz"])
            assert np.array_equal(t, a["tsklt"][d * (T + 1):(d + 1) * (T + 1)]), (c, d)


def test_bcoo_partitioners(sp):
    for c in cases("part_bcoo"):
        a = arrays(c)
        npb = a["npb"]
        bs, ns = sp.partition_bcoo(npb, int(npb.sum()), c["P"], c["by_nnz"])
        assert np.array_equal(bs, a["block_split"]) and np.array_equal(ns, a["nnzs"])
        T = c["T"]
        for d in range(c["P"]):
            t = sp.partition_tsklt_bcoo(npb, bs, ns, d, T, c["tsk_nnz"])
            assert np.array_equal(t, a["tsklt"][d * (T + 1):(d + 1) * (T + 1)]), (c, d)


@pytest.mark.parametrize("c", cases("mtx"), ids=ids("mtx"))
def test_mtx_compat_loader(sp, c):
    a = arrays(c)
    ctypes.CDLL("libc.so.6").srand(1)  # the reference reads values from unseeded rand()
    A = sp.read_mtx(os.path.join(GOLD, "small.mtx"), c["dtype"], c["flags"])
    assert [A.nrows, A.ncols, A.nnz] == c["dims"]
    assert np.array_equal(A.rowind, a["rowind"]) and np.array_equal(A.colind, a["colind"])
    assert np.array_equal(A.values, a["values"])


@pytest.mark.parametrize("c", cases("tiled_csr"), ids=ids("tiled_csr"))
def test_tiled_csr(sp, c):
    a = arrays(c)
    A = sp.Coo.from_arrays(c["nrows"], c["ncols"], a["r"], a["c"], a["v"])
    T = sp.coo2tiled(A, c["tiling"], sp.FMT_CSR, c["horz"], c["vert"])
    assert [T.ntiles, T.tile_height, T.tile_width] == c["dims"]
    assert np.array_equal(T.drowptr, a["drowptr"])
    assert np.array_equal(T.dcolind, a["dcolind"]) and np.array_equal(T.dval, a["dval"])
    if c["tiling"] == sp.TILE_BD:
        assert np.array_equal(T.vert_tile_widths, a["cuts"])
    if c["tiling"]:
        assert np.array_equal(sp.partition_tiled(T, c["horz"]), a["split"])


@pytest.mark.parametrize("c", cases("tiled_coo"), ids=ids("tiled_coo"))
def test_tiled_coo(sp, c):
    a = arrays(c)
    A = sp.Coo.from_arrays(c["nrows"], c["ncols"], a["r"], a["c"], a["v"])
    T = sp.coo2tiled(A, c["tiling"], sp.FMT_COO, c["horz"], c["vert"])
    assert T.ntiles == c["ntiles"]
    assert np.array_equal(T.drowind, a["drowind"]) and np.array_equal(T.dcolind, a["dcolind"])
    assert np.array_equal(T.dval, a["dval"])
    if c["tiling"]:
        assert np.array_equal(sp.partition_tiled(T, c["horz"]), a["split"])


@pytest.mark.parametrize("c", cases("tiled_bcoo"), ids=ids("tiled_bcoo"))
def test_tiled_bcoo(sp, c):
    a = arrays(c)
    A = sp.Coo.from_arrays(c["nrows"], c["ncols"], a["r"], a["c"], a["v"])
    T = sp.coo2tiled(A, c["tiling"], sp.FMT_BCOO, c["horz"], c["vert"], 4, 4, count_values=True)
    assert [T.ntiles, T.num_blocks, T.num_block_rows] == c["dims"]
    assert np.array_equal(T.blocks_per_partition, a["bpp"])
    assert np.array_equal(T.browind, a["browind"]) and np.array_equal(T.bcolind, a["bcolind"])
    assert np.array_equal(T.bval, a["bval"]) and np.array_equal(T.nnz_per_block, a["npb"])
    if c["tiling"]:
        split = sp.partition_tiled(T, c["horz"], c["by_nnz"]).reshape(c["vert"], c["horz"] + 1)
        assert np.array_equal(np.diff(split.astype(np.int64), axis=1).ravel(), a["blocks"])


@pytest.mark.parametrize("c", cases("coo_split"), ids=ids("coo_split"))
def test_coo_nnz_split(sp, c):
    a = arrays(c)
    s = sp.partition_coo_nnz(a["rows"], c["nnz"], c["P"])
    for name in ("prev_rows", "rows_per_part", "prev_nnz", "nnz_per_part", "merge"):
        assert np.array_equal(s[name], a[name]), name
    for i in range(c["tsk_parts"]):
        start, cnt = sp.partition_even(int(s["nnz_per_part"][i]), 16)
        assert np.array_equal(start, a["tstart"][i * 16:(i + 1) * 16])
        assert np.array_equal(cnt, a["tcount"][i * 16:(i + 1) * 16])
