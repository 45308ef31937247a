"""The COO-nnz inter-part split with straddling rows (sp_partition_coo_nnz,
restating spmv/1D/COO-nnz/host/app.c:118-156) and the even split
(sp_partition_even, app.c:186-204 and spmv/2D/DCOO/host/app.c:189-207).

Two kinds of check:
  * against the reference's own loop, compiled from the text of app.c by
    oracle/Makefile into oracle/_ref/libref_coo_split.so (skipped when the
    reference sources were not present at build time);
  * invariants that fix the split without the reference: parts tile the
    entries in order with sizes floor/ceil(nnz/P), a part is flagged for
    merging exactly when it starts inside a row, and each part's row range
    covers its entries.  A planted mistake (merge flagged by the row index
    instead of the entry offset) must fail them."""
import ctypes
import os

import numpy as np
import pytest

import oracle_lib as orc

LIB = os.path.join(orc.REFDIR, "libref_coo_split.so")


def cases():
    rng = np.random.default_rng(77)
    out = []
    for nrows, P in ((1, 1), (5, 3), (40, 7), (40, 64), (300, 16), (1000, 64), (17, 200), (2000, 2048)):
        for kind in ("uniform", "empty", "heavy", "zeros_tail"):
            if kind == "uniform":
                rows = rng.integers(1, 20, nrows)
            elif kind == "empty":
                rows = rng.integers(0, 4, nrows) * rng.integers(0, 2, nrows)
            elif kind == "heavy":
                rows = rng.integers(0, 3, nrows)
                rows[rng.integers(0, nrows)] = 5000
            else:
                rows = rng.integers(1, 9, nrows)
                rows[nrows // 2:] = 0
            rows = rows.astype(np.uint32)
            if rows.sum() == 0:
                rows[0] = 1
            out.append((rows, P))
    return out


def check_invariants(rows, P, s):
    nnz = int(rows.sum())
    rowptr = np.concatenate([[0], np.cumsum(rows.astype(np.int64))])
    per = s["nnz_per_part"].astype(np.int64)
    assert per.sum() == nnz
    assert set(per.tolist()) <= {nnz // P, -(-nnz // P)}
    assert np.array_equal(s["prev_nnz"].astype(np.int64), np.concatenate([[0], np.cumsum(per)[:-1]]))
    for i in range(P):
        e0 = int(s["prev_nnz"][i])
        inside = bool(np.any((rowptr[:-1] < e0) & (e0 < rowptr[1:])))
        assert bool(s["merge"][i]) == inside, ("merge", i)
        if per[i] > 0:
            r0, nr = int(s["prev_rows"][i]), int(s["rows_per_part"][i])
            rc = int(np.searchsorted(rowptr, e0, side="right")) - 1  # the row holding entry e0
            assert r0 <= rc and (r0 == rc or rowptr[r0] == e0), ("first row, or empty rows before it", i)
            assert rowptr[r0 + nr] >= e0 + per[i], ("rows cover the entries", i)
            assert rowptr[r0 + nr - 1] < e0 + per[i], ("last row holds an entry", i)


@pytest.mark.parametrize("idx", range(len(cases())))
def test_coo_nnz_split_invariants(sp, idx):
    rows, P = cases()[idx]
    s = sp.partition_coo_nnz(rows, int(rows.sum()), P)
    check_invariants(rows, P, s)


def test_planted_merge_mistake_fails(sp):
    rows, P = np.array([3, 5, 2, 7, 1, 6], np.uint32), 4
    s = sp.partition_coo_nnz(rows, int(rows.sum()), P)
    check_invariants(rows, P, s)
    assert np.any(s["merge"] == 0) and np.any(s["merge"] == 1)
    bad = dict(s)
    bad["merge"] = (s["prev_rows"] != 0).astype(np.uint32)  # flagged by row, not by entry offset
    with pytest.raises(AssertionError):
        check_invariants(rows, P, bad)


def test_even_split_invariants(sp):
    for count, T in ((0, 1), (1, 16), (15, 16), (16, 16), (17, 16), (1000, 7), (5, 64), (123457, 256)):
        start, cnt = sp.partition_even(count, T)
        c = cnt.astype(np.int64)
        assert c.sum() == count and set(c.tolist()) <= {count // T, -(-count // T)}
        assert np.all(np.diff(c) <= 0), "the remainder goes to the first chunks"
        assert np.array_equal(start.astype(np.int64), np.concatenate([[0], np.cumsum(c)[:-1]]))


@pytest.mark.skipif(not os.path.exists(LIB), reason="oracle/_ref/libref_coo_split.so not built")
@pytest.mark.parametrize("idx", range(len(cases())))
def test_coo_nnz_split_matches_reference(sp, idx):
    rows, P = cases()[idx]
    L = ctypes.CDLL(LIB)
    nnz = int(rows.sum())
    outs = [np.zeros(P, np.uint32) for _ in range(5)]
    tp = min(P, 8)
    ts, tc = np.zeros(tp * 16, np.uint32), np.zeros(tp * 16, np.uint32)
    V = lambda a: ctypes.c_void_p(a.ctypes.data)
    ntask = L.ref_coo_nnz_split(ctypes.c_uint32(len(rows)), ctypes.c_uint32(nnz), V(rows), ctypes.c_uint32(P),
                                *[V(o) for o in outs], ctypes.c_uint32(tp), V(ts), V(tc))
    assert ntask == 16
    s = sp.partition_coo_nnz(rows, nnz, P)
    for name, want in zip(["prev_rows", "rows_per_part", "prev_nnz", "nnz_per_part", "merge"], outs):
        assert np.array_equal(s[name], want), name
    # the tasklet split of each part's entries (NR_TASKLETS = 16 in the harness build)
    for i in range(tp):
        start, cnt = sp.partition_even(int(s["nnz_per_part"][i]), 16)
        assert np.array_equal(start, ts[i * 16:(i + 1) * 16]) and np.array_equal(cnt, tc[i * 16:(i + 1) * 16])
