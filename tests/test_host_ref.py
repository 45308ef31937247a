"""Host surface vs the reference's own support code, compiled from
support/matrix.h + support/partition.h into oracle/_ref (CPU only).

Skipped when oracle/_ref was not built (the reference sources are absent);
tests/test_golden.py then covers the same functions from committed vectors.
"""
import ctypes
import os

import numpy as np
import pytest

import oracle_lib as orc

pytestmark = pytest.mark.skipif(not orc.have_refs(), reason="oracle/_ref not built")

V = ctypes.c_void_p
U = ctypes.c_uint32


def P(a):
    return V(a.ctypes.data)


def rand_coo(rng, nrows, ncols, nnz, dt, unique=True):
    if unique:
        keys = rng.choice(nrows * ncols, size=min(nnz, nrows * ncols), replace=False)
        r, c = (keys // ncols).astype(np.uint32), (keys % ncols).astype(np.uint32)
    else:
        r = rng.integers(0, nrows, nnz).astype(np.uint32)
        c = rng.integers(0, ncols, nnz).astype(np.uint32)
    v = rng.integers(-4, 5, size=len(r)).astype(orc.NP[dt])
    if dt >= 4:
        v = (rng.standard_normal(len(r))).astype(orc.NP[dt])
    return r, c, v


@pytest.mark.parametrize("dt", [0, 2, 5])
def test_coo2csr_matches_reference(sp, dt):
    L = orc.ref("CSR-row", dt)
    rng = np.random.default_rng(dt)
    for nrows, ncols, nnz in [(50, 60, 300), (1, 1, 1), (200, 30, 1000), (17, 1000, 0)]:
        r, c, v = rand_coo(rng, nrows, ncols, nnz, dt, unique=False)
        nnz = len(r)
        rp, ci, vo = np.zeros(nrows + 1, np.uint32), np.zeros(max(nnz, 1), np.uint32), np.zeros(max(nnz, 1), orc.NP[dt])
        L.ref_coo2csr(U(nrows), U(ncols), U(nnz), P(r), P(c), P(v), P(rp), P(ci), P(vo))
        A = sp.Coo.from_arrays(nrows, ncols, r, c, v)
        C = A.to_csr()
        assert np.array_equal(C.rowptr, rp)
        assert np.array_equal(C.colind, ci[:nnz])
        assert np.array_equal(C.values, vo[:nnz])


def test_sort_coo_matches_reference(sp):
    L = orc.ref("COO-row", 2)
    rng = np.random.default_rng(7)
    for nrows, ncols, nnz in [(40, 50, 500), (300, 7, 1000), (1, 2000, 900)]:
        r, c, v = rand_coo(rng, nrows, ncols, nnz, 2)
        nnz = len(r)
        ro, co, vo, rows = (np.zeros(nnz, np.uint32), np.zeros(nnz, np.uint32), np.zeros(nnz, np.int32),
                            np.zeros(nrows, np.uint32))
        L.ref_sort_coo(U(nrows), U(ncols), U(nnz), P(r), P(c), P(v), P(ro), P(co), P(vo), P(rows))
        A = sp.Coo.from_arrays(nrows, ncols, r, c, v, sort=True)
        assert np.array_equal(A.rowind, ro) and np.array_equal(A.colind, co) and np.array_equal(A.values, vo)
        assert np.array_equal(A.rows, rows)


def ref_bcsr(L, dt, nrows, ncols, r, c, v, R, C, sort=1):
    dims = np.zeros(3, np.uint64)
    L.ref_build_bcsr(U(nrows), U(ncols), U(len(r)), P(r), P(c), P(v), U(R), U(C), ctypes.c_int(sort), P(dims))
    nbr, nbc, nb = (int(d) for d in dims)
    bp, bc = np.zeros(nbr + 1, np.uint32), np.zeros(max(nb, 1), np.uint32)
    bv, npb = np.zeros(max(nb, 1) * R * C, orc.NP[dt]), np.zeros(max(nb, 1), np.uint32)
    L.ref_copy_bcsr(P(bp), P(bc), P(bv), P(npb))
    return nbr, nbc, nb, bp, bc[:nb], bv[:nb * R * C], npb[:nb]


@pytest.mark.parametrize("dt", [0, 2, 4])
@pytest.mark.parametrize("rc", [(4, 4), (2, 3), (1, 1), (3, 2)])
def test_bcsr_builder_matches_reference(sp, dt, rc):
    R, C = rc
    L = orc.ref("BCSR-block", dt)
    rng = np.random.default_rng(11 + dt)
    for nrows, ncols, nnz in [(37, 41, 300), (64, 64, 900), (5, 200, 100), (130, 9, 400)]:
        r, c, v = rand_coo(rng, nrows, ncols, nnz, dt)
        A = sp.Coo.from_arrays(nrows, ncols, r, c, v)
        Cm = A.to_csr()
        for sort in (0, 1):
            nbr, nbc, nb, bp, bc, bv, npb = ref_bcsr(L, dt, nrows, ncols, r, c, v, R, C, sort)
            B = Cm.to_bcsr(R, C, count_values=True, sort=bool(sort))
            if not sort:
                B.count_nnz_per_block(count_values=True)
            assert (B.num_block_rows, B.num_block_cols, B.num_blocks) == (nbr, nbc, nb)
            assert np.array_equal(B.browptr, bp)
            assert np.array_equal(B.bcolind, bc), "block order (first-seen or sorted) differs"
            assert np.array_equal(B.bval, bv)
            assert np.array_equal(B.nnz_per_block, npb), "value count per block differs"


@pytest.mark.parametrize("dt", [2, 5])
def test_bcsr2bcoo_matches_reference(sp, dt):
    L = orc.ref("BCOO-block", dt)
    rng = np.random.default_rng(5)
    r, c, v = rand_coo(rng, 70, 90, 800, dt)
    ref_bcsr(L, dt, 70, 90, r, c, v, 4, 4)
    A = sp.Coo.from_arrays(70, 90, r, c, v)
    B = A.to_csr().to_bcsr(4, 4, count_values=True)
    O = B.to_bcoo()
    nb = O.num_blocks
    bi, bc, bv, npb = (np.zeros(nb, np.uint32), np.zeros(nb, np.uint32), np.zeros(nb * 16, orc.NP[dt]),
                       np.zeros(nb, np.uint32))
    L.ref_bcsr2bcoo(P(bi), P(bc), P(bv), P(npb))
    assert np.array_equal(O.browind, bi) and np.array_equal(O.bcolind, bc)
    assert np.array_equal(O.bval, bv) and np.array_equal(O.nnz_per_block, npb)


def row_lengths(rng, n, kind):
    if kind == "uniform":
        return rng.integers(0, 20, n)
    if kind == "skew":
        return (rng.pareto(1.2, n) * 3).astype(np.int64) % 5000
    if kind == "empty":
        return np.zeros(n, np.int64)
    return np.where(rng.random(n) < 0.5, 0, rng.integers(1, 50, n))


KINDS = ["uniform", "skew", "empty", "sparse"]


@pytest.mark.parametrize("by_nnz", [0, 1])
def test_1d_partitioners_match_reference(sp, by_nnz):
    rng = np.random.default_rng(100 + by_nnz)
    Lc = orc.ref("CSR-nnz" if by_nnz else "CSR-row", 2)
    Lo = orc.ref("COO-nnz-rgrn" if by_nnz else "COO-row", 2)
    for kind in KINDS:
        for n in (1, 7, 100, 1000):
            lens = row_lengths(rng, n, kind)
            rp = np.concatenate([[0], np.cumsum(lens)]).astype(np.uint32)
            nnz = int(rp[-1])
            for Pn in (1, 2, 3, 16, 64, 2048):
                ref = np.zeros(Pn + 1, np.uint32)
                Lc.ref_partition(U(n), U(nnz), P(rp), ctypes.c_int(Pn), P(ref))
                mine = sp.partition_by_nnz(rp, Pn) if by_nnz else sp.partition_by_row(n, Pn)
                assert np.array_equal(mine, ref), (kind, n, Pn)
                refo = np.zeros(Pn + 1, np.uint32)
                rows = lens.astype(np.uint32)
                Lo.ref_partition(U(n), U(nnz), P(rows), ctypes.c_int(Pn), P(refo))
                assert np.array_equal(mine, refo), ("coo", kind, n, Pn)


@pytest.mark.parametrize("dt", [0, 2, 5])
def test_tasklet_partitioners_match_reference(sp, dt):
    rng = np.random.default_rng(200 + dt)
    L = orc.ref("CSR-row", dt)
    Lo = orc.ref("COO-row", dt)
    gran = 8 // [1, 2, 4, 8, 4, 8][dt]
    for kind in KINDS:
        for n in (5, 64, 333):
            lens = row_lengths(rng, n, kind)
            rp = np.concatenate([[0], np.cumsum(lens)]).astype(np.uint32)
            for T in (1, 2, 16, 24):
                for r0, r1 in ((0, n), (n // 3, n - n // 4)):
                    rows = r1 - r0
                    nnzp = int(rp[r1] - rp[r0])
                    ref = np.zeros(T + 1, np.uint32)
                    L.ref_tsklt_by_row(ctypes.c_int(rows), ctypes.c_int(T), P(ref))
                    assert np.array_equal(sp.partition_tsklt_by_row(rows, T, gran), ref), (kind, n, T, "row")
                    L.ref_tsklt_by_nnz(U(n), U(int(rp[-1])), P(rp), ctypes.c_int(rows), ctypes.c_int(nnzp),
                                       ctypes.c_int(r0), ctypes.c_int(T), P(ref))
                    assert np.array_equal(sp.partition_tsklt_by_nnz(rp, r0, rows, nnzp, T, gran), ref), \
                        (kind, n, T, r0, "nnz")
                    # COO twins use granularity 1 and the per-row counts
                    cnt = lens.astype(np.uint32)
                    Lo.ref_tsklt_by_row(ctypes.c_int(rows), ctypes.c_int(T), P(ref))
                    assert np.array_equal(sp.partition_tsklt_by_row(rows, T, 1), ref)
                    Lo.ref_tsklt_by_nnz(U(n), U(int(rp[-1])), P(cnt), ctypes.c_int(rows), ctypes.c_int(nnzp),
                                        ctypes.c_int(r0), ctypes.c_int(T), P(ref))
                    assert np.array_equal(sp.partition_tsklt_by_nnz(rp, r0, rows, nnzp, T, 1), ref)


@pytest.mark.parametrize("by_nnz", [0, 1])
@pytest.mark.parametrize("tsk_nnz", [0, 1])
def test_bcsr_partitioners_match_reference(sp, by_nnz, tsk_nnz):
    L = orc.ref("BCSR-nnz" if by_nnz else "BCSR-block", 2)
    rng = np.random.default_rng(300 + 2 * by_nnz + tsk_nnz)
    for (nr, nc, nnz) in [(90, 80, 1500), (64, 400, 3000), (33, 33, 50)]:
        r, c, v = rand_coo(rng, nr, nc, nnz, 2)
        ref_bcsr(L, 2, nr, nc, r, c, v, 4, 4)
        B = sp.Coo.from_arrays(nr, nc, r, c, v).to_csr().to_bcsr(4, 4, count_values=False)
        npb = B.nnz_per_block.copy()
        tot = int(npb.sum())
        for Pn in (1, 2, 5, 16):
            for T in (1, 3, 16):
                bs, bl, ns = np.zeros(Pn + 1, np.uint32), np.zeros(Pn, np.uint32), np.zeros(Pn, np.uint32)
                ts = np.zeros(Pn * (T + 1), np.uint32)
                L.ref_partition_bcsr(P(npb), U(tot), ctypes.c_int(Pn), ctypes.c_int(T), ctypes.c_int(tsk_nnz),
                                     P(bs), P(bl), P(ns), P(ts))
                mbs, mbl, mns = sp.partition_bcsr(B.browptr, npb, tot, Pn, by_nnz)
                assert np.array_equal(mbs, bs) and np.array_equal(mbl, bl) and np.array_equal(mns, ns), (Pn,)
                for d in range(Pn):
                    mt = sp.partition_tsklt_bcsr(B.browptr, npb, mbs, mbl, mns, d, T, tsk_nnz)
                    assert np.array_equal(mt, ts[d * (T + 1):(d + 1) * (T + 1)]), (Pn, T, d)


@pytest.mark.parametrize("by_nnz", [0, 1])
@pytest.mark.parametrize("tsk_nnz", [0, 1])
def test_bcoo_partitioners_match_reference(sp, by_nnz, tsk_nnz):
    L = orc.ref("BCOO-nnz" if by_nnz else "BCOO-block", 2)
    rng = np.random.default_rng(400 + 2 * by_nnz + tsk_nnz)
    for (nr, nc, nnz) in [(90, 80, 1500), (40, 300, 2000)]:
        r, c, v = rand_coo(rng, nr, nc, nnz, 2)
        ref_bcsr(L, 2, nr, nc, r, c, v, 4, 4)
        B = sp.Coo.from_arrays(nr, nc, r, c, v).to_csr().to_bcsr(4, 4, count_values=False)
        O = B.to_bcoo()
        nb = O.num_blocks
        tmp = [np.zeros(nb, np.uint32), np.zeros(nb, np.uint32), np.zeros(nb * 16, np.int32), np.zeros(nb, np.uint32)]
        L.ref_bcsr2bcoo(*[P(a) for a in tmp])
        npb = O.nnz_per_block.copy()
        tot = int(npb.sum())
        for Pn in (1, 2, 7, 32):
            for T in (1, 4, 16):
                bs, ns, ts = np.zeros(Pn + 1, np.uint32), np.zeros(Pn, np.uint32), np.zeros(Pn * (T + 1), np.uint32)
                L.ref_partition_bcoo(P(npb), U(tot), ctypes.c_int(Pn), ctypes.c_int(T), ctypes.c_int(tsk_nnz),
                                     P(bs), P(ns), P(ts))
                mbs, mns = sp.partition_bcoo(npb, tot, Pn, by_nnz)
                assert np.array_equal(mbs, bs) and np.array_equal(mns, ns), (Pn, T)
                for d in range(Pn):
                    mt = sp.partition_tsklt_bcoo(npb, mbs, mns, d, T, tsk_nnz)
                    assert np.array_equal(mt, ts[d * (T + 1):(d + 1) * (T + 1)]), (Pn, T, d)


def test_mtx_loader_compat_matches_reference(sp, tmp_path):
    """readCOOMatrix: padding to 8/sizeof(val) and rand()%4+1 values."""
    path = str(tmp_path / "m.mtx")
    rng = np.random.default_rng(9)
    with open(path, "w") as f:
        f.write("%%MatrixMarket matrix coordinate real general\n% comment\n")
        f.write("13 11 40\n")
        keys = rng.choice(13 * 11, 40, replace=False)
        for k in keys:
            f.write(f"{k // 11 + 1} {k % 11 + 1} {rng.standard_normal():.6f}\n")
    for dt, dirname, flags in [(0, "CSR-row", 3), (2, "CSR-row", 3), (5, "CSR-row", 3), (2, "BCSR-block", 2)]:
        L = orc.ref(dirname, dt)
        dims = np.zeros(3, np.uint32)
        L.ref_srand(ctypes.c_uint(1))
        assert L.ref_load_mtx(path.encode(), P(dims)) == 0
        n = int(dims[2])
        ri, ci, v = np.zeros(n, np.uint32), np.zeros(n, np.uint32), np.zeros(n, orc.NP[dt])
        L.ref_copy_coo(P(ri), P(ci), P(v))
        L.ref_srand(ctypes.c_uint(1))  # the same libc generator the product's compat mode uses
        A = sp.read_mtx(path, dt, flags)
        assert (A.nrows, A.ncols, A.nnz) == tuple(int(d) for d in dims)
        assert np.array_equal(A.rowind, ri) and np.array_equal(A.colind, ci) and np.array_equal(A.values, v)


def test_mtx_loader_errors(sp, tmp_path):
    with pytest.raises(sp._lib.SparsePError) as e:
        sp.read_mtx(str(tmp_path / "missing.mtx"), 2)
    assert e.value.rc == -3
    bad = tmp_path / "bad.mtx"
    bad.write_text("%%MatrixMarket\n3 3 2\n1 1 1.0\n")  # one entry short
    with pytest.raises(sp._lib.SparsePError):
        sp.read_mtx(str(bad), 2)
    bad.write_text("3 3 1\n4 1 1.0\n")  # out of range
    with pytest.raises(sp._lib.SparsePError):
        sp.read_mtx(str(bad), 2)
