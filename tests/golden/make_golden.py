"""Writes tests/golden/host_surface.npz + host_surface.json: input/output
vectors of the reference's own host support code (matrix builders,
partitioners, Matrix Market loader), produced by the harness libraries that
oracle/Makefile compiles from the reference sources into oracle/_ref.

Run where the reference sources are present (after `make -C oracle`):
    python tests/golden/make_golden.py
The vectors are data (seeded inputs and the reference's outputs);
tests/test_golden.py replays them against libsparsep.so on any machine.
"""
import ctypes
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
import oracle_lib as orc  # noqa: E402

V, U, Ic = ctypes.c_void_p, ctypes.c_uint32, ctypes.c_int
ARR = {}
CASES = []


def P(a):
    return V(a.ctypes.data)


def put(case, **arrays):
    k = len(CASES)
    case["id"] = k
    for name, a in arrays.items():
        ARR[f"{k}/{name}"] = np.ascontiguousarray(a)
    case["arrays"] = sorted(arrays)
    CASES.append(case)


def rand_coo(rng, nrows, ncols, nnz, unique=True):
    if unique:
        keys = rng.choice(nrows * ncols, size=min(nnz, nrows * ncols), replace=False)
        r, c = (keys // ncols).astype(np.uint32), (keys % ncols).astype(np.uint32)
    else:
        r = rng.integers(0, nrows, nnz).astype(np.uint32)
        c = rng.integers(0, ncols, nnz).astype(np.uint32)
    v = rng.integers(-4, 5, size=len(r)).astype(np.int32)
    return r, c, v


def sorted_coo(rng, nrows, ncols, nnz):
    keys = np.sort(rng.choice(nrows * ncols, size=nnz, replace=False))
    r, c = (keys // ncols).astype(np.uint32), (keys % ncols).astype(np.uint32)
    v = (rng.integers(1, 5, size=nnz) * rng.choice([-1, 1], size=nnz)).astype(np.int32)
    return r, c, v


def ref_bcsr(L, nrows, ncols, r, c, v, R, C, sort=1):
    dims = np.zeros(3, np.uint64)
    L.ref_build_bcsr(U(nrows), U(ncols), U(len(r)), P(r), P(c), P(v), U(R), U(C), Ic(sort), P(dims))
    nbr, nbc, nb = (int(d) for d in dims)
    bp, bc = np.zeros(nbr + 1, np.uint32), np.zeros(max(nb, 1), np.uint32)
    bv, npb = np.zeros(max(nb, 1) * R * C, np.int32), np.zeros(max(nb, 1), np.uint32)
    L.ref_copy_bcsr(P(bp), P(bc), P(bv), P(npb))
    return nbr, nbc, nb, bp, bc[:nb], bv[:nb * R * C], npb[:nb]


def row_lengths(rng, n, kind):
    if kind == "uniform":
        return rng.integers(0, 20, n)
    if kind == "skew":
        return (rng.pareto(1.2, n) * 3).astype(np.int64) % 5000
    if kind == "empty":
        return np.zeros(n, np.int64)
    return np.where(rng.random(n) < 0.5, 0, rng.integers(1, 50, n))


def main():
    if not orc.have_refs():
        raise SystemExit("oracle/_ref is not built: run `make -C oracle` where the reference sources exist")
    rng = np.random.default_rng(2024)

    # coo2csr (spmv/1D/CSR-row/support/matrix.h:113-152), INT32
    L = orc.ref("CSR-row", 2)
    for nrows, ncols, nnz in [(50, 60, 300), (1, 1, 1), (17, 1000, 0), (200, 30, 1000)]:
        r, c, v = rand_coo(rng, nrows, ncols, nnz, unique=False)
        n = len(r)
        rp, ci, vo = np.zeros(nrows + 1, np.uint32), np.zeros(max(n, 1), np.uint32), np.zeros(max(n, 1), np.int32)
        L.ref_coo2csr(U(nrows), U(ncols), U(n), P(r), P(c), P(v), P(rp), P(ci), P(vo))
        put({"kind": "coo2csr", "nrows": nrows, "ncols": ncols}, r=r, c=c, v=v, rowptr=rp, colind=ci[:n],
            values=vo[:n])

    # sortCOOMatrix (spmv/1D/COO-row/support/matrix.h:86-125)
    L = orc.ref("COO-row", 2)
    for nrows, ncols, nnz in [(40, 50, 500), (300, 7, 1000)]:
        r, c, v = rand_coo(rng, nrows, ncols, nnz)
        n = len(r)
        ro, co, vo, rows = (np.zeros(n, np.uint32), np.zeros(n, np.uint32), np.zeros(n, np.int32),
                            np.zeros(nrows, np.uint32))
        L.ref_sort_coo(U(nrows), U(ncols), U(n), P(r), P(c), P(v), P(ro), P(co), P(vo), P(rows))
        put({"kind": "sort_coo", "nrows": nrows, "ncols": ncols}, r=r, c=c, v=v, ro=ro, co=co, vo=vo, rows=rows)

    # csr2bcsr + sort + value count (spmv/1D/BCSR-block/support/matrix.h:202-485)
    L = orc.ref("BCSR-block", 2)
    for (nrows, ncols, nnz) in [(37, 41, 300), (130, 9, 400)]:
        r, c, v = rand_coo(rng, nrows, ncols, nnz)
        v[::7] = 0  # value-zero entries: nnz_per_block counts values
        for R, C in [(4, 4), (2, 3)]:
            for sort in (0, 1):
                nbr, nbc, nb, bp, bc, bv, npb = ref_bcsr(L, nrows, ncols, r, c, v, R, C, sort)
                put({"kind": "bcsr", "nrows": nrows, "ncols": ncols, "R": R, "C": C, "sort": sort,
                     "nbr": nbr, "nbc": nbc, "nb": nb}, r=r, c=c, v=v, browptr=bp, bcolind=bc, bval=bv, npb=npb)

    # bcsr2bcoo (spmv/1D/BCOO-block/support/matrix.h:549)
    L = orc.ref("BCOO-block", 2)
    r, c, v = rand_coo(rng, 70, 90, 800)
    B = ref_bcsr(L, 70, 90, r, c, v, 4, 4)
    nbk = B[2]
    bi, bc, bv, npb = (np.zeros(nbk, np.uint32), np.zeros(nbk, np.uint32), np.zeros(nbk * 16, np.int32),
                       np.zeros(nbk, np.uint32))
    L.ref_bcsr2bcoo(P(bi), P(bc), P(bv), P(npb))
    put({"kind": "bcoo", "nrows": 70, "ncols": 90}, r=r, c=c, v=v, browind=bi, bcolind=bc, bval=bv, npb=npb)

    # 1D partitioners: partition_by_row / partition_by_nnz
    # (spmv/1D/CSR-row/support/partition.h:35-60, spmv/1D/CSR-nnz/support/partition.h:35-78)
    for by_nnz in (0, 1):
        Lc = orc.ref("CSR-nnz" if by_nnz else "CSR-row", 2)
        for kind in ("uniform", "skew", "empty", "sparse"):
            for n in (7, 100, 1000):
                lens = row_lengths(rng, n, kind)
                rp = np.concatenate([[0], np.cumsum(lens)]).astype(np.uint32)
                for Pn in (1, 3, 16, 64, 2048):
                    out = np.zeros(Pn + 1, np.uint32)
                    Lc.ref_partition(U(n), U(int(rp[-1])), P(rp), Ic(Pn), P(out))
                    put({"kind": "part1d", "by_nnz": by_nnz, "n": n, "P": Pn}, rowptr=rp, split=out)

    # tasklet partitioners (spmv/1D/CSR-row/support/partition.h:66,92), INT32 granularity 2
    L = orc.ref("CSR-row", 2)
    for kind in ("uniform", "skew", "sparse"):
        for n in (64, 333):
            lens = row_lengths(rng, n, kind)
            rp = np.concatenate([[0], np.cumsum(lens)]).astype(np.uint32)
            for T in (2, 16, 24):
                r0, r1 = n // 3, n - n // 4
                rows, nnzp = r1 - r0, int(rp[r1] - rp[r0])
                a, b = np.zeros(T + 1, np.uint32), np.zeros(T + 1, np.uint32)
                L.ref_tsklt_by_row(Ic(rows), Ic(T), P(a))
                L.ref_tsklt_by_nnz(U(n), U(int(rp[-1])), P(rp), Ic(rows), Ic(nnzp), Ic(r0), Ic(T), P(b))
                put({"kind": "tsklt", "r0": r0, "rows": rows, "nnz": nnzp, "T": T, "gran": 2},
                    rowptr=rp, by_row=a, by_nnz=b)

    # BCSR partitioners (spmv/1D/BCSR-block/support/partition.h:39-192)
    for by_nnz in (0, 1):
        L = orc.ref("BCSR-nnz" if by_nnz else "BCSR-block", 2)
        for tsk_nnz in (0, 1):
            r, c, v = rand_coo(rng, 90, 80, 1500)
            B = ref_bcsr(L, 90, 80, r, c, v, 4, 4)
            npb = B[6].copy()
            tot = int(npb.sum())
            for Pn, T in ((2, 3), (5, 16), (16, 1)):
                bs, bl, ns = np.zeros(Pn + 1, np.uint32), np.zeros(Pn, np.uint32), np.zeros(Pn, np.uint32)
                ts = np.zeros(Pn * (T + 1), np.uint32)
                L.ref_partition_bcsr(P(npb), U(tot), Ic(Pn), Ic(T), Ic(tsk_nnz), P(bs), P(bl), P(ns), P(ts))
                put({"kind": "part_bcsr", "by_nnz": by_nnz, "tsk_nnz": tsk_nnz, "P": Pn, "T": T, "nrows": 90,
                     "ncols": 80}, r=r, c=c, v=v, browptr=B[3], npb=npb, block_split=bs, blocks=bl, nnzs=ns,
                    tsklt=ts)

    # BCOO partitioners (spmv/1D/BCOO-block/support/partition.h:37-121)
    for by_nnz in (0, 1):
        L = orc.ref("BCOO-nnz" if by_nnz else "BCOO-block", 2)
        for tsk_nnz in (0, 1):
            r, c, v = rand_coo(rng, 40, 300, 2000)
            B = ref_bcsr(L, 40, 300, r, c, v, 4, 4)
            nbk = B[2]
            tmp = [np.zeros(nbk, np.uint32), np.zeros(nbk, np.uint32), np.zeros(nbk * 16, np.int32),
                   np.zeros(nbk, np.uint32)]
            L.ref_bcsr2bcoo(*[P(a) for a in tmp])
            npb = tmp[3].copy()
            tot = int(npb.sum())
            for Pn, T in ((2, 4), (7, 16), (32, 1)):
                bs, ns, ts = np.zeros(Pn + 1, np.uint32), np.zeros(Pn, np.uint32), np.zeros(Pn * (T + 1), np.uint32)
                L.ref_partition_bcoo(P(npb), U(tot), Ic(Pn), Ic(T), Ic(tsk_nnz), P(bs), P(ns), P(ts))
                put({"kind": "part_bcoo", "by_nnz": by_nnz, "tsk_nnz": tsk_nnz, "P": Pn, "T": T}, npb=npb,
                    block_split=bs, nnzs=ns, tsklt=ts)

    # readCOOMatrix compat mode (spmv/1D/CSR-row/support/matrix.h:44-96): padding + rand()%4+1 after srand(1)
    mtx = os.path.join(HERE, "small.mtx")
    r2 = np.random.default_rng(9)
    with open(mtx, "w") as f:
        f.write("%%MatrixMarket matrix coordinate real general\n% fixture for the compat loader\n13 11 40\n")
        for k in r2.choice(13 * 11, 40, replace=False):
            f.write(f"{k // 11 + 1} {k % 11 + 1} {r2.standard_normal():.6f}\n")
    for dt, dirname, flags in [(0, "CSR-row", 3), (2, "CSR-row", 3), (5, "CSR-row", 3), (2, "BCSR-block", 2)]:
        L = orc.ref(dirname, dt)
        dims = np.zeros(3, np.uint32)
        L.ref_srand(ctypes.c_uint(1))
        assert L.ref_load_mtx(mtx.encode(), P(dims)) == 0
        n = int(dims[2])
        ri, ci, vv = np.zeros(n, np.uint32), np.zeros(n, np.uint32), np.zeros(n, orc.NP[dt])
        L.ref_copy_coo(P(ri), P(ci), P(vv))
        put({"kind": "mtx", "dtype": dt, "flags": flags, "dims": [int(d) for d in dims]}, rowind=ri, colind=ci,
            values=vv)

    # 2D tilings (spmv/2D/*/support/matrix.h) and stripe partitioners (spmv/2D/*/support/partition.h), INT32
    shapes = [(64, 80, 900, 4, 2), (100, 37, 700, 3, 5)]
    for tiling, dirname in ((0, "DCSR"), (1, "RBDCSR"), (2, "BDCSR")):
        L = orc.ref(dirname, 2)
        for (nr, nc, nnz, horz, vert) in shapes:
            r, c, v = sorted_coo(rng, nr, nc, nnz)
            dims = np.zeros(3, np.uint32)
            L.ref2_build(U(nr), U(nc), U(nnz), P(r), P(c), P(v), Ic(horz), Ic(vert), P(dims))
            rp = np.zeros(int(dims[0]) * (int(dims[1]) + 1), np.uint32)
            ci, vo, cuts = np.zeros(nnz, np.uint32), np.zeros(nnz, np.int32), np.zeros(vert + 1, np.uint32)
            L.ref2_copy_csr(P(rp), P(ci), P(vo), P(cuts))
            arrays = dict(r=r, c=c, v=v, drowptr=rp, dcolind=ci, dval=vo, cuts=cuts)
            if tiling:
                split = np.zeros(vert * (horz + 1), np.uint32)
                L.ref2_partition(P(split))
                arrays["split"] = split
            put({"kind": "tiled_csr", "tiling": tiling, "nrows": nr, "ncols": nc, "horz": horz, "vert": vert,
                 "dims": [int(d) for d in dims]}, **arrays)
    for tiling, dirname in ((0, "DCOO"), (1, "RBDCOO"), (2, "BDCOO")):
        L = orc.ref(dirname, 2)
        for (nr, nc, nnz, horz, vert) in shapes:
            r, c, v = sorted_coo(rng, nr, nc, nnz)
            dims = np.zeros(3, np.uint32)
            L.ref2_build(U(nr), U(nc), U(nnz), P(r), P(c), P(v), Ic(horz), Ic(vert), P(dims))
            ri, ci, vo = np.zeros(nnz, np.uint32), np.zeros(nnz, np.uint32), np.zeros(nnz, np.int32)
            L.ref2_copy_coo(P(ri), P(ci), P(vo))
            arrays = dict(r=r, c=c, v=v, drowind=ri, dcolind=ci, dval=vo)
            if tiling:
                split = np.zeros(vert * (horz + 1), np.uint32)
                L.ref2_partition(P(split))
                arrays["split"] = split
            put({"kind": "tiled_coo", "tiling": tiling, "nrows": nr, "ncols": nc, "horz": horz, "vert": vert,
                 "ntiles": int(dims[0])}, **arrays)
    for tiling, dirname in ((0, "DBCOO"), (1, "RBDBCOO-block"), (1, "RBDBCOO-nnz"), (2, "BDBCOO-block"),
                            (2, "BDBCOO-nnz")):
        L = orc.ref(dirname, 2)
        for (nr, nc, nnz, horz, vert) in shapes:
            r, c, v = sorted_coo(rng, nr, nc, nnz)
            dims = np.zeros(3, np.uint32)
            L.ref2_build(U(nr), U(nc), U(nnz), P(r), P(c), P(v), Ic(horz), Ic(vert), U(4), U(4), P(dims))
            nt, nb = int(dims[0]), int(dims[1])
            bi, bc, npb, bpp = (np.zeros(nb, np.uint32), np.zeros(nb, np.uint32), np.zeros(nb, np.uint32),
                                np.zeros(nt, np.uint32))
            bv = np.zeros(nb * 16, np.int32)
            L.ref2_copy_bcoo(P(bi), P(bc), P(bv), P(npb), P(bpp))
            arrays = dict(r=r, c=c, v=v, browind=bi, bcolind=bc, bval=bv, npb=npb, bpp=bpp)
            if tiling:
                blocks, nnzs = np.zeros(vert * horz, np.uint32), np.zeros(vert * horz, np.uint32)
                L.ref2_partition(P(blocks), P(nnzs), P(np.zeros(vert * (horz + 1), np.uint32)))
                arrays["blocks"] = blocks
            put({"kind": "tiled_bcoo", "tiling": tiling, "by_nnz": int(dirname.endswith("-nnz")), "nrows": nr,
                 "ncols": nc, "horz": horz, "vert": vert, "dims": [int(d) for d in dims]}, **arrays)

    # COO-nnz inter-DPU split with straddling rows (spmv/1D/COO-nnz/host/app.c:118-205)
    L = ctypes.CDLL(os.path.join(orc.REFDIR, "libref_coo_split.so"))
    for nrows, Pn in ((40, 7), (300, 16), (17, 200), (1000, 64)):
        for kind in ("uniform", "skew", "empty"):
            rows = row_lengths(rng, nrows, kind).astype(np.uint32)
            rows[0] += 1
            nnz = int(rows.sum())
            outs = [np.zeros(Pn, np.uint32) for _ in range(5)]
            tp = min(Pn, 8)
            ts, tc = np.zeros(tp * 16, np.uint32), np.zeros(tp * 16, np.uint32)
            L.ref_coo_nnz_split(U(nrows), U(nnz), P(rows), U(Pn), *[P(o) for o in outs], U(tp), P(ts), P(tc))
            put({"kind": "coo_split", "P": Pn, "nnz": nnz, "tsk_parts": tp}, rows=rows,
                **dict(zip(["prev_rows", "rows_per_part", "prev_nnz", "nnz_per_part", "merge"], outs)),
                tstart=ts, tcount=tc)

    np.savez_compressed(os.path.join(HERE, "host_surface.npz"), **ARR)
    with open(os.path.join(HERE, "host_surface.json"), "w") as f:
        json.dump(CASES, f, indent=0)
    print(f"{len(CASES)} cases, {sum(a.nbytes for a in ARR.values())} bytes")


if __name__ == "__main__":
    main()
