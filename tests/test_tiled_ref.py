"""2D tilings and stripe partitioners vs the reference's spmv/2D support
code compiled into oracle/_ref (CPU only; skipped when it was not built)."""
import ctypes

import numpy as np
import pytest

import oracle_lib as orc

pytestmark = pytest.mark.skipif(not orc.have_refs(), reason="oracle/_ref not built")

V, U = ctypes.c_void_p, ctypes.c_uint32


def P(a):
    return V(a.ctypes.data)


def sorted_coo(sp, rng, nrows, ncols, nnz, dt, nonzero=True):
    keys = np.sort(rng.choice(nrows * ncols, size=nnz, replace=False))
    r, c = (keys // ncols).astype(np.uint32), (keys % ncols).astype(np.uint32)
    if dt == 2:
        v = rng.integers(1, 5, size=nnz) * rng.choice([-1, 1], size=nnz)
        v = v.astype(np.int32)
    else:
        v = (rng.standard_normal(nnz) + 3.0).astype(np.float32)
    return r, c, v, sp.Coo.from_arrays(nrows, ncols, r, c, v)


SHAPES = [(64, 80, 900, 4, 2), (100, 37, 700, 3, 5), (200, 300, 3000, 8, 8), (33, 1000, 500, 2, 16)]


@pytest.mark.parametrize("dt", [2, 4])
def test_csr_tilings_match_reference(sp, dt):
    rng = np.random.default_rng(10 + dt)
    for tiling, dirname in ((sp.TILE_D, "DCSR"), (sp.TILE_RB, "RBDCSR"), (sp.TILE_BD, "BDCSR")):
        L = orc.ref(dirname, dt)
        for (nr, nc, nnz, horz, vert) in SHAPES:
            r, c, v, A = sorted_coo(sp, rng, nr, nc, nnz, dt)
            dims = np.zeros(3, np.uint32)
            L.ref2_build(U(nr), U(nc), U(nnz), P(r), P(c), P(v), ctypes.c_int(horz), ctypes.c_int(vert), P(dims))
            T = sp.coo2tiled(A, tiling, sp.FMT_CSR, horz, vert)
            assert (T.ntiles, T.tile_height, T.tile_width) == tuple(int(d) for d in dims), dirname
            rp = np.zeros(T.ntiles * (T.tile_height + 1), np.uint32)
            ci, vo, cuts = np.zeros(nnz, np.uint32), np.zeros(nnz, v.dtype), np.zeros(vert + 1, np.uint32)
            L.ref2_copy_csr(P(rp), P(ci), P(vo), P(cuts))
            assert np.array_equal(T.drowptr, rp), dirname
            assert np.array_equal(T.dcolind, ci) and np.array_equal(T.dval, vo), dirname
            if tiling == sp.TILE_BD:
                assert np.array_equal(T.vert_tile_widths, cuts), "BD column cut points"
            if tiling != sp.TILE_D:
                split = np.zeros(vert * (horz + 1), np.uint32)
                L.ref2_partition(P(split))
                assert np.array_equal(sp.partition_tiled(T, horz), split), (dirname, "row_split")


@pytest.mark.parametrize("dt", [2, 4])
def test_coo_tilings_match_reference(sp, dt):
    rng = np.random.default_rng(20 + dt)
    for tiling, dirname in ((sp.TILE_D, "DCOO"), (sp.TILE_RB, "RBDCOO"), (sp.TILE_BD, "BDCOO")):
        L = orc.ref(dirname, dt)
        for (nr, nc, nnz, horz, vert) in SHAPES:
            r, c, v, A = sorted_coo(sp, rng, nr, nc, nnz, dt)
            dims = np.zeros(3, np.uint32)
            L.ref2_build(U(nr), U(nc), U(nnz), P(r), P(c), P(v), ctypes.c_int(horz), ctypes.c_int(vert), P(dims))
            T = sp.coo2tiled(A, tiling, sp.FMT_COO, horz, vert)
            ri, ci, vo = np.zeros(nnz, np.uint32), np.zeros(nnz, np.uint32), np.zeros(nnz, v.dtype)
            L.ref2_copy_coo(P(ri), P(ci), P(vo))
            assert int(dims[0]) == T.ntiles
            assert np.array_equal(T.drowind, ri) and np.array_equal(T.dcolind, ci) and np.array_equal(T.dval, vo)
            if tiling != sp.TILE_D:
                split = np.zeros(vert * (horz + 1), np.uint32)
                L.ref2_partition(P(split))
                assert np.array_equal(sp.partition_tiled(T, horz), split), (dirname, "nnz_split")


@pytest.mark.parametrize("dt", [2, 4])
@pytest.mark.parametrize("rc", [(4, 4), (2, 3)])
def test_block_tilings_match_reference(sp, dt, rc):
    R, C = rc
    rng = np.random.default_rng(30 + dt + R)
    for tiling, dirname in ((sp.TILE_D, "DBCOO"), (sp.TILE_RB, "RBDBCOO-block"), (sp.TILE_RB, "RBDBCOO-nnz"),
                            (sp.TILE_BD, "BDBCOO-block"), (sp.TILE_BD, "BDBCOO-nnz")):
        L = orc.ref(dirname, dt)
        by_nnz = dirname.endswith("-nnz")
        for (nr, nc, nnz, horz, vert) in SHAPES:
            r, c, v, A = sorted_coo(sp, rng, nr, nc, nnz, dt)
            dims = np.zeros(3, np.uint32)
            L.ref2_build(U(nr), U(nc), U(nnz), P(r), P(c), P(v), ctypes.c_int(horz), ctypes.c_int(vert), U(R), U(C),
                         P(dims))
            T = sp.coo2tiled(A, tiling, sp.FMT_BCOO, horz, vert, R, C, count_values=True)
            nt, nb = int(dims[0]), int(dims[1])
            assert (T.ntiles, T.num_blocks, T.num_block_rows) == (nt, nb, int(dims[2])), dirname
            bi, bc, npb, bpp = (np.zeros(nb, np.uint32), np.zeros(nb, np.uint32), np.zeros(nb, np.uint32),
                                np.zeros(nt, np.uint32))
            bv = np.zeros(nb * R * C, v.dtype)
            L.ref2_copy_bcoo(P(bi), P(bc), P(bv), P(npb), P(bpp))
            assert np.array_equal(T.blocks_per_partition, bpp), dirname
            assert np.array_equal(T.browind, bi) and np.array_equal(T.bcolind, bc), dirname
            assert np.array_equal(T.bval, bv) and np.array_equal(T.nnz_per_block, npb), dirname
            if tiling != sp.TILE_D:
                blocks, nnzs = np.zeros(vert * horz, np.uint32), np.zeros(vert * horz, np.uint32)
                L.ref2_partition(P(blocks), P(nnzs), P(np.zeros(vert * (horz + 1), np.uint32)))
                split = sp.partition_tiled(T, horz, by_nnz).reshape(vert, horz + 1)
                assert np.array_equal(np.diff(split.astype(np.int64), axis=1).ravel(), blocks), (dirname, "blocks")


@pytest.mark.parametrize("by_nnz", [0, 1])
def test_rbd_bcsr_stripe_partition_matches_reference(sp, by_nnz):
    rng = np.random.default_rng(40 + by_nnz)
    L = orc.ref("RBDBCSR-nnz" if by_nnz else "RBDBCSR-block", 2)
    for (nr, nc, nnz, horz, vert) in SHAPES:
        r, c, v, A = sorted_coo(sp, rng, nr, nc, nnz, 2)
        dims = np.zeros(3, np.uint32)
        L.ref2_build(U(nr), U(nc), U(nnz), P(r), P(c), P(v), ctypes.c_int(horz), ctypes.c_int(vert), U(4), U(4),
                     P(dims))
        T = sp.coo2tiled(A, sp.TILE_RB, sp.FMT_BCSR, horz, vert, 4, 4, count_values=True)
        blocks, nnzs, bsplit = (np.zeros(vert * horz, np.uint32), np.zeros(vert * horz, np.uint32),
                                np.zeros(vert * (horz + 1), np.uint32))
        L.ref2_partition(P(blocks), P(nnzs), P(bsplit))
        assert np.array_equal(sp.partition_tiled(T, horz, by_nnz), bsplit)
