"""GPU generators against the oracle: every structure array bit-exact.

fill_cqmat / fill_cdiag on the GPU (tbsla_hip_fill_*) against the CPU
restatement on the reference's own test shapes (src/cpp/tests/
test_spmv_cqmat.cpp:162-198) over 2-D block grids, for all five formats,
plus conversions from COO and the configuration-sized structure.
"""
import numpy as np
import pytest

import oracle as O
import tbsla_amd as T

pytestmark = pytest.mark.gpu

SHAPES = ([(10, 10, c) for c in range(13)] + [(5, 10, c) for c in range(13)] +
          [(10, 5, c) for c in range(13)] + [(30, 30, 2 * c) for c in range(13)] +
          [(20, 30, 2 * c) for c in range(13)] + [(30, 20, 2 * c) for c in range(13)])
GRIDS = [(0, 0, 1, 1), (0, 1, 2, 2), (1, 0, 2, 2), (1, 1, 2, 2), (2, 1, 3, 2), (0, 2, 1, 3),
         (1, 0, 3, 1)]
FMTS = [O.COO, O.SCOO, O.CSR, O.ELL, O.DENSE]


def same(gpu, orc, fmt, ctx):
    go, gg = orc.geom, gpu.geom
    if fmt == O.DENSE and not orc.s.values:
        # nothing generated in the block: the reference leaves nnz
        # uninitialised there; the GPU reports 0
        go.pop("nnz")
        gg.pop("nnz")
    assert gg == go, ctx
    for a, b in zip(gpu.arrays(), orc.arrays()):
        if a.size == 0 and b.size == 0:
            continue
        assert a.shape == b.shape, ctx
        assert np.array_equal(a, b), ctx


@pytest.mark.parametrize("fmt", FMTS)
def test_fill_cqmat_matches_oracle(fmt):
    m = T.Matrix(fmt)
    for (nr, nc, c) in SHAPES:
        for q in (0.0, 0.1, 1.0):
            for s in (0, 3):
                for (pr, pc, NR, NC) in GRIDS:
                    ctx = (fmt, nr, nc, c, q, s, pr, pc, NR, NC)
                    m.fill_cqmat(nr, nc, c, q, s, pr, pc, NR, NC)
                    same(m, O.fill_cqmat(fmt, nr, nc, c, q, s, pr, pc, NR, NC), fmt, ctx)


@pytest.mark.parametrize("fmt", FMTS)
def test_fill_cqmat_values_match_oracle(fmt):
    """Project values V(seed, k) (uniform [0,1)) are bit-identical too."""
    m = T.Matrix(fmt)
    for (nr, nc, c, q, s, pr, pc, NR, NC) in [(30, 30, 6, 0.3, 1, 0, 0, 1, 1),
                                              (300, 280, 17, 0.5, 0, 1, 1, 2, 3),
                                              (1000, 1000, 32, 1.0, 2, 0, 0, 1, 1),
                                              (997, 1013, 9, 0.1, 0, 2, 1, 4, 2)]:
        m.fill_cqmat(nr, nc, c, q, s, pr, pc, NR, NC, value_seed=12345)
        same(m, O.fill_cqmat(fmt, nr, nc, c, q, s, pr, pc, NR, NC, value_seed=12345), fmt,
             (fmt, nr, nc))


@pytest.mark.parametrize("fmt", FMTS)
def test_fill_cdiag_matches_oracle(fmt):
    m = T.Matrix(fmt)
    for (nr, nc, c) in SHAPES:
        for (pr, pc, NR, NC) in GRIDS:
            m.fill_cdiag(nr, nc, c, pr, pc, NR, NC)
            same(m, O.fill_cdiag(fmt, nr, nc, c, pr, pc, NR, NC), fmt, (fmt, nr, nc, c, pr, pc))


@pytest.mark.parametrize("fmt", [O.SCOO, O.CSR, O.ELL, O.DENSE])
def test_convert_from_coo_matches_oracle(fmt):
    for (nr, nc, c) in SHAPES[::2] + [(2000, 1500, 24)]:
        for q in (0.0, 0.5, 1.0):
            for (pr, pc, NR, NC) in ((0, 0, 1, 1), (1, 0, 2, 2), (2, 1, 3, 2)):
                if fmt == O.DENSE and nr * nc > 10**6:
                    continue
                gcoo = T.Matrix(T.COO).fill_cqmat(nr, nc, c, q, 1, pr, pc, NR, NC, value_seed=7)
                ocoo = O.fill_cqmat(O.COO, nr, nc, c, q, 1, pr, pc, NR, NC, value_seed=7)
                same(gcoo.to(fmt), ocoo.to(fmt), fmt, (fmt, nr, nc, c, q, pr, pc))


def test_cfg1_structure_bit_exact():
    """n=100000, c=10, q=0.1, seed 0; values V(0, k)."""
    n, c, q = 100000, 10, 0.1
    g = T.Matrix(T.CSR).fill_cqmat(n, n, c, q, 0, value_seed=0)
    o = O.fill_cqmat(O.CSR, n, n, c, q, 0, value_seed=0)
    assert g.geom == o.geom
    assert g.geom["nnz"] == 1000000 and g.geometry().stored == 999955
    for a, b in zip(g.arrays(), o.arrays()):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("fmt", [T.CSR, T.ELL, T.SCOO])
def test_large_block_rows_equal_full_matrix_rows(fmt):
    """n = 2^22, c = 32, q = 0.1 (the SpMV configuration).  With NC = 1 a row
    block of a 64-block split holds exactly the full matrix's rows, so the
    full GPU fill must agree with oracle blocks at the start, the middle and
    the tail (where the reference's quirks live), and with the counts."""
    n, c, q = 4194304, 32, 0.1
    gnv = O.cqmat_gnv(n, n, c)
    assert gnv == 134217232
    full = T.Matrix(fmt).fill_cqmat(n, n, c, q, 0)
    g = full.geometry()
    if fmt == T.CSR:
        assert g.stored == gnv and g.nnz == gnv + c * (c - 1) // 2
    elif fmt == T.ELL:
        assert g.nnz == gnv and g.max_col == c
    else:
        assert g.nnz == gnv
    arrays = full.arrays()
    for pr in (0, 31, 63):
        blk = O.fill_cqmat(fmt, n, n, c, q, 0, pr, 0, 64, 1)
        bg = blk.geom
        lo, hi = bg["f_row"], bg["f_row"] + bg["ln_row"]
        ob = blk.arrays()
        if fmt == T.CSR:
            rp, ci, va = arrays
            assert np.array_equal(rp[lo:hi + 1] - rp[lo], ob[0])
            assert np.array_equal(ci[rp[lo]:rp[hi]], ob[1])
            assert np.array_equal(va[rp[lo]:rp[hi]], ob[2])
        elif fmt == T.ELL:
            assert np.array_equal(arrays[0][lo:hi], ob[0])
            assert np.array_equal(arrays[1][lo:hi], ob[1])
        else:
            r = arrays[0]
            a, b = np.searchsorted(r, lo), np.searchsorted(r, hi)
            for x, y in zip(arrays, ob):
                assert np.array_equal(x[a:b], y)
