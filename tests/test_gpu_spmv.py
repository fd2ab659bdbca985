"""GPU SpMV and composites against the oracle and the reference's closed forms.

Tolerance: integer-valued cases (cdiag closed forms, unit values with iota
inputs) must be exact; FP64 cases with values and inputs in [0,1) must agree
with the oracle within a relative 1e-12 per element (every product is
non-negative, so the bound holds element-wise).
"""
import numpy as np
import pytest

import oracle as O
import tbsla_amd as T
from test_oracle import compute_spmv_res_cdiag

pytestmark = pytest.mark.gpu
RTOL = 1e-12

SHAPES = ([(10, 10, c) for c in range(13)] + [(5, 10, c) for c in range(13)] +
          [(10, 5, c) for c in range(13)] + [(30, 30, 2 * c) for c in range(13)] +
          [(20, 30, 2 * c) for c in range(13)] + [(30, 20, 2 * c) for c in range(13)])
FMTS = [T.COO, T.SCOO, T.CSR, T.ELL, T.DENSE]


def close(a, b, rtol=RTOL):
    a, b = np.asarray(a), np.asarray(b)
    assert a.shape == b.shape
    err = np.abs(a - b)
    assert np.all(err <= rtol * np.abs(b) + 1e-300), float(np.max(err / (np.abs(b) + 1e-300)))


@pytest.mark.parametrize("fmt", FMTS)
def test_spmv_cdiag_closed_form(fmt):
    """src/cpp/tests/test_spmv_cdiag.cpp on the GPU: generated and converted
    from COO, spmv(iota) equals the closed form exactly."""
    for (nr, nc, c) in SHAPES:
        v = np.arange(nc, dtype=np.float64)
        exp = compute_spmv_res_cdiag(nr, nc, c, v)
        m = T.Matrix(fmt).fill_cdiag(nr, nc, c)
        assert np.array_equal(m.spmv(v), exp), (fmt, nr, nc, c)
        if fmt != T.COO:
            m2 = T.Matrix(T.COO).fill_cdiag(nr, nc, c).to(fmt)
            assert np.array_equal(m2.spmv(v), exp), (fmt, nr, nc, c, "from COO")


@pytest.mark.parametrize("fmt", FMTS)
def test_Ax_cdiag_closed_form(fmt):
    """src/cpp/tests/test_Ax_cdiag.cpp: Ax into a pre-zeroed r."""
    for (nr, nc, c) in SHAPES[::3]:
        v = np.arange(nc, dtype=np.float64)
        m = T.Matrix(fmt).fill_cdiag(nr, nc, c)
        assert np.array_equal(m.Ax(np.zeros(nr), v), compute_spmv_res_cdiag(nr, nc, c, v))


@pytest.mark.parametrize("fmt", FMTS)
def test_a_axpx_cdiag_closed_form(fmt):
    """src/cpp/tests/test_a_axpx__cdiag.cpp and test_AAxpAx_cdiag.cpp."""
    for (n, c) in [(10, i) for i in range(13)] + [(30, 2 * i) for i in range(13)]:
        v = 2 * np.arange(n, dtype=np.float64) + 1
        exp = compute_spmv_res_cdiag(n, n, c, compute_spmv_res_cdiag(n, n, c, v) + v)
        m = T.Matrix(fmt).fill_cdiag(n, n, c)
        assert np.array_equal(m.a_axpx_(v), exp), (fmt, n, c)
        r, _ = m.AAxpAx(np.zeros(n), v, np.zeros(n))
        assert np.array_equal(r, exp), (fmt, n, c)
        r, _ = m.AAxpAxpx(np.zeros(n), v, np.zeros(n))
        assert np.array_equal(r, exp + v), (fmt, n, c)


@pytest.mark.parametrize("fmt", FMTS)
def test_spmv_cqmat_unit_values_exact(fmt):
    """Reference values (all 1.0) and iota inputs: every result is an exact
    integer, so GPU and oracle agree bit for bit whatever the sum order."""
    for (nr, nc, c) in SHAPES[::2]:
        for q in (0.0, 0.4, 1.0):
            for (pr, pc, NR, NC) in ((0, 0, 1, 1), (1, 1, 2, 2), (0, 1, 1, 3)):
                o = O.fill_cqmat(fmt, nr, nc, c, q, 1, pr, pc, NR, NC)
                g = T.Matrix(fmt).fill_cqmat(nr, nc, c, q, 1, pr, pc, NR, NC)
                if fmt == T.DENSE and not o.s.values:
                    continue
                geo = o.geom
                v = np.arange(geo["n_col"], dtype=np.float64)[geo["f_col"]:]
                assert np.array_equal(g.spmv(v), o.spmv(v)), (fmt, nr, nc, c, q, pr, pc)


@pytest.mark.parametrize("fmt", FMTS)
def test_spmv_cqmat_fp64_values(fmt):
    for (n, c, q, grid) in [(1000, 8, 0.1, (0, 0, 1, 1)), (5000, 32, 0.3, (0, 0, 1, 1)),
                            (4000, 16, 0.1, (1, 2, 4, 4)), (3000, 70, 0.05, (0, 0, 1, 1)),
                            (2048, 200, 0.5, (1, 0, 2, 1))]:
        if fmt == T.DENSE and n > 4000:
            continue
        o = O.fill_cqmat(fmt, n, n, c, q, 0, *grid, value_seed=0)
        g = T.Matrix(fmt).fill_cqmat(n, n, c, q, 0, *grid, value_seed=0)
        geo = o.geom
        v = O.unit_vector(1, geo["n_col"])[geo["f_col"]:] if fmt != T.COO else \
            O.unit_vector(1, geo["n_col"])
        if fmt == T.COO:
            v = O.unit_vector(1, n)
        close(g.spmv(v), o.spmv(v))


@pytest.mark.parametrize("fmt", [T.COO, T.SCOO, T.DENSE])
def test_Ax_accumulates_and_vect_incr(fmt):
    """COO/SCOO/DENSE add into r (MatrixCOO.cpp:140, MatrixSCOO.cpp:162,
    MatrixDENSE.cpp:83); COO/SCOO shift rows by vect_incr."""
    n = 400
    o = O.fill_cqmat(fmt, n, n, 9, 0.2, 1, value_seed=3)
    g = T.Matrix(fmt).fill_cqmat(n, n, 9, 0.2, 1, value_seed=3)
    v = O.unit_vector(5, n)
    r0 = O.unit_vector(6, n + 8)
    incr = 5 if fmt != T.DENSE else 0
    close(g.Ax(r0, v, incr)[: n + 8], o.Ax(r0, v, incr)[: n + 8])


@pytest.mark.parametrize("fmt", [T.CSR, T.ELL])
def test_Ax_overwrites(fmt):
    n = 300
    o = O.fill_cqmat(fmt, n, n, 7, 0.2, 1, value_seed=3)
    g = T.Matrix(fmt).fill_cqmat(n, n, 7, 0.2, 1, value_seed=3)
    v = O.unit_vector(5, n)
    close(g.Ax(np.full(n, 123.0), v), o.spmv(v))


def test_coo_unsorted_import_and_long_rows():
    """Triplets out of row order take the atomic path; long rows crossing
    many 512-entry chunks take the deterministic fix-up path."""
    rng = np.random.default_rng(0)
    n, nnz = 3000, 60000
    rows = np.sort(rng.integers(0, 40, nnz)).astype(np.int32)  # few, very long rows
    cols = rng.integers(0, n, nnz).astype(np.int32)
    vals = rng.random(nnz)
    geom = dict(n_row=n, n_col=n, f_row=0, f_col=0, ln_row=n, ln_col=n, pr=0, pc=0, NR=1, NC=1,
                nnz=nnz)
    v = rng.random(n)
    oc = O.coo_from_arrays(n, n, rows, cols, vals)
    g = T.Matrix.from_arrays(T.COO, geom, rows, cols, vals)
    close(g.spmv(v), oc.spmv(v))
    d1 = g.spmv(v)
    assert np.array_equal(d1, g.spmv(v))  # deterministic
    perm = rng.permutation(nnz)
    g2 = T.Matrix.from_arrays(T.COO, geom, rows[perm], cols[perm], vals[perm])
    close(g2.spmv(v), oc.spmv(v), rtol=1e-11)


@pytest.mark.parametrize("fmt", [T.COO, T.SCOO])
def test_coo_row_pointer_path_gaps_and_refill(fmt):
    """Row-sorted triplets go through a derived row pointer and the CSR row
    kernel: rows before the first and after the last entry, empty rows in
    between, a block with f_row > 0, and a refill (the row pointer is
    rebuilt) all match the oracle exactly on integer data."""
    rng = np.random.default_rng(2)
    n = 3000
    rows = np.sort(rng.choice(np.arange(700, 2300), 9000)).astype(np.int32)
    rows = rows[(rows < 1000) | (rows > 1400)]  # a gap of 400 empty rows
    cols = rng.integers(0, n, rows.size).astype(np.int32)
    vals = rng.integers(-3, 4, rows.size).astype(np.float64)
    v = np.arange(n, dtype=np.float64)
    geom = dict(n_row=n, n_col=n, f_row=0, f_col=0, ln_row=n, ln_col=n, pr=0, pc=0, NR=1, NC=1,
                nnz=int(rows.size))
    g = T.Matrix.from_arrays(fmt, geom, rows, cols, vals)
    oc = O.coo_from_arrays(n, n, rows, cols, vals)
    assert np.array_equal(g.spmv(v), oc.spmv(v))
    for q, grid in ((0.3, (1, 1, 3, 2)), (0.0, (0, 0, 1, 1))):
        g.fill_cqmat(2000, 2000, 12, q, 1, *grid)
        o = O.fill_cqmat(fmt, 2000, 2000, 12, q, 1, *grid)
        geo = o.geom
        x = np.arange(geo["n_col"], dtype=np.float64)[geo["f_col"]:] if fmt == T.SCOO else \
            np.arange(2000, dtype=np.float64)
        assert np.array_equal(g.spmv(x), o.spmv(x)), (fmt, grid)


def test_cfg1_ten_spmvs():
    """Configuration 1: n=100000, c=10, q=0.1, seed 0, values U[0,1) (seed 0),
    x U[0,1) (seed 1); ten repeated SpMVs, each within 1e-12 of the oracle."""
    n = 100000
    o = O.fill_cqmat(O.CSR, n, n, 10, 0.1, 0, value_seed=0)
    g = T.Matrix(T.CSR).fill_cqmat(n, n, 10, 0.1, 0, value_seed=0)
    x = O.unit_vector(1, n)
    dx = T.DeviceBuffer(n * 8)
    T.check(T.lib().tbsla_hip_fill_unit_vector(dx.ptr, n, 1, None))
    assert np.array_equal(dx.to_numpy(), x)
    y_ref = o.spmv(x)
    dy = T.DeviceBuffer(n * 8)
    for _ in range(10):
        g.spmv_device(dy, dx)
        close(dy.to_numpy(), y_ref)


@pytest.mark.parametrize("lanes", [1, 2, 4, 8, 16, 32, 64])
def test_csr_schedules_agree(lanes):
    """Every (lanes, rows) schedule gives the oracle's result within 1e-12,
    on regular rows and on rows of very different lengths."""
    n = 20000
    o = O.fill_cqmat(O.CSR, n, n, 40, 0.2, 0, value_seed=2)
    g = T.Matrix(T.CSR).fill_cqmat(n, n, 40, 0.2, 0, value_seed=2)
    rng = np.random.default_rng(lanes)
    lens = rng.integers(0, 300, 3000)
    rp = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
    cols = rng.integers(0, 3000, rp[-1]).astype(np.int32)
    vals = rng.random(rp[-1])
    geom = dict(n_row=3000, n_col=3000, f_row=0, f_col=0, ln_row=3000, ln_col=3000, pr=0, pc=0,
                NR=1, NC=1, nnz=int(rp[-1]))
    irr = T.Matrix.from_arrays(T.CSR, geom, rp, cols, vals)
    oirr = O.csr_from_arrays(3000, 3000, rp, cols, vals)
    x = O.unit_vector(1, n)
    for rows in (1, 2, 4):
        g.set_csr_schedule(lanes, rows)
        close(g.spmv(x), o.spmv(x))
        irr.set_csr_schedule(lanes, rows)
        close(irr.spmv(x[:3000]), oirr.spmv(x[:3000]))


VEC_KERNELS = ([(1, L, R, nt) for L in (1, 4, 16, 64) for R in (1, 4) for nt in (0, 1)]
               + [(2, L, R, nt) for L in (4, 8, 16, 32) for R in (1, 2) for nt in (0, 1)]
               + [(4, L, R, nt) for L in (4, 8, 16) for R in (1, 2) for nt in (0, 1)])


@pytest.mark.parametrize("vec,lanes,rows,nt", VEC_KERNELS)
def test_csr_vector_kernels_agree(vec, lanes, rows, nt):
    """Every compiled row kernel (entries per lane load, lanes, rows,
    non-temporal hint) stays within 1e-12 of the oracle: regular cqmat rows,
    rows of very different lengths (starts not aligned to the vector width),
    a block with f_col > 0, and the split layout."""
    n = 6000
    o = O.fill_cqmat(O.CSR, n, n, 40, 0.2, 0, value_seed=2)
    g = T.Matrix(T.CSR).fill_cqmat(n, n, 40, 0.2, 0, value_seed=2)
    x = O.unit_vector(1, n)
    g.set_csr_kernel(lanes, rows, vec, nt)
    assert g.get_csr_kernel() == {"lanes": lanes, "rows": rows, "vec": vec, "nontemporal": nt}
    y = g.spmv(x)
    close(y, o.spmv(x))
    assert np.array_equal(y, g.spmv(x))
    g.set_csr_layout(T.CSR_SPLIT, 64, 8)
    close(g.spmv(x), o.spmv(x))
    rng = np.random.default_rng(vec * 100 + lanes)
    lens = rng.integers(0, 97, 1500)
    rp = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
    cols = rng.integers(0, 1500, rp[-1]).astype(np.int32)
    vals = rng.random(rp[-1])
    geom = dict(n_row=1500, n_col=1500, f_row=0, f_col=0, ln_row=1500, ln_col=1500, pr=0, pc=0,
                NR=1, NC=1, nnz=int(rp[-1]))
    irr = T.Matrix.from_arrays(T.CSR, geom, rp, cols, vals)
    irr.set_csr_kernel(lanes, rows, vec, nt)
    close(irr.spmv(x[:1500]), O.csr_from_arrays(1500, 1500, rp, cols, vals).spmv(x[:1500]))
    ob = O.fill_cqmat(O.CSR, n, n, 12, 0.3, 0, 1, 1, 2, 2, value_seed=4)
    gb = T.Matrix(T.CSR).fill_cqmat(n, n, 12, 0.3, 0, 1, 1, 2, 2, value_seed=4)
    gb.set_csr_kernel(lanes, rows, vec, nt)
    xb = O.unit_vector(5, ob.geom["ln_col"])
    close(gb.spmv(xb), ob.spmv(xb))


def test_autotune_keeps_results():
    n = 50000
    o = O.fill_cqmat(O.CSR, n, n, 24, 0.1, 0, value_seed=2)
    g = T.Matrix(T.CSR).fill_cqmat(n, n, 24, 0.1, 0, value_seed=2)
    x = O.unit_vector(1, n)
    dx, dy = T.DeviceBuffer.from_numpy(x), T.DeviceBuffer(n * 8)
    us = g.autotune(dy, dx, 3)
    L, R = g.get_csr_schedule()
    assert us > 0 and L in (1, 2, 4, 8, 16, 32, 64) and R in (1, 2, 4)
    k = g.get_csr_kernel()
    assert k["vec"] in (1, 2, 4) and k["nontemporal"] in (0, 1)
    close(g.spmv(x), o.spmv(x))


@pytest.mark.parametrize("fmt", [T.CSR, T.ELL, T.COO, T.SCOO])
def test_spmv_config2_sampled_rows(fmt):
    """n = 2^22, c = 32, q = 0.1, values and x in [0,1): the full GPU SpMV on
    the rows of three oracle row blocks (start, middle, tail)."""
    n, c, q = 4194304, 32, 0.1
    g = T.Matrix(fmt).fill_cqmat(n, n, c, q, 0, value_seed=0)
    x = O.unit_vector(1, n)
    dx = T.DeviceBuffer.from_numpy(x)
    y = g.spmv(dx)
    for pr in (0, 37, 255):
        blk = O.fill_cqmat(O.CSR, n, n, c, q, 0, pr, 0, 256, 1, value_seed=0)
        geo = blk.geom
        close(y[geo["f_row"]:geo["f_row"] + geo["ln_row"]], blk.spmv(x))


@pytest.mark.parametrize("fmt", [T.CSR, T.ELL, T.COO, T.SCOO])
def test_spmv_config2_diag_layout_sampled_rows(fmt):
    """The same matrix in the diagonal layout (32 stored diagonals, the
    13.4 M rewired entries multiplied through LDS column slices): the rows of
    three oracle row blocks, and the same bits on a second SpMV."""
    n, c, q = 4194304, 32, 0.1
    g = T.Matrix(fmt).fill_cqmat(n, n, c, q, 0, value_seed=0)
    g.set_csr_layout(T.CSR_DIAG, 65536, 16)
    x = O.unit_vector(1, n)
    dx = T.DeviceBuffer.from_numpy(x)
    y = g.spmv(dx)
    assert np.array_equal(y, g.spmv(dx))
    for pr in (0, 37, 255):
        blk = O.fill_cqmat(O.CSR, n, n, c, q, 0, pr, 0, 256, 1, value_seed=0)
        geo = blk.geom
        close(y[geo["f_row"]:geo["f_row"] + geo["ln_row"]], blk.spmv(x))


def test_dense_slice_of_config2():
    """DENSE block (0,0) of a 256 x 256 grid of the n = 2^22 cqmat:
    16384 x 16384."""
    n = 4194304
    g = T.Matrix(T.DENSE).fill_cqmat(n, n, 32, 0.1, 0, 0, 0, 256, 256, value_seed=0)
    o = O.fill_cqmat(O.CSR, n, n, 32, 0.1, 0, 0, 0, 256, 256, value_seed=0)
    assert g.geom["ln_row"] == 16384 and g.geom["nnz"] == o.geom["nnz"]
    x = O.unit_vector(1, 16384)
    close(g.spmv(x), o.spmv(x))
