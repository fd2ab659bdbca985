"""The split and diagonal CSR execution layouts (tbsla_hip_set_csr_layout)
on the GPU.

SpMV through either layout stays within a relative 1e-12 (max-norm) of the
oracle for every window / slice count / block position, is bitwise
repeatable, is rebuilt by normalize_cols, and is dropped when the arrays
change.  The diagonal layout keeps explicit zeros and duplicate entries
out of its diagonals, so Inf and NaN in x reach exactly the rows they reach
in the row-major sum.
"""
import numpy as np
import pytest

import oracle as O
import tbsla_amd as T

pytestmark = pytest.mark.gpu


def rel_max(a, b):
    d = np.max(np.abs(b))
    return float(np.max(np.abs(a - b)) / (d if d > 0 else 1.0))


CASES = [
    (20000, 16, 0.1, (0, 0, 1, 1), 64, 8),
    (20000, 16, 0.1, (0, 0, 1, 1), 0, 16),
    (20000, 16, 1.0, (0, 0, 1, 1), 64, 64),
    (30000, 32, 0.1, (0, 0, 4, 4), 64, 8),
    (30000, 32, 0.1, (0, 1, 4, 4), 64, 8),
    (30000, 9, 0.3, (3, 2, 4, 4), 5, 8),
    (5000, 8, 0.0, (0, 0, 1, 1), 1 << 30, 8),
    (1, 1, 0.5, (0, 0, 1, 1), 64, 8),
]


@pytest.mark.parametrize("n,c,q,grid,window,slices", CASES)
def test_split_spmv_matches_oracle(n, c, q, grid, window, slices):
    o = O.fill_cqmat(O.CSR, n, n, c, q, 0, *grid, value_seed=3)
    g = T.Matrix(T.CSR).fill_cqmat(n, n, c, q, 0, *grid, value_seed=3)
    x = O.unit_vector(7, o.geom["ln_col"])
    rows = g.spmv(x)
    g.set_csr_layout(T.CSR_SPLIT, window, slices)
    lay = g.get_csr_layout()
    assert lay["layout"] == T.CSR_SPLIT and lay["slices"] == slices
    y1, y2 = g.spmv(x), g.spmv(x)
    ref = o.spmv(x)
    assert rel_max(y1, ref) < 1e-12
    assert np.array_equal(y1, y2)
    g.set_csr_layout(T.CSR_ROWS)
    assert np.array_equal(g.spmv(x), rows)


@pytest.mark.parametrize("n,c,q,grid,window,slices", CASES)
def test_diag_spmv_matches_oracle(n, c, q, grid, window, slices):
    window = min(window, 65536)
    o = O.fill_cqmat(O.CSR, n, n, c, q, 0, *grid, value_seed=3)
    g = T.Matrix(T.CSR).fill_cqmat(n, n, c, q, 0, *grid, value_seed=3)
    x = O.unit_vector(7, o.geom["ln_col"])
    rows = g.spmv(x)
    g.set_csr_layout(T.CSR_DIAG, window, slices)
    lay = g.get_csr_layout()
    assert lay["layout"] == T.CSR_DIAG and lay["slices"] == slices
    y1, y2 = g.spmv(x), g.spmv(x)
    assert rel_max(y1, o.spmv(x)) < 1e-12
    assert np.array_equal(y1, y2)
    g.set_csr_layout(T.CSR_ROWS)
    assert np.array_equal(g.spmv(x), rows)


# The diagonal layout's remote form: LDS slices with 4096-row blocks (n =
# 100000: 7 column slices, 25 blocks, a short last slice and block); 128-row
# blocks (64 remote entries a row: 256 rows would hold 16384 > 14336); no
# block size fits (300 remote entries a row: 64 rows hold 19200), so the
# L2-slice form; the same matrix as a block of a 3 x 3 grid (f_col > 0).
@pytest.mark.parametrize("n,c,q,grid", [
    (100000, 16, 0.1, (0, 0, 1, 1)),
    (20000, 64, 1.0, (0, 0, 1, 1)),
    (20000, 300, 1.0, (0, 0, 1, 1)),
    (100000, 16, 0.1, (1, 2, 3, 3)),
])
def test_diag_remote_forms_match_oracle(n, c, q, grid):
    o = O.fill_cqmat(O.CSR, n, n, c, q, 0, *grid, value_seed=11)
    g = T.Matrix(T.CSR).fill_cqmat(n, n, c, q, 0, *grid, value_seed=11)
    x = O.unit_vector(5, o.geom["ln_col"]) - 0.5
    g.set_csr_layout(T.CSR_DIAG, 65536, 16)
    y1, y2 = g.spmv(x), g.spmv(x)
    assert rel_max(y1, o.spmv(x)) < 1e-12
    assert np.array_equal(y1, y2)
    # Ax (MatrixCSR::Ax overwrites its result)
    r = O.unit_vector(9, o.geom["ln_row"])
    assert rel_max(g.Ax(r.copy(), x), o.Ax(r.copy(), x)) < 1e-12


def test_diag_remote_count_zeros_duplicates_and_nonfinite_x():
    """Rows r of a 3000 x 3000 block: (r, r) = 1 + r/7; (r, r+1) = 0.0 for
    odd r, 2.5 otherwise; a second (r, r) = -0.5 for r % 3 == 0; (r, r+500)
    for r % 5 == 0.  The diagonals 0 and +1 are stored (+500 holds a fifth
    of the rows, below the quarter threshold); the explicit zeros, the
    second diagonal entries and the +500 entries are remote."""
    n = 3000
    rp, ci, va = [0], [], []
    for r in range(n):
        ent = [(r, 1.0 + r / 7.0)]
        if r + 1 < n:
            ent.append((r + 1, 0.0 if r % 2 else 2.5))
        if r % 3 == 0:
            ent.append((r, -0.5))
        if r % 5 == 0 and r + 500 < n:
            ent.append((r + 500, 0.25))
        for c, v in ent:
            ci.append(c)
            va.append(v)
        rp.append(len(ci))
    rp, ci, va = np.array(rp, np.int32), np.array(ci, np.int32), np.array(va)
    o = O.csr_from_arrays(n, n, rp, ci, va)
    g = T.Matrix.from_arrays(T.CSR, o.geom, rp, ci, va)
    g.set_csr_layout(T.CSR_DIAG, 1000, 8)
    zeros = int(np.sum(va == 0.0))
    dups = len(range(0, n, 3))
    far = len([r for r in range(0, n, 5) if r + 500 < n])
    assert g.get_csr_layout()["remote"] == zeros + dups + far
    for bad in (np.inf, -np.inf, np.nan):
        x = O.unit_vector(4, n) + 0.5
        x[[17, 1200, n - 1]] = bad
        want, got = o.spmv(x), g.spmv(x)
        assert np.array_equal(np.isnan(got), np.isnan(want))
        assert np.array_equal(np.isinf(got), np.isinf(want))
        fin = np.isfinite(want)
        assert rel_max(got[fin], want[fin]) < 1e-12


def test_split_remote_count_and_normalize():
    n = 40000
    o = O.fill_cqmat(O.CSR, n, n, 16, 0.1, 0, value_seed=1)
    g = T.Matrix(T.CSR).fill_cqmat(n, n, 16, 0.1, 0, value_seed=1)
    g.set_csr_layout(T.CSR_SPLIT, 64, 8)
    rp, ci, _ = o.arrays()
    rows = np.repeat(np.arange(n), np.diff(rp))
    assert g.get_csr_layout()["remote"] == int(np.sum(np.abs(ci - rows) > 64))
    g.make_stochastic()
    o.make_stochastic()
    assert g.get_csr_layout()["layout"] == T.CSR_SPLIT
    x = O.unit_vector(2, n)
    assert rel_max(g.spmv(x), o.spmv(x)) < 1e-12
    b, it, _ = g.page_rank(0.85, 1e-10, 200)
    bo, ito, _ = o.page_rank(0.85, 1e-10, 200)
    assert it == ito and rel_max(b, bo) < 1e-12
    # new arrays drop the layout
    g.fill_cqmat(n, n, 16, 0.1, 0, value_seed=1)
    assert g.get_csr_layout()["layout"] == T.CSR_ROWS


@pytest.mark.parametrize("layout", [T.CSR_SPLIT, T.CSR_DIAG])
@pytest.mark.parametrize("grid", [(0, 0, 1, 1), (1, 0, 2, 2)])
def test_split_ell_matches_oracle(grid, layout):
    n = 20000
    o = O.fill_cqmat(O.ELL, n, n, 16, 0.1, 0, *grid, value_seed=5)
    g = T.Matrix(T.ELL).fill_cqmat(n, n, 16, 0.1, 0, *grid, value_seed=5)
    x = O.unit_vector(3, o.geom["ln_col"])
    g.set_csr_layout(layout, 64, 8)
    y = g.spmv(x)
    assert rel_max(y, o.spmv(x)) < 1e-12
    assert np.array_equal(y, g.spmv(x))


def test_split_bad_arguments():
    g = T.Matrix(T.CSR).fill_cqmat(100, 100, 4, 0.1, 0)
    for w, s in [(-1, 8), (64, 0), (64, 12), (64, 128)]:
        with pytest.raises(T.TbslaHipError):
            g.set_csr_layout(T.CSR_SPLIT, w, s)
    with pytest.raises(T.TbslaHipError):
        g.set_csr_layout(T.CSR_DIAG, (1 << 20) + 1, 8)
    with pytest.raises(T.TbslaHipError):
        g.set_csr_layout(3, 64, 8)
    e = T.Matrix(T.DENSE).fill_cqmat(100, 100, 4, 0.1, 0)
    with pytest.raises(T.TbslaHipError):
        e.set_csr_layout(T.CSR_SPLIT, 64, 8)


def coo_x(o, seed):
    """x as the block's SpMV reads it: COO by the global column (n_col
    entries), SCOO by the column less f_col (ln_col entries)."""
    geo = o.geom
    return O.unit_vector(seed, geo["n_col"] if o.fmt == O.COO else geo["ln_col"]) - 0.25


@pytest.mark.parametrize("layout", [T.CSR_SPLIT, T.CSR_DIAG])
@pytest.mark.parametrize("fmt", [T.COO, T.SCOO])
@pytest.mark.parametrize("grid", [(0, 0, 1, 1), (1, 1, 2, 2), (1, 0, 2, 2), (2, 1, 3, 3)])
def test_coo_layouts_match_oracle(fmt, layout, grid):
    """COO and SCOO blocks in the derived layouts over their triplets' rows
    (rows from the first to the last row holding a triplet; y is zero
    elsewhere): SpMV within 1e-12 of the oracle and bitwise repeatable; on a
    block of the matrix's diagonal the stored diagonals are the matrix's (the
    rows shifted by the block's first row), so few entries stay remote; Ax
    (accumulating for COO and SCOO, MatrixCOO.cpp / MatrixSCOO.cpp) keeps
    its own path; a refill drops the layout."""
    n = 30000
    o = O.fill_cqmat(fmt, n, n, 16, 0.1, 0, *grid, value_seed=6)
    g = T.Matrix(fmt).fill_cqmat(n, n, 16, 0.1, 0, *grid, value_seed=6)
    x = coo_x(o, 4)
    before = g.spmv(x)
    g.set_csr_layout(layout, 65536 if layout == T.CSR_DIAG else 64, 8)
    lay = g.get_csr_layout()
    assert lay["layout"] == layout
    y1, y2 = g.spmv(x), g.spmv(x)
    ref = o.spmv(x)
    assert rel_max(y1, ref) < 1e-12
    assert np.array_equal(y1, y2)
    if layout == T.CSR_DIAG and grid[0] == grid[1]:
        assert lay["remote"] < 0.2 * o.geom["nnz"], lay
    r = O.unit_vector(9, ref.size)
    assert rel_max(g.Ax(r.copy(), x), o.Ax(r.copy(), x)) < 1e-12
    g.set_csr_layout(T.CSR_ROWS)
    assert rel_max(g.spmv(x), before) < 1e-15
    g.set_csr_layout(layout, 65536 if layout == T.CSR_DIAG else 64, 8)
    g.fill_cqmat(n, n, 16, 0.1, 0, *grid, value_seed=6)
    assert g.get_csr_layout()["layout"] == T.CSR_ROWS


@pytest.mark.parametrize("fmt", [T.COO, T.SCOO])
def test_coo_layout_gaps_values_changed_and_unsorted(fmt):
    """Row-sorted triplets with empty rows before, between and after them
    (integer data: exact) in the diagonal layout; new values written through
    the device arrays and announced by values_changed rebuild the layout;
    triplets out of row order refuse a layout (UNSUPPORTED) and keep their
    atomic path."""
    rng = np.random.default_rng(3)
    n = 4000
    rows = np.sort(rng.choice(np.arange(500, 3500), 30000)).astype(np.int32)
    rows = rows[(rows < 1500) | (rows > 1800)]
    cols = np.clip(rows + rng.integers(-3, 4, rows.size), 0, n - 1).astype(np.int32)
    cols[::17] = rng.integers(0, n, cols[::17].size)
    vals = rng.integers(-3, 4, rows.size).astype(np.float64)
    geom = dict(n_row=n, n_col=n, f_row=0, f_col=0, ln_row=n, ln_col=n, pr=0, pc=0, NR=1, NC=1,
                nnz=int(rows.size))
    g = T.Matrix.from_arrays(fmt, geom, rows, cols, vals)
    x = np.arange(n, dtype=np.float64) - 1000.0
    g.set_csr_layout(T.CSR_DIAG, 64, 8)
    assert g.get_csr_layout()["layout"] == T.CSR_DIAG
    assert np.array_equal(g.spmv(x), O.coo_from_arrays(n, n, rows, cols, vals).spmv(x))
    # new values through the device arrays
    _, _, dvals = g.device_arrays()
    vals2 = vals * 2.0 - 1.0
    T.check(T.lib().tbsla_hip_memcpy_htod(dvals, vals2.ctypes.data, vals2.nbytes, None))
    g.values_changed()
    assert g.get_csr_layout()["layout"] == T.CSR_DIAG
    assert np.array_equal(g.spmv(x), O.coo_from_arrays(n, n, rows, cols, vals2).spmv(x))
    perm = rng.permutation(rows.size)
    u = T.Matrix.from_arrays(fmt, geom, rows[perm], cols[perm], vals[perm])
    with pytest.raises(T.TbslaHipError):
        u.set_csr_layout(T.CSR_DIAG, 64, 8)
    assert u.get_csr_layout()["layout"] == T.CSR_ROWS
    ref = O.coo_from_arrays(n, n, rows, cols, vals).spmv(x)
    assert rel_max(u.spmv(x), ref) < 1e-14


@pytest.mark.parametrize("fmt", [T.COO, T.SCOO])
def test_coo_autotune_keeps_results(fmt):
    """The autotuner on a COO / SCOO block (n = 200000, c = 24, q = 0.1)
    leaves a layout whose SpMV is within 1e-12 of the oracle."""
    n = 200000
    o = O.fill_cqmat(fmt, n, n, 24, 0.1, 0, value_seed=2)
    g = T.Matrix(fmt).fill_cqmat(n, n, 24, 0.1, 0, value_seed=2)
    us = g.tune_layout(3)
    assert us > 0
    assert g.get_csr_layout()["layout"] in (T.CSR_ROWS, T.CSR_SPLIT, T.CSR_DIAG)
    x = coo_x(o, 1)
    assert rel_max(g.spmv(x), o.spmv(x)) < 1e-12
