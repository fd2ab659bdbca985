"""The oracle's row operations (get_row_sums, normalize_rows,
get_row_max_abs, apply_exponential; Matrix.hpp:46-51) pinned two ways:

* against the compiled reference (oracle/_ref, where it was built), on the
  reference's own cqmat blocks with signed values written into both objects,
  bit-exact, every format and several block grids;
* against what the mathematics fixes, with no reference needed: row sums of
  the block rebuilt as a dense array (numpy's sum, a different summation
  route), row sums of 1 after normalisation, exp(z - max|z|) <= 1 with
  equality at the largest entry, and numpy's exp/power elementwise.
"""
import numpy as np
import pytest

import oracle as O

needs_ref = pytest.mark.skipif(not O.ref_available(), reason="oracle/_ref not built here")

FORMATS = [O.CSR, O.ELL, O.SCOO, O.DENSE, O.COO]
CASES = [(30, 30, 6, 0.0), (30, 20, 10, 0.3), (20, 30, 4, 0.5), (200, 200, 9, 0.2)]
GRIDS = [(0, 0, 1, 1), (1, 1, 2, 2), (2, 0, 3, 2)]


def signed(n, seed):
    """Values in [-2, 2) with about one in eight exactly zero."""
    rng = np.random.default_rng(seed)
    v = rng.uniform(-2.0, 2.0, n)
    v[rng.random(n) < 0.125] = 0.0
    return v


def with_signed_values(m, seed):
    """Multiply the block's values (1 on entries, 0 on ELL/DENSE padding) by
    signed values, keeping padding at zero."""
    cur = m.arrays()[-1].reshape(-1)
    v = cur * signed(cur.size, seed)
    m.set_values(v)
    return v


def dense_block(m):
    """ln_row x ln_col array of the block, built from its arrays."""
    g = m.geom
    D = np.zeros((g["ln_row"], g["ln_col"]))
    arr = m.arrays()
    if m.fmt == O.CSR:
        rp, ci, va = arr
        rows = np.repeat(np.arange(g["ln_row"]), np.diff(rp))
        np.add.at(D, (rows, ci - g["f_col"]), va)
    elif m.fmt == O.ELL:
        cols, va = arr
        rows = np.repeat(np.arange(cols.shape[0]), cols.shape[1])
        keep = va.reshape(-1) != 0
        np.add.at(D, (rows[keep], cols.reshape(-1)[keep] - g["f_col"]), va.reshape(-1)[keep])
    elif m.fmt in (O.SCOO, O.COO):
        ro, co, va = arr
        np.add.at(D, (ro - g["f_row"], co - g["f_col"]), va)
    else:
        D[:] = arr[0]
    return D


def sums_len(fmt, g):
    return g["n_row"] if fmt in (O.ELL, O.SCOO) else g["ln_row"]


@needs_ref
@pytest.mark.parametrize("fmt", FORMATS)
def test_row_ops_match_reference(fmt):
    for k, (nr, nc, c, q) in enumerate(CASES):
        for (pr, pc, NR, NC) in GRIDS:
            o = O.fill_cqmat(fmt, nr, nc, c, q, 1, pr, pc, NR, NC)
            r = O.ref_fill_cqmat(fmt, nr, nc, c, q, 1, pr, pc, NR, NC)
            if not r._ptrs()[0]:
                # an empty block whose arrays the reference never allocated:
                # its CSR and DENSE row loops would read through null
                # pointers; the oracle reads such a block as all zeros
                assert o.geom["nnz"] == 0 or fmt == O.DENSE
                continue
            v = with_signed_values(o, 100 + k)
            r.set_values(v)
            g = o.geom
            n = max(g["n_row"], 1)
            s0 = np.random.default_rng(k).uniform(-1, 1, n) if fmt == O.SCOO else np.zeros(n)
            so = O.Matrix.get_row_sums(o, s0)
            sr = r.get_row_sums(s0)
            assert np.array_equal(so, sr), (fmt, nr, nc, c, q, pr, pc)
            # zero a few sums: CSR's safeguard, the others' division by zero
            s1 = so.copy()
            s1[::7] = 0.0
            o.normalize_rows(s1)
            r.normalize_rows(s1)
            assert np.array_equal(o.arrays()[-1], r.arrays()[-1], equal_nan=True)
            # max_abs and the exponential on a fresh block with signed values
            o = O.fill_cqmat(fmt, nr, nc, c, q, 1, pr, pc, NR, NC)
            r = O.ref_fill_cqmat(fmt, nr, nc, c, q, 1, pr, pc, NR, NC)
            v = with_signed_values(o, 200 + k)
            r.set_values(v)
            mo = o.get_row_max_abs()
            mr = r.get_row_max_abs(g["ln_row"])
            assert np.array_equal(mo, mr)
            for base, shift in ((-1, True), (0, False), (2, True), (3, False)):
                o.set_values(v)
                r.set_values(v)
                o.apply_exponential(mo if shift else None, base)
                r.apply_exponential(mr if shift else None, base)
                assert np.array_equal(o.arrays()[-1], r.arrays()[-1]), (fmt, base, shift)


@pytest.mark.parametrize("fmt", [O.CSR, O.ELL, O.SCOO, O.DENSE])
def test_row_sums_equal_dense_row_sums(fmt):
    for k, (nr, nc, c, q) in enumerate(CASES):
        for (pr, pc, NR, NC) in GRIDS:
            o = O.fill_cqmat(fmt, nr, nc, c, q, 1, pr, pc, NR, NC)
            with_signed_values(o, 300 + k)
            g = o.geom
            D = dense_block(o)
            s = o.get_row_sums()
            assert s.size == sums_len(fmt, g)
            off = g["f_row"] if fmt in (O.ELL, O.SCOO) else 0
            got = s[off:off + g["ln_row"]]
            scale = np.maximum(np.abs(D).sum(axis=1), 1e-300)
            assert np.all(np.abs(got - D.sum(axis=1)) <= 1e-12 * scale), (fmt, nr, nc, pr, pc)
            # ELL and SCOO write nothing outside the block's global rows
            if fmt in (O.ELL, O.SCOO):
                assert not np.any(s[:off]) and not np.any(s[off + g["ln_row"]:])


@pytest.mark.parametrize("fmt", [O.CSR, O.ELL, O.SCOO, O.DENSE])
def test_cdiag_row_sums_count_entries(fmt):
    # fill_cdiag puts 1.0 at (i, i - c) and (i, i + c) where they fall inside
    for (n_row, n_col, cd) in ((40, 40, 0), (40, 40, 3), (30, 50, 7), (50, 30, 11)):
        o = O.fill_cdiag(fmt, n_row, n_col, cd)
        s = o.get_row_sums()[:n_row]
        i = np.arange(n_row)
        want = ((i - cd >= 0) & (i - cd < n_col)).astype(float)
        if cd > 0:
            want += ((i + cd < n_col)).astype(float)
        assert np.array_equal(s, want), (fmt, n_row, n_col, cd)


@pytest.mark.parametrize("fmt", [O.CSR, O.ELL, O.SCOO, O.DENSE])
def test_normalised_rows_sum_to_one(fmt):
    o = O.fill_cqmat(fmt, 200, 200, 9, 0.2, 1, value_seed=4)  # values in [0, 1)
    s = o.get_row_sums()
    o.normalize_rows(s)
    t = O.Matrix.get_row_sums(o)
    off = 0
    nz = s[off:off + 200] != 0
    assert np.all(np.abs(t[nz] - 1.0) < 1e-14)


def test_coo_row_ops_are_noops():
    o = O.fill_cqmat(O.COO, 50, 50, 5, 0.3, 1, value_seed=2)
    before = o.arrays()[-1].copy()
    s = o.get_row_sums(np.full(50, 7.0))
    assert np.array_equal(s, np.full(50, 7.0))
    o.normalize_rows(np.zeros(50))
    o.apply_exponential(None, -1)
    assert np.array_equal(o.arrays()[-1], before)


def test_csr_zero_sum_rows_become_zero():
    # row 0: 1 + (-1) = 0 -> the reference's safeguard sets the row to zero
    o = O.csr_from_arrays(3, 3, np.array([0, 2, 3, 3]), np.array([0, 2, 1]),
                          np.array([1.0, -1.0, 4.0]))
    s = o.get_row_sums()
    assert np.array_equal(s, [0.0, 4.0, 0.0])
    o.normalize_rows(s)
    assert np.array_equal(o.arrays()[-1], [0.0, 0.0, 1.0])


def test_max_abs_and_exponential_closed_forms():
    o = O.fill_cqmat(O.CSR, 300, 300, 12, 0.3, 1, value_seed=8)
    v = with_signed_values(o, 17)
    rp = o.arrays()[0]
    m = o.get_row_max_abs()
    for i in range(300):
        row = v[rp[i]:rp[i + 1]]
        nzr = row[row != 0]
        assert m[i] == (np.abs(nzr).max() if nzr.size else 0.0)
    o.apply_exponential(m, -1)
    e = o.arrays()[-1]
    rows = np.repeat(np.arange(300), np.diff(rp))
    nz = v != 0
    assert np.array_equal(e[~nz], v[~nz])  # zeros stay zero
    assert np.allclose(e[nz], np.exp(v[nz] - m[rows[nz]]), rtol=1e-15, atol=0)
    assert np.all(e[nz] <= 1.0)
    # the entry of largest |z| in a row whose maximum is positive becomes 1
    top = [i for i in range(300) if rp[i + 1] > rp[i] and v[rp[i]:rp[i + 1]].max() == m[i] > 0]
    assert top and all(e[rp[i]:rp[i + 1]].max() == 1.0 for i in top)
    o.set_values(v)
    o.apply_exponential(None, 3)
    assert np.allclose(e[nz] * 0 + o.arrays()[-1][nz], np.power(3.0, v[nz]), rtol=1e-15, atol=0)
