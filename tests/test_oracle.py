"""Pins the CPU oracle (oracle/tbsla_oracle.c) before it is trusted.

Each check is chosen so that a plausible mistake in the restatement (wrong
seed, dropped tail quirk, off-by-one block bound, transposed index) fails:
  * glibc's own rand_r (the third-party arithmetic under cqmat_value);
  * the compiled reference library (oracle/_ref, when built here) on the
    reference's own test shapes (src/cpp/tests/test_spmv_cqmat.cpp:162-198)
    over 2-D block grids, all five formats, bit-exact;
  * committed golden fixtures written from the compiled reference
    (tests/golden/), which also pin the oracle where oracle/_ref is absent;
  * the closed forms of the reference tests (src/cpp/utils/array.cpp:17-83);
  * mathematics for the drivers the reference cannot run here (PageRank
    against a direct linear solve, CG against a direct solve).
"""
import ctypes
import glob
import os

import numpy as np
import pytest

import oracle as O

SHAPES = ([(10, 10, c) for c in range(13)] + [(5, 10, c) for c in range(13)] +
          [(10, 5, c) for c in range(13)] + [(30, 30, 2 * c) for c in range(13)] +
          [(20, 30, 2 * c) for c in range(13)] + [(30, 20, 2 * c) for c in range(13)])
GRIDS = [(0, 0, 1, 1), (0, 1, 2, 2), (1, 0, 2, 2), (1, 1, 2, 2), (2, 1, 3, 2), (0, 2, 1, 3),
         (1, 0, 3, 1)]
needs_ref = pytest.mark.skipif(not O.ref_available(), reason="oracle/_ref not built here")


def test_rand_r_matches_glibc():
    libc = ctypes.CDLL("libc.so.6")
    libc.rand_r.argtypes = [ctypes.POINTER(ctypes.c_uint)]
    rng = np.random.default_rng(7)
    seeds = [0, 1, 2, 2**31 - 1, 2**31, 2**32 - 1] + list(rng.integers(0, 2**32, 2000))
    for s in seeds:
        a = ctypes.c_uint(int(s))
        v1 = libc.rand_r(ctypes.byref(a))
        v2, s2 = O.rand_r(int(s))
        assert (v1, a.value) == (v2, s2), s


def test_cqmat_value_structure():
    """Invariants of values_generation.cpp:28-47 that need no reference."""
    n, c = 40, 6
    # q = 0: pure upper band, row i holds columns i .. i+c-1 (full rows)
    for k in range((n - c + 1) * c):
        assert O.cqmat_value(k, n, n, c, 0.0, 1) == (k // c, k // c + k % c)
    # tail rows: row n-c+t has c-t entries (t = 1 .. c-1), band continues
    k = (n - c + 1) * c
    for t in range(1, c):
        for j in range(c - t):
            assert O.cqmat_value(k, n, n, c, 0.0, 1) == (n - c + t, n - c + t + j)
            k += 1
    assert k == O.cqmat_gnv(n, n, c)
    # q > 0: every tail entry is rewired (integer division quirk, :45) -- its
    # column is the second rand_r draw of its own seed
    for kk in range((n - c + 1) * c, O.cqmat_gnv(n, n, c)):
        d1, s1 = O.rand_r(kk)
        d2, _ = O.rand_r(s1)
        assert d1 != 2147483647
        assert O.cqmat_value(kk, n, n, c, 1e-9, 1)[1] == d2 % n
    # seed_mult 0 behaves as 1 (:29-32)
    for kk in range(0, 300, 7):
        assert O.cqmat_value(kk, n, n, c, 0.3, 0) == O.cqmat_value(kk, n, n, c, 0.3, 1)


def test_lnv_pflv_partition():
    for size in (0, 1, 7, 100, 4194304, 4194307):
        for npos in (1, 2, 3, 4, 7, 256):
            lens = [O.lnv(size, p, npos) for p in range(npos)]
            firsts = [O.pflv(size, p, npos) for p in range(npos)]
            assert sum(lens) == size
            assert firsts[0] == 0
            assert all(firsts[p] + lens[p] == firsts[p + 1] for p in range(npos - 1))
            assert max(lens) - min(lens) <= 1


@needs_ref
def test_cqmat_value_matches_reference():
    rng = np.random.default_rng(3)
    for (nr, nc, c) in [(30, 30, 8), (100, 80, 13), (17, 40, 40), (4194304, 4194304, 32)]:
        gnv = O.cqmat_gnv(nr, nc, c)
        ks = list(range(min(gnv, 400))) + list(rng.integers(0, gnv + 50, 400))
        for q in (0.0, 0.1, 1.0):
            for s in (0, 1, 5):
                for k in ks:
                    assert O.cqmat_value(int(k), nr, nc, c, q, s) == \
                        O.ref_cqmat_value(int(k), nr, nc, c, q, s)


@needs_ref
def test_cdiag_value_matches_reference():
    for nc in (5, 10, 30):
        for cd in range(0, 14):
            nv = 2 * nc
            for i in range(nv):
                assert O.cdiag_value(i, nv, nc, nc, cd) == O.ref_cdiag_value(i, nv, nc, nc, cd)


def _compare(o, r, fmt, skip_dense_empty_nnz=True):
    go, gr = o.geom, r.geom
    if fmt == O.DENSE and skip_dense_empty_nnz and not o.s.values:
        # DENSE leaves nnz uninitialised when the block holds no generator
        # entries (MatrixDENSE.cpp:227-228); nothing to compare there.
        go.pop("nnz")
        gr.pop("nnz")
    assert go == gr
    ao = o.arrays()
    ar = r.arrays(stored=len(ao[0]) if fmt in (O.COO, O.SCOO) else None)
    assert len(ao) == len(ar)
    for x, y in zip(ao, ar):
        assert x.shape == y.shape and np.array_equal(x, y)


@needs_ref
@pytest.mark.parametrize("fmt", [O.COO, O.SCOO, O.CSR, O.ELL, O.DENSE])
def test_fill_cqmat_matches_reference(fmt):
    for (nr, nc, c) in SHAPES:
        for q in (0.0, 0.1, 0.5, 1.0):
            for s in (0, 1, 3):
                for (pr, pc, NR, NC) in GRIDS:
                    o = O.fill_cqmat(fmt, nr, nc, c, q, s, pr, pc, NR, NC)
                    r = O.ref_fill_cqmat(fmt, nr, nc, c, q, s, pr, pc, NR, NC)
                    _compare(o, r, fmt)


@needs_ref
@pytest.mark.parametrize("fmt", [O.COO, O.SCOO, O.CSR, O.ELL, O.DENSE])
def test_fill_cdiag_matches_reference(fmt):
    for (nr, nc, c) in SHAPES:
        for (pr, pc, NR, NC) in GRIDS:
            o = O.fill_cdiag(fmt, nr, nc, c, pr, pc, NR, NC)
            r = O.ref_fill_cdiag(fmt, nr, nc, c, pr, pc, NR, NC)
            _compare(o, r, fmt)


@needs_ref
def test_fill_cqmat_large_block_matches_reference():
    """One 4x4-grid block of a larger matrix, all formats but DENSE."""
    for fmt in (O.COO, O.SCOO, O.CSR, O.ELL):
        for (pr, pc) in ((0, 0), (0, 1), (3, 3)):
            o = O.fill_cqmat(fmt, 20000, 20000, 16, 0.1, 0, pr, pc, 4, 4)
            r = O.ref_fill_cqmat(fmt, 20000, 20000, 16, 0.1, 0, pr, pc, 4, 4)
            _compare(o, r, fmt)


def compute_spmv_res_cdiag(nr, nc, c, v):
    """Closed form of src/cpp/utils/array.cpp:17-30."""
    r = np.zeros(nr)
    i = 0
    while i < min(c, nr):
        r[i] = v[i + c] if i < nc - c else 0
        i += 1
    while i < min(nr, nc - c):
        r[i] = v[i] if c == 0 else v[i + c] + v[i - c]
        i += 1
    while i < nr:
        r[i] = v[i - c] if i < nc + c else 0
        i += 1
    return r


@pytest.mark.parametrize("fmt", [O.COO, O.SCOO, O.CSR, O.ELL, O.DENSE])
def test_spmv_cdiag_closed_form(fmt):
    """src/cpp/tests/test_spmv_cdiag.cpp: every format, generated and
    converted from COO, spmv(iota) equal to the closed form."""
    for (nr, nc, c) in SHAPES:
        v = np.arange(nc, dtype=np.float64)
        m = O.fill_cdiag(fmt, nr, nc, c)
        assert np.array_equal(m.spmv(v), compute_spmv_res_cdiag(nr, nc, c, v))
        if fmt != O.COO:
            m2 = O.fill_cdiag(O.COO, nr, nc, c).to(fmt)
            assert np.array_equal(m2.spmv(v), compute_spmv_res_cdiag(nr, nc, c, v))


@pytest.mark.parametrize("fmt", [O.COO, O.SCOO, O.CSR, O.ELL, O.DENSE])
def test_a_axpx_cdiag_closed_form(fmt):
    """src/cpp/tests/test_a_axpx__cdiag.cpp with array.cpp:54-83."""
    for (n, c) in [(10, i) for i in range(13)] + [(30, 2 * i) for i in range(13)]:
        v = 2 * np.arange(n, dtype=np.float64) + 1
        exp = compute_spmv_res_cdiag(n, n, c, v) + v
        exp = compute_spmv_res_cdiag(n, n, c, exp)
        m = O.fill_cdiag(fmt, n, n, c)
        assert np.array_equal(m.a_axpx_(v), exp)


@needs_ref
@pytest.mark.parametrize("fmt", [O.COO, O.SCOO, O.CSR, O.ELL, O.DENSE])
def test_spmv_matches_reference(fmt):
    """spmv and a_axpx_ on cqmat blocks (values 1, integer inputs: exact)."""
    for (nr, nc, c) in SHAPES[::3]:
        for q in (0.0, 0.3, 1.0):
            for (pr, pc, NR, NC) in GRIDS[:4]:
                o = O.fill_cqmat(fmt, nr, nc, c, q, 1, pr, pc, NR, NC)
                r = O.ref_fill_cqmat(fmt, nr, nc, c, q, 1, pr, pc, NR, NC)
                g = o.geom
                if fmt == O.DENSE and not o.s.values:
                    continue
                v = np.arange(g["n_col"], dtype=np.float64)[g["f_col"]:]
                assert np.array_equal(o.spmv(v), r.spmv(v, o.result_len))
                if nr == nc and NR == NC == 1:
                    assert np.array_equal(o.a_axpx_(v), r.a_axpx_(v, o.result_len))


@needs_ref
@pytest.mark.parametrize("fmt", [O.SCOO, O.CSR, O.ELL, O.DENSE])
def test_conversion_from_coo_matches_reference(fmt):
    for (nr, nc, c) in SHAPES[::2]:
        for q in (0.0, 0.5, 1.0):
            for s in (0, 2):
                for (pr, pc, NR, NC) in ((0, 0, 1, 1), (1, 0, 2, 2), (2, 1, 3, 2)):
                    co = O.fill_cqmat(O.COO, nr, nc, c, q, s, pr, pc, NR, NC)
                    cr = O.ref_fill_cqmat(O.COO, nr, nc, c, q, s, pr, pc, NR, NC)
                    _compare(co.to(fmt), cr.to(fmt), fmt, skip_dense_empty_nnz=False)


@needs_ref
@pytest.mark.parametrize("fmt", [O.SCOO, O.CSR, O.ELL])
def test_make_stochastic_matches_reference(fmt):
    for (nr, nc, c) in [(30, 30, 6), (30, 20, 10), (20, 30, 4), (200, 200, 9)]:
        for q in (0.0, 0.2, 1.0):
            for (pr, pc, NR, NC) in ((0, 0, 1, 1), (1, 1, 2, 2), (0, 1, 1, 2)):
                o = O.fill_cqmat(fmt, nr, nc, c, q, 1, pr, pc, NR, NC)
                r = O.ref_fill_cqmat(fmt, nr, nc, c, q, 1, pr, pc, NR, NC)
                so = o.make_stochastic()
                sr = r.make_stochastic(o.geom["ln_col"])
                assert np.array_equal(so, sr)
                _compare(o, r, fmt)


GOLDEN = sorted(glob.glob(os.path.join(os.path.dirname(__file__), "golden", "cqmat_*.npz")))


@pytest.mark.skipif(not GOLDEN, reason="no golden fixtures")
def test_oracle_matches_golden_fixtures():
    """Fixtures written from the compiled reference by tests/golden/make_golden.py."""
    for path in GOLDEN:
        z = np.load(path, allow_pickle=False)
        fmt, nr, nc, c, q, s, pr, pc, NR, NC = (int(z["fmt"]), int(z["n_row"]), int(z["n_col"]),
                                                int(z["c"]), float(z["q"]), int(z["seed_mult"]),
                                                int(z["pr"]), int(z["pc"]), int(z["NR"]),
                                                int(z["NC"]))
        m = O.fill_cqmat(fmt, nr, nc, c, q, s, pr, pc, NR, NC)
        g = m.geom
        assert [g[f] for f in O.GEOM_FIELDS] == list(z["geom"]), path
        assert g["nnz"] == int(z["nnz"]), path
        for name, arr in zip(z["names"], m.arrays()):
            assert np.array_equal(arr, z[str(name)]), (path, name)
        v = z["x"]
        assert np.array_equal(m.spmv(v), z["y"]), path


def test_unit_value_vectorised_equals_scalar():
    for seed in (0, 1, 12345):
        vec = O.unit_vector(seed, 3000)
        assert all(vec[k] == O.unit_value(seed, k) for k in range(0, 3000, 37))
        assert vec.min() >= 0.0 and vec.max() < 1.0
        assert abs(vec.mean() - 0.5) < 0.03


def _dense_of(m):
    rp, ci, va = m.arrays()
    n = m.geom["ln_row"]
    A = np.zeros((n, m.geom["ln_col"]))
    for i in range(n):
        np.add.at(A[i], ci[rp[i]:rp[i + 1]], va[rp[i]:rp[i + 1]])
    return A


def test_page_rank_fixed_point():
    """page_rank on a column-stochastic matrix with no empty column converges
    to the solution of (I - beta A) b = (1 - beta)/n 1, normalised to sum 1:
    a direct solve, not an iteration."""
    n = 300
    gen = O.fill_cqmat(O.CSR, n, n, 7, 0.2, 3, value_seed=11)
    m = O.csr_from_dense(_dense_of(gen) + 0.5 * np.eye(n))  # no empty column
    s = m.make_stochastic()
    A = _dense_of(m)
    assert np.allclose(A.sum(axis=0), 1.0, rtol=0, atol=1e-14)
    assert np.all(s > 0)
    b, it, err = m.page_rank(0.85, 1e-13, 10000)
    exact = np.linalg.solve(np.eye(n) - 0.85 * A, np.full(n, 0.15 / n))
    exact /= exact.sum()
    assert err < 1e-13 and it < 10000
    assert abs(b.sum() - 1.0) < 1e-14
    # stopping at an L1 change < eps leaves an L1 error <= eps*beta/(1-beta)
    assert np.abs(b - exact).sum() < 1e-13 * 0.85 / 0.15 * 1.5


def test_page_rank_cycle_uniform():
    """A directed cycle is a permutation matrix: the fixed point is 1/n."""
    n = 50
    P = np.zeros((n, n))
    P[np.arange(n), (np.arange(n) + 1) % n] = 1.0
    b, it, err = O.csr_from_dense(P).page_rank(0.85, 1e-12, 1000)
    # b starts at 1 (Matrix.cpp:549-552): one step reaches 1/n, the second
    # measures a zero change
    assert it == 2 and err < 1e-15
    assert np.allclose(b, 1.0 / n, rtol=0, atol=1e-15)


def test_page_rank_max_norm_known_answer():
    """MatrixCOO's max-normalised PageRank on A = [[0, 1], [0, 0]], beta 0.5,
    worked on paper: b_t = (1, 1) -> y = (1, 0), T = (0 + 2)/4 -> (1, .5);
    y = (.5, 0), T = (.5 + 1.5)/4 (the carried T) -> (.75, .5) -> (1, 2/3);
    y = (2/3, 0), T = (.5 + 5/3)/4 -> (1, 13/21).  Without the carry the
    third iterate would be (1, 4/7)."""
    want = {1: ((2 / 3, 1 / 3), 0.5), 2: ((0.6, 0.4), 1 / 6), 3: ((21 / 34, 13 / 34), 1 / 21)}
    for m in (O.coo_from_arrays(2, 2, [0], [1], [1.0]),
              O.csr_from_dense(np.array([[0.0, 1.0], [0.0, 0.0]]))):
        for k, (b_want, e_want) in want.items():
            b, it, err = m.page_rank_max_norm(0.5, 1e-300, k)
            assert it == k
            assert np.allclose(b, b_want, rtol=0, atol=1e-15)
            assert abs(err - e_want) < 1e-15


def test_page_rank_max_norm_fixed_point():
    """On a column-stochastic A the carried teleport term settles at
    T = c S / (1 - c), c = (1 - beta)/n, S = sum(b); summing M b = beta A b + T
    gives the fixed point's max M = beta + (1 - beta) n / (n - 1 + beta), so
    the result is (M I - beta A)^-1 1 normalised -- a direct solve.  (With
    T = c S, M would be 1: the plain PageRank vector, 1e-5 away here.)"""
    n = 300
    gen = O.fill_cqmat(O.CSR, n, n, 7, 0.2, 3, value_seed=11)
    m = O.csr_from_dense(_dense_of(gen) + 0.5 * np.eye(n))
    m.make_stochastic()
    A = _dense_of(m)
    beta = 0.85
    M = beta + (1 - beta) * n / (n - 1 + beta)
    exact = np.linalg.solve(M * np.eye(n) - beta * A, np.ones(n))
    exact /= exact.sum()
    plain = np.linalg.solve(np.eye(n) - beta * A, np.ones(n))
    plain /= plain.sum()
    assert np.abs(exact - plain).max() > 1e-6 * exact.max()
    for mm in (m, O.coo_from_arrays(n, n, *np.nonzero(A), A[np.nonzero(A)])):
        b, it, err = mm.page_rank_max_norm(beta, 1e-14, 10000)
        assert err < 1e-14 and it < 10000
        assert abs(b.sum() - 1.0) < 1e-14
        assert np.abs(b - exact).max() < 1e-12 * exact.max()


def test_cg_solves_laplacian():
    """CG on the 1-D Laplacian tridiag(-1, 2, -1) against a direct solve,
    within n iterations; absolute and relative stopping rules."""
    n = 64
    A = 2 * np.eye(n) - np.eye(n, k=1) - np.eye(n, k=-1)
    m = O.csr_from_dense(A)
    rhs = np.sin(np.arange(n) + 1.0)
    x, it, rn = m.cg(rhs, max_iterations=1000, tolerance=1e-10)
    exact = np.linalg.solve(A, rhs)
    assert it <= n and rn < 1e-10
    assert np.max(np.abs(x - exact)) < 1e-8 * np.max(np.abs(exact))
    x2, it2, rn2 = m.cg(rhs, max_iterations=1000, tolerance=1e-8, relative=True)
    assert rn2 < 1e-8 * np.linalg.norm(rhs)
    assert np.linalg.norm(rhs - A @ x2) < 2e-8 * np.linalg.norm(rhs)
    # a diagonal system converges in one step per distinct eigenvalue
    D = np.diag(np.repeat([1.0, 3.0, 7.0], 10))
    xd, itd, _ = O.csr_from_dense(D).cg(np.ones(30), tolerance=1e-12)
    assert itd == 2 and np.allclose(xd, 1 / np.diag(D), rtol=1e-14)


def test_spmm_columns_equal_spmv():
    m = O.fill_cqmat(O.CSR, 500, 500, 9, 0.2, 1, value_seed=4)
    B = np.stack([O.unit_vector(100 + k, 500) for k in range(8)], axis=1)
    C = m.spmm(B)
    for k in range(8):
        assert np.array_equal(C[:, k], m.spmv(B[:, k]))
