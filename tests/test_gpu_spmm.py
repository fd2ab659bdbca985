"""SpMM and personalised PageRank on the GPU against the oracle.

* SpMM (dense_multiply): bit-exact -- the GPU adds each row's products in
  storage order without FMA, as the reference loop.
* Personalised PageRank (64 one-hot teleports as one SpMM per iteration):
  every column within a relative 1e-12 of the oracle's single-vector
  iteration with the same teleport, and at convergence the solution of
  (I - beta A) b = (1 - beta) e_s (normalised), from a direct solve.
"""
import numpy as np
import pytest

import oracle as O
import tbsla_amd as T

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("k", [1, 7, 64, 100, 256, 300, 512])
def test_spmm_bit_exact(k):
    for (n, c, q, grid) in [(3000, 12, 0.2, (0, 0, 1, 1)), (8000, 32, 0.1, (1, 2, 4, 4))]:
        o = O.fill_cqmat(O.CSR, n, n, c, q, 0, *grid, value_seed=5)
        g = T.Matrix(T.CSR).fill_cqmat(n, n, c, q, 0, *grid, value_seed=5)
        ln_col = o.geom["ln_col"]
        B = np.stack([O.unit_vector(50 + j, ln_col) for j in range(k)], axis=1)
        assert np.array_equal(g.spmm(B), o.spmm(B))


def test_spmm_unsupported_formats():
    g = T.Matrix(T.ELL).fill_cqmat(100, 100, 4, 0.1, 0)
    with pytest.raises(T.TbslaHipError):
        g.spmm(np.ones((100, 4)))


def stochastic_pair(n, c, q, seed):
    o = O.fill_cqmat(O.CSR, n, n, c, q, seed, value_seed=seed)
    g = T.Matrix(T.CSR).fill_cqmat(n, n, c, q, seed, value_seed=seed)
    o.make_stochastic()
    g.make_stochastic()
    return o, g


@pytest.mark.parametrize("n", [20000, 70000])
def test_ppr_columns_match_single_vector_oracle(n):
    """n = 70000: 1094 tiles over 768 workgroups (runs of two tiles, the ring
    refilled once per workgroup); n = 20000: 313 tiles, one per workgroup."""
    o, g = stochastic_pair(n, 16, 0.1, 2)
    rng = np.random.default_rng(1)
    src = rng.choice(n, 64, replace=False).astype(np.int32)
    X, it, err = g.personalized_page_rank(src, 0.85, 1e-30, 25)
    assert it == 25
    for j in range(0, 64, 7):
        b, ito, _ = o.page_rank_source(int(src[j]), 0.85, 1e-30, 25)
        assert ito == 25
        rel = np.max(np.abs(X[:, j] - b)) / np.max(b)
        assert rel < 1e-12, (j, rel)


def test_ppr_fixed_point_direct_solve():
    n = 400
    gen = O.fill_cqmat(O.CSR, n, n, 7, 0.2, 3, value_seed=11)
    rp, ci, va = gen.arrays()
    A = np.zeros((n, n))
    for i in range(n):
        np.add.at(A[i], ci[rp[i]:rp[i + 1]], va[rp[i]:rp[i + 1]])
    A += 0.5 * np.eye(n)
    A /= A.sum(axis=0)
    rows, cols = np.nonzero(A)
    rpa = np.zeros(n + 1, np.int64)
    np.add.at(rpa, rows + 1, 1)
    rpa = np.cumsum(rpa)
    geom = dict(n_row=n, n_col=n, f_row=0, f_col=0, ln_row=n, ln_col=n, pr=0, pc=0, NR=1, NC=1,
                nnz=len(rows))
    g = T.Matrix.from_arrays(T.CSR, geom, rpa, cols, A[rows, cols])
    src = np.arange(64, dtype=np.int32) * 5
    X, it, err = g.personalized_page_rank(src, 0.85, 1e-14, 5000)
    assert err < 1e-14
    for j in range(64):
        e = np.zeros(n)
        e[src[j]] = 0.15
        exact = np.linalg.solve(np.eye(n) - 0.85 * A, e)
        exact /= exact.sum()
        assert np.abs(X[:, j] - exact).sum() < 1e-12


def ppr_against_oracle(n, rp, ci, va, src, iters=12):
    geom = dict(n_row=n, n_col=n, f_row=0, f_col=0, ln_row=n, ln_col=n, pr=0, pc=0, NR=1, NC=1,
                nnz=len(ci))
    g = T.Matrix.from_arrays(T.CSR, geom, rp, ci, va)
    o = O.csr_from_arrays(n, n, rp, ci, va)
    X, it, _ = g.personalized_page_rank(src, 0.85, 1e-30, iters)
    assert it == iters
    worst = 0.0
    for j in range(0, 64, 5):
        b, ito, _ = o.page_rank_source(int(src[j]), 0.85, 1e-30, iters)
        assert ito == iters
        worst = max(worst, float(np.max(np.abs(X[:, j] - b)) / np.max(np.abs(b))))
    assert worst < 1e-12, worst


def test_ppr_long_rows_and_many_remote_entries():
    """Rows of 100 entries (past the first 64-entry chunk) with 30 % rewired
    columns (more remote rows than the pipeline's slots), n = 5000."""
    n = 5000
    gen = O.fill_cqmat(O.CSR, n, n, 100, 0.3, 4, value_seed=4)
    gen.make_stochastic()
    rp, ci, va = gen.arrays()
    src = np.random.default_rng(3).choice(n, 64, replace=False).astype(np.int32)
    ppr_against_oracle(n, rp, ci, va, src)


def test_ppr_band_off_the_diagonal():
    """Every row's entries sit 40..49 columns right of the diagonal (wrapping
    at the end), so a tile's own rows lie outside its LDS window: the old
    values come from HBM.  Also a row with no entries."""
    n = 3000
    rows, cols = [], []
    for i in range(n):
        if i == 1234:
            continue
        for k in range(10):
            rows.append(i)
            cols.append((i + 40 + k) % n)
    rows, cols = np.array(rows), np.array(cols)
    rp = np.zeros(n + 1, np.int64)
    np.add.at(rp, rows + 1, 1)
    rp = np.cumsum(rp)
    va = O.unit_vector(6, len(cols)) + 0.1
    src = np.random.default_rng(4).choice(n, 64, replace=False).astype(np.int32)
    ppr_against_oracle(n, rp, cols, va, src)


def test_ppr_off_diagonal_band_long_rows_many_remote():
    """Old values from HBM with rows past the first 32 entries and more
    remote entries than the kernel fetches ahead: each row holds 24 entries
    50..73 columns right of the diagonal (wrapping) and 16 random columns,
    a few rows 90 entries."""
    n = 4000
    rng = np.random.default_rng(9)
    rows, cols = [], []
    for i in range(n):
        k_band, k_rand = (24, 16) if i % 97 else (60, 30)
        for k in range(k_band):
            rows.append(i)
            cols.append((i + 50 + k) % n)
        for c in rng.integers(0, n, k_rand):
            rows.append(i)
            cols.append(int(c))
    rows, cols = np.array(rows), np.array(cols)
    rp = np.zeros(n + 1, np.int64)
    np.add.at(rp, rows + 1, 1)
    rp = np.cumsum(rp)
    va = O.unit_vector(8, len(cols)) + 0.1
    src = np.random.default_rng(5).choice(n, 64, replace=False).astype(np.int32)
    ppr_against_oracle(n, rp, cols, va, src)


@pytest.mark.parametrize("n", [64, 100, 130])
def test_ppr_small_blocks(n):
    """One or three tiles: a partial last tile and a ring larger than the
    block."""
    gen = O.fill_cqmat(O.CSR, n, n, 6, 0.2, 2, value_seed=2)
    gen.make_stochastic()
    rp, ci, va = gen.arrays()
    src = np.random.default_rng(n).choice(n, 64, replace=False).astype(np.int32)
    ppr_against_oracle(n, rp, ci, va, src)
