"""Column normalisation, PageRank and CG on the GPU against the oracle.

* get_col_sums / make_stochastic: bit-exact with the oracle's serial loop
  (the GPU sums each column in the reference's loop order).
* PageRank (page_rank semantics): same iteration count, iterate within a
  relative 1e-12 (max-norm) of the oracle.
* CG (corrected): early iterates within a relative 1e-12 of the oracle's
  (the iteration amplifies the last-bit differences of the dot products
  only slowly there), the converged solution against a direct solve.
"""
import numpy as np
import pytest

import oracle as O
import tbsla_amd as T

pytestmark = pytest.mark.gpu


def rel_max(a, b):
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


def to_gpu_csr(m):
    """Import an oracle CSR block into a GPU CSR handle."""
    rp, ci, va = m.arrays()
    return T.Matrix.from_arrays(T.CSR, m.geom, rp, ci, va)


def laplacian_1d(n):
    A = 2 * np.eye(n) - np.eye(n, k=1) - np.eye(n, k=-1)
    return A


def sym_cqmat(n, c, q, seed):
    """A = (B + B^T)/2 with the diagonal set to the row's off-diagonal
    absolute sum + 1 (strictly diagonally dominant, SPD)."""
    B = O.fill_cqmat(O.CSR, n, n, c, q, seed, value_seed=seed)
    rp, ci, va = B.arrays()
    D = np.zeros((n, n))
    for i in range(n):
        np.add.at(D[i], ci[rp[i]:rp[i + 1]], va[rp[i]:rp[i + 1]])
    A = (D + D.T) / 2
    np.fill_diagonal(A, 0.0)
    np.fill_diagonal(A, np.abs(A).sum(axis=1) + 1.0)
    return A


@pytest.mark.parametrize("fmt", [T.CSR, T.ELL, T.SCOO])
def test_col_sums_and_normalisation_bit_exact(fmt):
    for (n, c, q, grid) in [(3000, 12, 0.2, (0, 0, 1, 1)), (5000, 20, 0.5, (1, 2, 3, 4)),
                            (20000, 32, 0.1, (0, 0, 1, 1))]:
        o = O.fill_cqmat(fmt, n, n, c, q, 0, *grid, value_seed=9)
        g = T.Matrix(fmt).fill_cqmat(n, n, c, q, 0, *grid, value_seed=9)
        so = o.get_col_sums()
        sg = g.get_col_sums()
        assert np.array_equal(sg, so), (fmt, n)
        so = O.fill_cqmat(fmt, n, n, c, q, 0, *grid, value_seed=9)
        s1 = so.make_stochastic()
        s2 = g.make_stochastic()
        assert np.array_equal(s2, s1)
        for a, b in zip(g.arrays(), so.arrays()):
            assert np.array_equal(a, b)


@pytest.mark.parametrize("fmt", [T.COO, T.DENSE])
def test_col_sums_noop_formats(fmt):
    g = T.Matrix(fmt).fill_cqmat(500, 500, 8, 0.2, 0, value_seed=1)
    before = g.arrays()
    s = g.make_stochastic(np.full(500, 3.0))
    assert np.array_equal(s, np.full(500, 3.0))
    for a, b in zip(g.arrays(), before):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("fmt", [T.CSR, T.ELL, T.SCOO, T.COO])
def test_page_rank_matches_oracle(fmt):
    n, c, q = 60000, 16, 0.1
    o = O.fill_cqmat(O.CSR, n, n, c, q, 0, value_seed=0)
    o.make_stochastic()
    bo, ito, eo = o.page_rank(0.85, 1e-10, 10000)
    g = T.Matrix(fmt).fill_cqmat(n, n, c, q, 0, value_seed=0)
    if fmt == T.COO:
        # COO's make_stochastic is a no-op in the reference: import the
        # oracle's normalised CSR entries as triplets instead
        rp, ci, va = o.arrays()
        rows = np.repeat(np.arange(n, dtype=np.int32), np.diff(rp))
        geom = dict(o.geom, nnz=len(rows))
        g = T.Matrix.from_arrays(T.COO, geom, rows, ci, va)
    else:
        g.make_stochastic()
    bg, itg, eg = g.page_rank(0.85, 1e-10, 10000)
    assert itg == ito, (itg, ito)
    assert rel_max(bg, bo) < 1e-12
    assert abs(bg.sum() - 1.0) < 1e-12


@pytest.mark.parametrize("fmt", [T.COO, T.CSR, T.ELL, T.SCOO, T.DENSE])
def test_page_rank_max_norm_matches_oracle(fmt):
    """MatrixCOO's max-normalised PageRank: same iteration count, iterate
    within a relative 1e-12 of the oracle."""
    n, c, q = (2000, 12, 0.2) if fmt == T.DENSE else (60000, 16, 0.1)
    o = O.fill_cqmat(O.CSR, n, n, c, q, 0, value_seed=0)
    o.make_stochastic()
    rp, ci, va = o.arrays()
    rows = np.repeat(np.arange(n, dtype=np.int32), np.diff(rp))
    oc = O.coo_from_arrays(n, n, rows, ci, va)
    bo, ito, eo = oc.page_rank_max_norm(0.85, 1e-10, 10000)
    geom = dict(o.geom, nnz=len(rows))
    g = T.Matrix.from_arrays(T.COO, geom, rows, ci, va)
    if fmt != T.COO:
        g = g.to(fmt)
    bg, itg, eg = g.page_rank_max_norm(0.85, 1e-10, 10000)
    assert itg == ito, (itg, ito)
    assert eg < 1e-10 and eo < 1e-10
    assert rel_max(bg, bo) < 1e-12
    assert abs(bg.sum() - 1.0) < 1e-12


def test_page_rank_max_norm_known_answer_and_geometry():
    """The 2 x 2 case worked on paper (see the oracle's test); a block that
    does not cover the matrix is refused."""
    geom = dict(n_row=2, n_col=2, ln_row=2, ln_col=2, f_row=0, f_col=0, pr=0, pc=0, NR=1,
                NC=1, nnz=1)
    g = T.Matrix.from_arrays(T.COO, geom, np.array([0], np.int32), np.array([1], np.int32),
                             np.array([1.0]))
    for k, b_want, e_want in [(1, (2 / 3, 1 / 3), 0.5), (3, (21 / 34, 13 / 34), 1 / 21)]:
        b, it, err = g.page_rank_max_norm(0.5, 1e-300, k)
        assert it == k and np.allclose(b, b_want, rtol=0, atol=1e-15)
        assert abs(err - e_want) < 1e-15
    blk = T.Matrix(T.CSR).fill_cqmat(1000, 1000, 8, 0.1, 0, 0, 0, 2, 2, value_seed=1)
    with pytest.raises(T.TbslaHipError):
        blk.page_rank_max_norm()


@pytest.mark.parametrize("fmt", [T.CSR, T.ELL])
def test_page_rank_fused_in_diagonal_layout(fmt):
    """A block in the diagonal layout's LDS-slice form runs each PageRank
    iteration as one pass (the update inside the diagonal kernel, one
    partial per row block): same iteration count as the oracle, iterate
    within a relative 1e-12, within 1e-13 of the row layout's, and the same
    bits on a second solve."""
    n, c, q = 100000, 16, 0.1
    o = O.fill_cqmat(O.CSR, n, n, c, q, 0, value_seed=0)
    o.make_stochastic()
    bo, ito, _ = o.page_rank(0.85, 1e-10, 10000)
    g = T.Matrix(fmt).fill_cqmat(n, n, c, q, 0, value_seed=0)
    g.make_stochastic()
    b_rows, it_rows, _ = g.page_rank(0.85, 1e-10, 10000)
    g.set_csr_layout(T.CSR_DIAG, 65536, 16)
    assert g.get_csr_layout()["lds_block_rows"] > 0
    b1, it1, e1 = g.page_rank(0.85, 1e-10, 10000)
    b2, it2, e2 = g.page_rank(0.85, 1e-10, 10000)
    assert it1 == ito == it_rows, (it1, ito, it_rows)
    assert rel_max(b1, bo) < 1e-12
    assert rel_max(b1, b_rows) < 1e-13
    assert np.array_equal(b1, b2) and e1 == e2 and it2 == it1
    assert abs(b1.sum() - 1.0) < 1e-12
    # a fixed number of iterations (epsilon 0)
    bk, itk, _ = g.page_rank(0.85, 0.0, 7)
    bok, itok, _ = o.page_rank(0.85, 0.0, 7)
    assert itk == itok == 7 and rel_max(bk, bok) < 1e-12


def test_page_rank_max_iterations_and_determinism():
    n = 30000
    g = T.Matrix(T.CSR).fill_cqmat(n, n, 8, 0.3, 2, value_seed=4)
    g.make_stochastic()
    b1, it1, e1 = g.page_rank(0.85, 1e-14, 37)
    b2, it2, e2 = g.page_rank(0.85, 1e-14, 37)
    assert it1 == 37 and np.array_equal(b1, b2) and e1 == e2
    o = O.fill_cqmat(O.CSR, n, n, 8, 0.3, 2, value_seed=4)
    o.make_stochastic()
    bo, ito, _ = o.page_rank(0.85, 1e-14, 37)
    assert ito == 37 and rel_max(b1, bo) < 1e-12


def test_cg_early_iterates_match_oracle():
    n = 4000
    A = sym_cqmat(n, 9, 0.1, 3)
    o = O.csr_from_dense(A)
    g = to_gpu_csr(o)
    rhs = 2 * O.unit_vector(7, n) - 1
    for k in (1, 3, 10, 30):
        xo, ito, ro = o.cg(rhs, max_iterations=k, tolerance=1e-30)
        xg, itg, rg = g.cg(rhs, max_iterations=k, tolerance=1e-30)
        assert itg == ito == k
        assert rel_max(xg, xo) < 1e-12, (k, rel_max(xg, xo))
        assert abs(rg - ro) <= 1e-10 * ro


def test_cg_converges_like_oracle():
    """Without the divisor clamp: same iteration count as the oracle within
    two, the solution of a direct solve."""
    n = 2000
    A = laplacian_1d(n) + np.diag(np.linspace(0.01, 1.0, n))
    o = O.csr_from_dense(A)
    g = to_gpu_csr(o)
    rhs = 2 * O.unit_vector(11, n) - 1
    exact = np.linalg.solve(A, rhs)
    xo, ito, ro = o.cg(rhs, tolerance=1e-10, near_zero=-1)
    xg, itg, rg = g.cg(rhs, tolerance=1e-10, near_zero=-1)
    assert abs(itg - ito) <= 2, (itg, ito)
    assert rg < 1e-10
    assert np.linalg.norm(A @ xg - rhs) < 2e-10
    assert rel_max(xg, exact) < 1e-8


def test_cg_reference_clamp_like_oracle():
    """The reference's absolute 1e-10 clamp (the default for the absolute
    test): the GPU follows the oracle's clamped iteration, which needs
    several times the unclamped count (656 against 140 to 1e-10 here)."""
    n = 2000
    A = laplacian_1d(n) + np.diag(np.linspace(0.01, 1.0, n))
    o = O.csr_from_dense(A)
    g = to_gpu_csr(o)
    rhs = 2 * O.unit_vector(11, n) - 1
    xo, ito, ro = o.cg(rhs, max_iterations=3000, tolerance=1e-6)
    xg, itg, rg = g.cg(rhs, max_iterations=3000, tolerance=1e-6)
    assert abs(itg - ito) <= max(2, 0.02 * ito), (itg, ito)
    assert rg < 1e-6 and ro < 1e-6
    _, it_free, _ = g.cg(rhs, max_iterations=3000, tolerance=1e-6, near_zero=-1)
    assert it_free < itg / 2


def test_cg_relative_jacobi_and_mixed():
    n = 20000
    A = None
    # symmetric strongly diagonally dominant matrix with varying diagonal,
    # built as COO on the host and imported
    rng = np.random.default_rng(5)
    i = np.arange(n)
    rows = np.concatenate([i, i[1:], i[:-1], i[2:], i[:-2]])
    cols = np.concatenate([i, i[:-1], i[1:], i[:-2], i[2:]])
    off = -rng.random(n - 1)
    off2 = -0.5 * rng.random(n - 2)
    vals_off = np.concatenate([off, off, off2, off2])
    diag = np.zeros(n)
    np.add.at(diag, rows[n:], -vals_off)
    diag = diag * (1 + 100 * rng.random(n)) + 1e-3
    vals = np.concatenate([diag, vals_off])
    order = np.lexsort((cols, rows))
    rows, cols, vals = rows[order], cols[order], vals[order]
    rp = np.zeros(n + 1, np.int64)
    np.add.at(rp, rows + 1, 1)
    rp = np.cumsum(rp)
    o = O.csr_from_arrays(n, n, rp, cols, vals)
    g = to_gpu_csr(o)
    rhs = 2 * O.unit_vector(3, n) - 1
    x0, it0, _ = g.cg(rhs, tolerance=1e-8, relative=True)
    xj, itj, _ = g.cg(rhs, tolerance=1e-8, relative=True, preconditioner=T.PRECOND_JACOBI)
    xm, itm, _ = g.cg(rhs, tolerance=1e-8, relative=True, precision=T.CG_MIXED, replace_every=50)
    xmj, itmj, _ = g.cg(rhs, tolerance=1e-8, relative=True, precision=T.CG_MIXED,
                        preconditioner=T.PRECOND_JACOBI, replace_every=50)

    def true_rel(x):
        y = o.spmv(x)
        return np.linalg.norm(rhs - y) / np.linalg.norm(rhs)
    assert true_rel(x0) < 1.5e-8
    assert true_rel(xj) < 1.5e-8
    assert itj < it0  # the varying diagonal is what Jacobi removes
    # mixed precision: convergence is judged on the FP64 true residual
    assert true_rel(xm) < 1.5e-8
    assert true_rel(xmj) < 1.5e-8
    # the oracle's plain CG agrees on the iteration count within a couple
    xo, ito, _ = o.cg(rhs, tolerance=1e-8, relative=True)
    assert abs(ito - it0) <= 2


def stochastic_parts(sizes, seed=0):
    """Column-stochastic cqmat CSR blocks of the given (n, c) and their
    block diagonal."""
    parts = []
    for i, (n, c) in enumerate(sizes):
        g = T.Matrix(T.CSR).fill_cqmat(n, n, c, 0.1, 0, value_seed=seed + i)
        g.make_stochastic()
        parts.append(g)
    offs = np.concatenate([[0], np.cumsum([n for n, _ in sizes])]).astype(np.int32)
    return parts, T.Matrix.block_diag(parts), offs


def test_page_rank_batched_equals_single():
    """Every instance of the batch follows the single-block solve of its own
    block bit for bit (same row kernel), iteration count and error included,
    and the oracle's PageRank within 1e-12 with the same count."""
    sizes = [(4096, 8), (777, 31), (20000, 16), (3, 2), (9000, 20), (50000, 32)]
    parts, bd, offs = stochastic_parts(sizes)
    bd.set_csr_kernel(8, 1, 1, 1)
    b, its, errs, tus = bd.page_rank_batched(offs, 0.85, 1e-10, 500, times=True)
    # solve times: positive, and an instance that needs more iterations ends later
    assert np.all(tus > 0)
    assert tus[np.argmax(its)] >= tus[np.argmin(its)]
    for k, (g, (n, c)) in enumerate(zip(parts, sizes)):
        g.set_csr_kernel(8, 1, 1, 1)
        bs, it, err = g.page_rank(0.85, 1e-10, 500)
        assert its[k] == it and errs[k] == err, (k, its[k], it, errs[k], err)
        assert np.array_equal(b[offs[k]:offs[k + 1]], bs), k
        o = O.fill_cqmat(O.CSR, n, n, c, 0.1, 0, value_seed=k)
        o.make_stochastic()
        bo, ito, _ = o.page_rank(0.85, 1e-10, 500)
        assert ito == it and rel_max(bs, bo) < 1e-12, (k, ito, it)


def test_page_rank_batched_iteration_cap_and_one_instance():
    sizes = [(3000, 12), (6000, 24)]
    parts, bd, offs = stochastic_parts(sizes, seed=7)
    b, its, errs = bd.page_rank_batched(offs, 0.85, 1e-30, 21)  # never converges: capped
    assert list(its) == [21, 21]
    for k, g in enumerate(parts):
        bd_k = b[offs[k]:offs[k + 1]]
        assert abs(bd_k.sum() - 1.0) < 1e-12
    one, its1, _ = parts[1].page_rank_batched(np.array([0, 6000], np.int32), 0.85, 1e-10, 500)
    single, it1, _ = parts[1].page_rank(0.85, 1e-10, 500)
    assert its1[0] == it1 and np.array_equal(one, single)
    with pytest.raises(T.TbslaHipError):
        bd.page_rank_batched(np.array([0, 100, 8000], np.int32))  # offsets must end at ln_row
