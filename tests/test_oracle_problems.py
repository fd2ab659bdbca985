"""Pins for the oracle's SPD generators and preconditioned CG (CPU only).

* Poisson with k = 1 equals the Kronecker-sum 7-point Laplacian; with a
  variable k it is symmetric, its row sums equal the boundary faces' k, and
  a two-cell grid gives the harmonic mean by hand.
* The noise field: corr_len 0 is exp(sigma * Box-Muller(V)) written in
  numpy; the smoothing equals scipy.ndimage.gaussian_filter (reflect,
  truncate 3) of the same noise.
* Symmetrised cqmat against the dense formula (D + D^T)/2, diagonal
  row-abs-sum + 1.
* Transpose against scipy; uniform vectors against the numpy generator.
* PCG: kind 0 is bit-identical to the reference-semantics CG restatement;
  Jacobi and G G^T against a numpy PCG; solutions against a direct solve.
"""
import numpy as np
import pytest
import scipy.ndimage as nd
import scipy.sparse as sp
import scipy.sparse.linalg as spla

import oracle as O


def dense(m, n=None):
    rp, ci, va = m.arrays()
    n = n or len(rp) - 1
    return sp.csr_matrix((va, ci, rp), shape=(n, m.geom["ln_col"])).toarray()


def lap1(k):
    return 2 * np.eye(k) - np.eye(k, k=1) - np.eye(k, k=-1)


@pytest.mark.parametrize("dims", [(1, 1, 1), (5, 4, 3), (2, 7, 1)])
def test_poisson_constant_is_kronecker_laplacian(dims):
    nx, ny, nz = dims
    K = (np.kron(np.eye(nz), np.kron(np.eye(ny), lap1(nx)))
         + np.kron(np.eye(nz), np.kron(lap1(ny), np.eye(nx)))
         + np.kron(lap1(nz), np.eye(nx * ny)))
    A = dense(O.poisson3d(*dims))
    assert np.array_equal(A, K)


def test_poisson_variable_coefficient():
    # two cells along x: faces k0 (boundary), 2 k0 k1 / (k0 + k1), k1 (boundary);
    # y and z contribute 4 boundary faces each
    k = np.array([2.0, 0.5])
    A = dense(O.poisson3d(2, 1, 1, kappa=k))
    h = 2 * 2.0 * 0.5 / 2.5
    assert np.allclose(A, [[6 * 2.0 - 2.0 + h, -h], [-h, 6 * 0.5 - 0.5 + h]], rtol=0, atol=1e-15)
    dims = (6, 5, 4)
    kk = np.exp(np.random.default_rng(0).standard_normal(np.prod(dims)))
    A = dense(O.poisson3d(*dims, kappa=kk))
    assert np.array_equal(A, A.T)
    # row sum = k_cell times the number of boundary faces of the cell
    x, y, z = np.meshgrid(np.arange(6), np.arange(5), np.arange(4), indexing="ij")
    x, y, z = (a.transpose(2, 1, 0).ravel() for a in (x, y, z))
    nb = sum(c.astype(int) for c in (x == 0, x == 5, y == 0, y == 4, z == 0, z == 3))
    assert np.allclose(A.sum(axis=1), kk * nb, rtol=1e-13, atol=1e-13)
    assert np.all(np.linalg.eigvalsh(A) > 0)


def test_noise_is_box_muller_of_the_value_generator():
    dims = (8, 6, 5)
    n = np.prod(dims)
    u = O.unit_vector(11, 2 * n)
    z = np.sqrt(-2.0 * np.log(1.0 - u[0::2])) * np.cos(6.283185307179586 * u[1::2])
    f = O.lognormal_field(*dims, 0.0, 0.7, 0.0, 11)
    assert np.max(np.abs(f - np.exp(0.7 * z)) / f) < 1e-14


def test_smoothing_is_scipy_gaussian_filter():
    dims = (20, 18, 16)
    n = np.prod(dims)
    raw = np.log(O.lognormal_field(*dims, 0.0, 1.0, 0.0, 5))
    ell = 2.5
    sm = nd.gaussian_filter(raw.reshape(dims[::-1]), ell, mode="reflect", truncate=3.0).ravel()
    R = int(np.ceil(3 * ell))
    w = np.exp(-0.5 * (np.arange(-R, R + 1) / ell) ** 2)
    w /= w.sum()
    scale = 1 / np.sqrt((w * w).sum() ** 3)
    f = O.lognormal_field(*dims, 0.3, 1.0, ell, 5)
    assert np.max(np.abs(np.log(f) - (0.3 + sm * scale))) < 1e-12
    # unit variance of the smoothed field (whole-grid statistics, loose)
    big = np.log(O.lognormal_field(48, 48, 48, 0.0, 1.0, 2.0, 1))
    assert 0.8 < big.std() < 1.2 and abs(big.mean()) < 0.2


@pytest.mark.parametrize("n,c,q", [(1, 1, 0.0), (300, 8, 0.1), (500, 40, 0.5), (64, 64, 1.0)])
def test_sym_cqmat_dense_formula(n, c, q):
    B = O.fill_cqmat(O.CSR, n, n, c, q, 0, value_seed=7)
    D = np.zeros((n, n))
    rp, ci, va = B.arrays()
    for i in range(n):
        np.add.at(D[i], ci[rp[i]:rp[i + 1]], va[rp[i]:rp[i + 1]])
    S = (D + D.T) / 2
    np.fill_diagonal(S, 0.0)
    np.fill_diagonal(S, np.abs(S).sum(axis=1) + 1.0)
    A = O.sym_cqmat(n, c, q, 0, value_seed=7)
    Ad = dense(A)
    off = ~np.eye(n, dtype=bool)
    assert np.array_equal(Ad[off], S[off])
    assert np.allclose(np.diag(Ad), np.diag(S), rtol=1e-14, atol=0)
    rp, ci, _ = A.arrays()
    assert all(np.all(np.diff(ci[rp[i]:rp[i + 1]]) > 0) for i in range(n))


def test_transpose_matches_scipy():
    m = O.fill_cqmat(O.CSR, 700, 700, 9, 0.3, 0, value_seed=1)
    t = O.transpose(m)
    assert np.array_equal(dense(t), dense(m).T)
    # a block with offsets keeps global indices
    b = O.fill_cqmat(O.CSR, 900, 900, 7, 0.2, 0, 1, 2, 3, 4, value_seed=2)
    tb = O.transpose(b)
    g, gt = b.geom, tb.geom
    assert (gt["f_row"], gt["f_col"], gt["ln_row"], gt["ln_col"]) == \
        (g["f_col"], g["f_row"], g["ln_col"], g["ln_row"])


def test_uniform_vector():
    assert np.array_equal(O.uniform_vector(1000, 3), O.unit_vector(3, 1000))
    v = O.uniform_vector(1000, 3, -1.0, 1.0)
    assert np.array_equal(v, -1.0 + 2.0 * O.unit_vector(3, 1000))


def numpy_pcg(A, b, Minv, k):
    x = np.zeros_like(b)
    r = b.copy()
    z = Minv(r)
    p = z.copy()
    rz = r @ z
    for _ in range(k):
        Ap = A @ p
        a = rz / (p @ Ap)
        x = x + a * p
        r = r - a * Ap
        z = Minv(r)
        rzn = r @ z
        p = z + (rzn / rz) * p
        rz = rzn
    return x


def test_pcg_pins():
    o = O.sym_cqmat(400, 6, 0.2, 0, value_seed=2)
    A = dense(o)
    b = 2 * O.unit_vector(1, 400) - 1
    # kind 0 = the reference-semantics CG restatement, bit for bit
    x0, it0, r0 = O.pcg(o, b, max_iterations=15, tolerance=1e-30, relative=True)
    x1, it1, r1 = o.cg(b, max_iterations=15, tolerance=1e-30, relative=True)
    assert it0 == it1 == 15 and np.array_equal(x0, x1) and r0 == r1
    # Jacobi and G G^T against numpy
    d = np.diag(A)
    xj, _, _ = O.pcg(o, b, max_iterations=6, tolerance=1e-30, kind=1)
    assert np.max(np.abs(xj - numpy_pcg(A, b, lambda r: r / d, 6))) < 1e-12
    L = np.tril(A) / np.sqrt(d)[None, :] * 0.3 + np.diag(0.7 / np.sqrt(d))
    G = sp.csr_matrix(L)
    Go = O.csr_from_arrays(400, 400, G.indptr, G.indices, G.data)
    xf, _, _ = O.pcg(o, b, max_iterations=6, tolerance=1e-30, kind=2, G=Go)
    assert np.max(np.abs(xf - numpy_pcg(A, b, lambda r: L @ (L.T @ r), 6))) < 1e-12
    # converged solutions
    exact = np.linalg.solve(A, b)
    for kind, Gm in ((0, None), (1, None), (2, Go)):
        x, it, rn = O.pcg(o, b, tolerance=1e-12, kind=kind, G=Gm)
        assert np.max(np.abs(x - exact)) < 1e-10, kind


def test_poisson_pcg_direct():
    dims = (10, 9, 8)
    k = O.lognormal_field(*dims, 0.0, 1.0, 2.0, 4)
    o = O.poisson3d(*dims, kappa=k)
    A = sp.csr_matrix(dense(o))
    b = O.uniform_vector(np.prod(dims), 9, -1.0, 1.0)
    exact = spla.spsolve(A.tocsc(), b)
    x, it, rn = O.pcg(o, b, tolerance=1e-12, kind=1)
    assert np.max(np.abs(x - exact)) / np.max(np.abs(exact)) < 1e-9
