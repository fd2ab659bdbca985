"""SPD problem generators on the GPU against the oracle.

* 7-point Poisson (k = 1 and a log-normal k), symmetrised cqmat, CSR
  transpose, block-diagonal concatenation, uniform vectors: bit-exact.
* Log-normal field: within a relative 1e-13 of the oracle (device log, cos
  and exp against libm's).
"""
import numpy as np
import pytest
import scipy.sparse as sp

import oracle as O
import tbsla_amd as T

pytestmark = pytest.mark.gpu


def same_csr(g, o):
    a, b = g.arrays(), o.arrays()
    return all(np.array_equal(x, y) for x, y in zip(a, b))


@pytest.mark.parametrize("dims", [(1, 1, 1), (5, 4, 3), (16, 16, 16), (7, 1, 9)])
def test_poisson_constant_bit_exact(dims):
    g = T.Matrix(T.CSR).fill_poisson3d(*dims)
    o = O.poisson3d(*dims)
    assert same_csr(g, o)
    assert g.geom["nnz"] == o.geom["nnz"]


def test_lognormal_field_and_variable_poisson():
    dims = (24, 20, 18)
    dk = T.lognormal_field(*dims, mu=0.0, sigma=1.0, corr_len=4.0, seed=9)
    kg = dk.to_numpy(count=np.prod(dims))
    ko = O.lognormal_field(*dims, 0.0, 1.0, 4.0, 9)
    assert np.max(np.abs(kg - ko) / ko) < 1e-13
    # the matrix from one field is bit-exact
    g = T.Matrix(T.CSR).fill_poisson3d(*dims, kappa=dk)
    o = O.poisson3d(*dims, kappa=kg)
    assert same_csr(g, o)
    # symmetric, diagonally dominant
    rp, ci, va = g.arrays()
    A = sp.csr_matrix((va, ci, rp))
    assert abs(A - A.T).max() == 0.0
    d = A.diagonal()
    off = np.asarray(abs(A).sum(axis=1)).ravel() - d
    assert np.all(d >= off - 1e-12 * d)


@pytest.mark.parametrize("n,c,q,s", [(1, 1, 0.0, 0), (50, 60, 0.5, 1), (3000, 16, 0.05, 0),
                                     (20000, 16, 0.05, 0), (20000, 8, 0.1, 3)])
def test_sym_cqmat_bit_exact(n, c, q, s):
    g = T.Matrix(T.CSR).fill_sym_cqmat(n, c, q, s, value_seed=4)
    o = O.sym_cqmat(n, c, q, s, value_seed=4)
    assert same_csr(g, o)


def test_transpose_bit_exact():
    for (n, c, q, grid) in [(3000, 12, 0.3, (0, 0, 1, 1)), (4000, 9, 0.2, (1, 2, 3, 4))]:
        g = T.Matrix(T.CSR).fill_cqmat(n, n, c, q, 0, *grid, value_seed=2)
        o = O.fill_cqmat(O.CSR, n, n, c, q, 0, *grid, value_seed=2)
        gt, ot = g.transpose(), O.transpose(o)
        assert same_csr(gt, ot)
        assert gt.geom == ot.geom


def test_block_diag():
    parts = [T.Matrix(T.CSR).fill_sym_cqmat(n, 5, 0.1, 1, value_seed=n) for n in (10, 1, 333, 64)]
    bd = T.Matrix.block_diag(parts)
    dense = [sp.csr_matrix((a[2], a[1], a[0])).toarray() for a in (p.arrays() for p in parts)]
    rp, ci, va = bd.arrays()
    got = sp.csr_matrix((va, ci, rp), shape=(408, 408)).toarray()
    from scipy.linalg import block_diag
    assert np.array_equal(got, block_diag(*dense))


def test_uniform_vector_bit_exact():
    for lo, hi in [(0.0, 1.0), (-1.0, 1.0)]:
        g = T.uniform_vector(10001, 17, lo, hi).to_numpy(count=10001)
        assert np.array_equal(g, O.uniform_vector(10001, 17, lo, hi))
    assert np.array_equal(T.uniform_vector(999, 5).to_numpy(count=999), O.unit_vector(5, 999))
