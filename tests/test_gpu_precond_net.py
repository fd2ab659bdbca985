"""The learned preconditioner's GPU kernels against the oracle
(oracle/precond_net.py, float64 with the GPU's BF16 roundings mirrored).

* Initial parameters: the oracle's draws, cast to FP32.
* Forward: G within 1e-4 (relative, Frobenius) of the oracle's.
* Loss and gradient: loss within 1e-5, gradient within 1e-3 (relative,
  Euclidean, and per tensor within 1e-2) -- FP32 against float64 arithmetic on
  the same BF16-rounded GEMM operands.
* At initialisation G G^T is Jacobi: CG with it follows Jacobi-PCG.
* Training steps are deterministic and lower the loss.
"""
import numpy as np
import pytest
import scipy.sparse as sp

import oracle as O
from oracle import precond_net as PN
import tbsla_amd as T

pytestmark = pytest.mark.gpu


def problem(dims, seed):
    dk = T.lognormal_field(*dims, seed=seed, corr_len=2.0)
    return T.Matrix(T.CSR).fill_poisson3d(*dims, kappa=dk)


def random_params(seed):
    rng = np.random.default_rng(seed)
    p = PN.unflatten(PN.init_params(seed))
    p["wo"] = rng.standard_normal(64) * 0.05
    p["bo"] = np.array(0.02)
    p["wd"] = rng.standard_normal(64) * 0.05
    p["bd"] = np.array(-0.05)
    p["b1"] = rng.standard_normal(64) * 0.05
    p["c"] = rng.standard_normal(64) * 0.05
    return PN.flatten(p).astype(np.float32).astype(np.float64)


def graph_of(m):
    rp, ci, va = m.arrays()
    return PN.Graph(rp, ci, va)


def rel(a, b):
    return float(np.linalg.norm(np.ravel(a - b)) / np.linalg.norm(np.ravel(b)))


def test_initial_parameters():
    net = T.PrecondNet(seed=5)
    assert np.array_equal(net.get_params(), PN.init_params(5).astype(np.float32))


@pytest.mark.parametrize("kind", ["poisson", "sym"])
def test_forward_matches_oracle(kind):
    A = problem((12, 10, 8), 3) if kind == "poisson" else \
        T.Matrix(T.CSR).fill_sym_cqmat(900, 8, 0.1, 1, value_seed=2)
    theta = random_params(4)
    net = T.PrecondNet()
    net.set_params(theta.astype(np.float32))
    G, GT = net.apply(A)
    g = graph_of(A)
    gd, gl, _ = PN.forward(PN.unflatten(theta), g, bf=True)
    Go = PN.G_csr(g, gd, gl)
    rp, ci, va = G.arrays()
    assert np.array_equal(rp, Go.indptr) and np.array_equal(ci, Go.indices)
    assert rel(va, Go.data) < 1e-4
    # G^T is the transpose, bit for bit
    n = len(rp) - 1
    Gd = sp.csr_matrix((va, ci, rp), shape=(n, n))
    trp, tci, tva = GT.arrays()
    Gt = sp.csr_matrix((tva, tci, trp), shape=(n, n))
    assert (Gt - Gd.T).count_nonzero() == 0 or abs(Gt - Gd.T).max() == 0.0


def test_loss_and_gradient_match_oracle():
    A = problem((10, 9, 8), 7)
    theta = random_params(8)
    net = T.PrecondNet()
    net.set_params(theta.astype(np.float32))
    L, grad = net.loss_grad(A, 8, 123)
    g = graph_of(A)
    Z = PN.rademacher(g.n, 8, 123)
    Lo, go = PN.loss_grad(theta, g, Z, bf=True)
    assert abs(L - Lo) <= 1e-5 * Lo
    assert rel(grad, go) < 1e-3, rel(grad, go)
    o = 0
    for name, shp in PN.SHAPES:
        k = int(np.prod(shp))
        if np.linalg.norm(go[o:o + k]) > 1e-6 * np.linalg.norm(go):
            assert rel(grad[o:o + k], go[o:o + k]) < 1e-2, name
        o += k


def test_initial_network_is_jacobi():
    A = problem((16, 16, 16), 1)
    net = T.PrecondNet(seed=0)
    G, GT = net.apply(A)
    rp, ci, va = G.arrays()
    n = len(rp) - 1
    diag = np.repeat(np.arange(n), np.diff(rp)) == ci
    assert np.all(va[~diag] == 0.0)
    rhs = T.uniform_vector(n, 2, -1.0, 1.0).to_numpy(count=n)
    xj, itj, _ = A.cg(rhs, tolerance=1e-8, relative=True, preconditioner=T.PRECOND_JACOBI)
    xg, itg, _ = A.cg(rhs, tolerance=1e-8, relative=True, preconditioner=T.PRECOND_FACTORED,
                      factor=G, factor_t=GT)
    assert abs(itg - itj) <= 1
    assert np.max(np.abs(xg - xj)) / np.max(np.abs(xj)) < 1e-7


def test_training_is_deterministic_and_lowers_the_loss():
    probs = [problem((12, 12, 12), s) for s in range(4)]
    nets = [T.PrecondNet(seed=1), T.PrecondNet(seed=1)]
    losses = [[], []]
    for k, net in enumerate(nets):
        for step in range(40):
            losses[k].append(net.train_step(probs[step % 4], 8, 1000 + step, 2e-3, 1.0))
    assert np.array_equal(nets[0].get_params(), nets[1].get_params())
    assert losses[0] == losses[1]
    first, last = np.mean(losses[0][:4]), np.mean(losses[0][-4:])
    assert last < 0.8 * first, (first, last)
