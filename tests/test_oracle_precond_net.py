"""Pins for the oracle of the learned preconditioner (CPU only).

* bf16 rounding: nearest-even on known values.
* The analytic gradient of the float64 network equals central finite
  differences (relative to the largest component, 1e-8).
* Zero output heads give G^ = I, and the loss is then |I - A^|_F^2-like:
  for k probes its expectation is sum_{i != j} A^_ij^2 / n (checked with many
  probes).
* The Rademacher probes are +-1 with the documented generator.
"""
import numpy as np

import oracle as O
from oracle import precond_net as PN


def small_graph():
    k = O.lognormal_field(6, 5, 4, 0.0, 1.0, 1.5, 3)
    m = O.poisson3d(6, 5, 4, kappa=k)
    return PN.Graph(*m.arrays())


def test_bf16_rounding():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -9, 1.0 + 2 ** -9, -2.5, 0.0, 3e38])
    r = PN.bf16(x)
    assert r[0] == 1.0
    assert r[1] == 1.0  # 1 + 2^-8 is halfway between 1 and 1 + 2^-7: even is 1
    assert r[2] == 1.0 + 2 ** -7
    assert r[3] == 1.0
    assert r[4] == -2.5 and r[5] == 0.0


def test_gradient_finite_differences():
    g = small_graph()
    rng = np.random.default_rng(0)
    p = PN.unflatten(PN.init_params(7))
    p["wo"] = rng.standard_normal(64) * 0.05
    p["bo"] = np.array(0.01)
    p["wd"] = rng.standard_normal(64) * 0.05
    p["bd"] = np.array(-0.1)
    theta = PN.flatten(p)
    Z = PN.rademacher(g.n, 4, 11)
    L, gr = PN.loss_grad(theta, g, Z)
    idx = np.concatenate([rng.choice(PN.N_PARAMS, 40, replace=False),
                          [PN.N_PARAMS - 1, PN.N_PARAMS - 66, 0, 520]])
    for i in idx:
        e = np.zeros_like(theta)
        h = 1e-6 * max(1.0, abs(theta[i]))
        e[i] = h
        fd = (PN.loss_grad(theta + e, g, Z)[0] - PN.loss_grad(theta - e, g, Z)[0]) / (2 * h)
        assert abs(fd - gr[i]) < 1e-8 * np.abs(gr).max(), i


def test_zero_heads_are_identity_and_loss_expectation():
    g = small_graph()
    gd, gl, _ = PN.forward(PN.unflatten(PN.init_params(3)), g)
    assert np.all(gd == 1.0) and np.all(gl == 0.0)
    # E_Z |Z - A^ Z|^2 / (n k) = sum_{i != j} A^_ij^2 / n
    Z = PN.rademacher(g.n, 4096, 5)
    L, _, _ = PN.loss_and_ghat_grad(g, gd, gl, Z)
    # G^ = I: loss uses A^ (with unit diagonal) -> (I - A^) = -E
    expect = (g.ev ** 2).sum() / g.n
    assert abs(L - expect) < 0.03 * expect


def test_rademacher():
    Z = PN.rademacher(100, 8, 9)
    assert set(np.unique(Z)) == {-1.0, 1.0}
    u = O.unit_vector(9, 800).reshape(100, 8)
    assert np.array_equal(Z, np.where(u < 0.5, -1.0, 1.0))
