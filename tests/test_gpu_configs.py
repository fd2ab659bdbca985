"""The configured workloads at their full sizes, on the GPU.

Each test runs the HIP path on the workload the benchmark times and checks
it: against the oracle where the oracle finishes in seconds at that size
(fills, early iterates), and otherwise through properties fixed by the
mathematics (true residuals of converged solves, fixed points of PageRank,
column sums, agreement of batched and single solves).

* cqmat n = 2^22, c = 32, q = 0.1 (values U[0,1) seed 0), column-normalised:
  PageRank (beta 0.85, L1 change < 1e-10); the 4x4 grid's blocks (0,0) and
  (0,1), each in the layout the autotuner picks.
* CG: Poisson 128^3 and the symmetrised cqmat n = 2^21, c = 16, q = 0.05,
  plain FP64 and mixed precision to a relative residual of 1e-8; Poisson
  32^3 against a sparse direct solve.
* 512 symmetrised cqmat systems of uneven size in one batched mixed CG;
  personalised PageRank with 64 one-hot teleports at n = 2^22.
* The learned preconditioner's committed weights on held-out 64^3 and 128^3
  variable-coefficient Poisson problems, and its Jacobi fallback.
"""
import sys

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

import oracle as O
import tbsla_amd as T
from tbsla_amd import batched, learned

pytestmark = pytest.mark.gpu

N2 = 1 << 22


def say(*a):
    print(*a, file=sys.stderr, flush=True)


def rel_max(a, b):
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


def csr_scipy(m):
    rp, ci, va = m.arrays()
    if isinstance(m, T.Matrix):
        geo = m.geometry()
        nr, nc, fc = geo.ln_row, geo.ln_col, geo.f_col
    else:
        nr, nc, fc = m.geom["ln_row"], m.geom["ln_col"], m.geom["f_col"]
    return sp.csr_matrix((va, ci - fc, rp), shape=(nr, nc))


# ---------------------------------------------------------------------------
# cfg2: cqmat n = 2^22, c = 32, q = 0.1
# ---------------------------------------------------------------------------

@pytest.fixture(scope="module")
def stochastic_cfg2():
    say("cfg2 stochastic fixture: n=2^22 c=32 q=0.1 value seed 0")
    o = O.fill_cqmat(O.CSR, N2, N2, 32, 0.1, 0, value_seed=0)
    o.make_stochastic()
    g = T.Matrix(T.CSR).fill_cqmat(N2, N2, 32, 0.1, 0, value_seed=0)
    g.make_stochastic()
    return o, g


def test_cfg2_stochastic_matrix_bit_exact(stochastic_cfg2):
    o, g = stochastic_cfg2
    for a, b in zip(g.arrays(), o.arrays()):
        assert np.array_equal(a, b)


def test_cfg2_page_rank(stochastic_cfg2):
    """Early iterates within 1e-12 of the oracle; at convergence (65
    iterations, the oracle's own count for this matrix) the vector sums to
    1, is non-negative and is the fixed point of the reference iteration:
    beta A b + (1 - beta)/n is proportional to b (Matrix.cpp:106-171)."""
    o, g = stochastic_cfg2
    for k in (1, 3):
        say(f"cfg2 page_rank: {k} iterations vs oracle")
        bo, ito, _ = o.page_rank(0.85, 0.0, k)
        bg, itg, _ = g.page_rank(0.85, 0.0, k)
        assert ito == itg == k
        assert rel_max(bg, bo) < 1e-12
    say("cfg2 page_rank: to 1e-10")
    b, it, err = g.page_rank(0.85, 1e-10, 10000)
    assert err < 1e-10 and it == 65
    assert abs(b.sum() - 1.0) < 1e-12 and b.min() >= 0.0
    A = csr_scipy(g)
    v = 0.85 * (A @ b) + 0.15 / N2
    assert np.abs(v / v.sum() - b).sum() < 1e-9
    # the diagonal layout's LDS-slice form: each iteration one fused pass
    say("cfg2 page_rank: diagonal layout, fused update")
    g.set_csr_layout(T.CSR_DIAG, 65536, 16)
    try:
        assert g.get_csr_layout()["lds_block_rows"] > 0
        bo, _, _ = o.page_rank(0.85, 0.0, 3)
        bg, itg, _ = g.page_rank(0.85, 0.0, 3)
        assert itg == 3 and rel_max(bg, bo) < 1e-12
        bd, itd, errd = g.page_rank(0.85, 1e-10, 10000)
        assert errd < 1e-10 and itd == 65
        assert rel_max(bd, b) < 1e-12
    finally:
        g.set_csr_layout(T.CSR_ROWS)


@pytest.mark.parametrize("pc", [0, 1])
def test_cfg2_grid_block(pc):
    """Block (0, pc) of the 4x4 grid at n = 2^22: structure bit-exact, and
    SpMV within 1e-12 of the oracle both in the row layout and in the
    layout the autotuner picks (the diagonal layout for (0,0))."""
    say(f"cfg2 4x4 block (0,{pc})")
    o = O.fill_cqmat(O.CSR, N2, N2, 32, 0.1, 0, 0, pc, 4, 4, value_seed=0)
    g = T.Matrix(T.CSR).fill_cqmat(N2, N2, 32, 0.1, 0, 0, pc, 4, 4, value_seed=0)
    assert o.geom["ln_row"] == g.geometry().ln_row == N2 // 4
    for a, b in zip(g.arrays(), o.arrays()):
        assert np.array_equal(a, b)
    x = O.unit_vector(1, o.geom["ln_col"])
    yo = o.spmv(x)
    scale = max(float(np.max(np.abs(yo))), 1e-300)
    assert float(np.max(np.abs(g.spmv(x) - yo))) / scale < 1e-12
    g.tune_layout()
    say(f"  autotuned layout {g.get_csr_layout()}")
    assert float(np.max(np.abs(g.spmv(x) - yo))) / scale < 1e-12


# ---------------------------------------------------------------------------
# cfg3: CG on SPD systems
# ---------------------------------------------------------------------------

def true_rel_res(A, x, b):
    return float(np.linalg.norm(b - A @ x) / np.linalg.norm(b))


def test_cfg3_poisson128():
    """7-point Poisson 128^3 (Dirichlet): the matrix bit-exact with the
    oracle; plain FP64 CG's first 10 iterates within 1e-12 of the oracle's
    CG; plain FP64 and mixed CG (FP32 matrix, FP64 dots, true residual
    every 50) to a relative residual of 1e-8, checked on the true FP64
    residual, on the row layout and on the autotuned diagonal layout."""
    say("cfg3 Poisson 128^3")
    o = O.poisson3d(128, 128, 128)
    g = T.Matrix(T.CSR).fill_poisson3d(128, 128, 128)
    for a, b in zip(g.arrays(), o.arrays()):
        assert np.array_equal(a, b)
    n = 128 ** 3
    rhs = T.uniform_vector(n, 3, -1.0, 1.0).to_numpy(count=n)
    xo, ito, _ = O.pcg(o, rhs, max_iterations=10, tolerance=1e-30, kind=0)
    xg, itg, _ = g.cg(rhs, max_iterations=10, tolerance=1e-30, relative=True)
    assert ito == itg == 10 and rel_max(xg, xo) < 1e-12
    A = csr_scipy(o)
    its = {}
    for layout in ("rows", "tuned"):
        if layout == "tuned":
            g.tune_layout()
        for prec in (T.CG_FP64, T.CG_MIXED):
            x, it, _ = g.cg(rhs, tolerance=1e-8, relative=True, precision=prec,
                            replace_every=50 if prec == T.CG_MIXED else 0)
            r = true_rel_res(A, x, rhs)
            say(f"  {layout} prec {prec}: {it} iterations, true rel res {r:.3e}")
            assert r <= 1e-8
            its[(layout, prec)] = it
    assert abs(its[("rows", T.CG_FP64)] - its[("tuned", T.CG_FP64)]) <= 2


def test_cfg3_sym_cqmat():
    """Symmetrised cqmat n = 2^21, c = 16, q = 0.05: bit-exact with the
    oracle; the first 10 plain CG iterates within 1e-12; plain FP64 and
    mixed CG to a relative residual of 1e-8 on the true residual."""
    n = 1 << 21
    say("cfg3 symmetrised cqmat n=2^21 c=16 q=0.05")
    o = O.sym_cqmat(n, 16, 0.05, 1, value_seed=0)
    g = T.Matrix(T.CSR).fill_sym_cqmat(n, 16, 0.05, 1, value_seed=0)
    for a, b in zip(g.arrays(), o.arrays()):
        assert np.array_equal(a, b)
    rhs = T.uniform_vector(n, 4, -1.0, 1.0).to_numpy(count=n)
    xo, ito, _ = O.pcg(o, rhs, max_iterations=10, tolerance=1e-30, kind=0)
    xg, itg, _ = g.cg(rhs, max_iterations=10, tolerance=1e-30, relative=True)
    assert ito == itg == 10 and rel_max(xg, xo) < 1e-12
    A = csr_scipy(o)
    for prec in (T.CG_FP64, T.CG_MIXED):
        x, it, _ = g.cg(rhs, tolerance=1e-8, relative=True, precision=prec,
                        replace_every=50 if prec == T.CG_MIXED else 0)
        r = true_rel_res(A, x, rhs)
        say(f"  prec {prec}: {it} iterations, true rel res {r:.3e}")
        assert r <= 1e-8


def test_cfg3_poisson32_direct():
    """Poisson 32^3 to a relative residual of 1e-8 against scipy's sparse
    direct solve: the error is bounded by cond(A) x 1e-8 (cond ~ 440 for
    the 32^3 Dirichlet Laplacian): measured ~4e-8 relative, asserted 1e-6."""
    say("cfg3 Poisson 32^3 direct")
    g = T.Matrix(T.CSR).fill_poisson3d(32, 32, 32)
    A = csr_scipy(g).tocsc()
    n = A.shape[0]
    rhs = T.uniform_vector(n, 3, -1.0, 1.0).to_numpy(count=n)
    exact = spla.spsolve(A, rhs)
    for prec in (T.CG_FP64, T.CG_MIXED):
        x, it, _ = g.cg(rhs, tolerance=1e-8, relative=True, precision=prec,
                        replace_every=50 if prec == T.CG_MIXED else 0)
        err = rel_max(x, exact)
        say(f"  prec {prec}: {it} iterations, max rel error vs direct {err:.3e}")
        assert err < 1e-6


# ---------------------------------------------------------------------------
# cfg4: batched workloads
# ---------------------------------------------------------------------------

def test_cfg4_batched_cg_512():
    """512 SPD systems (n = 2^U(12,18), c = 8..32, q = 0.1) in one batched
    CG.  Mixed precision (the benchmark's form): every system converged with
    a true FP64 relative residual <= 1e-8 (A x by the library's FP64 SpMV of
    the whole block).  FP64: 8 sampled systems bit-identical to solving each
    alone (mixed solves verify on a schedule that depends on the other
    systems, so only their converged residuals are compared)."""
    say("cfg4 batched CG: building 512 systems")
    bd, offs, rhs, n, c = batched.build(512)
    lanes = 8
    nrows = int(offs[-1])
    x = T.DeviceBuffer(nrows * 8)
    _, res, x = batched.run(bd, offs, rhs, batched.options(bd, lanes), x)
    st = np.array([r.status for r in res])
    assert np.all(st == T.CG_CONVERGED), np.unique(st, return_counts=True)
    y = T.DeviceBuffer(nrows * 8)
    T.check(T.lib().tbsla_hip_spmv(bd.h, y.ptr, x.ptr, 0, None))
    b, ax = rhs.to_numpy(count=nrows), y.to_numpy(count=nrows)
    worst = max(float(np.linalg.norm(b[offs[k]:offs[k + 1]] - ax[offs[k]:offs[k + 1]]) /
                      np.linalg.norm(b[offs[k]:offs[k + 1]])) for k in range(512))
    say(f"  mixed: worst true rel res {worst:.3e}, iterations {min(r.iterations for r in res)}"
        f"..{max(r.iterations for r in res)}")
    assert worst <= 1e-8
    opt64 = bd.cg_options(5000, 1e-8, True, T.CG_FP64, T.PRECOND_NONE, 0, lanes_per_row=lanes)
    x64 = T.DeviceBuffer(nrows * 8)
    _, res64, x64 = batched.run(bd, offs, rhs, opt64, x64)
    xs = x64.to_numpy(count=nrows)
    rng = np.random.default_rng(5)
    for k in sorted(rng.choice(512, 8, replace=False)):
        p = T.Matrix(T.CSR).fill_sym_cqmat(int(n[k]), int(c[k]), 0.1, 1, value_seed=int(k))
        xk, itk, rk = p.cg(b[offs[k]:offs[k + 1]], tolerance=1e-8, relative=True,
                           lanes_per_row=lanes, max_iterations=5000,
                           max_blocks_per_system=256)
        assert itk == res64[k].iterations and rk == res64[k].res_norm
        assert np.array_equal(xk, xs[offs[k]:offs[k + 1]])


def test_cfg4_personalized_page_rank(stochastic_cfg2):
    """64 one-hot teleports at n = 2^22 as one SpMM per iteration: the first
    iterates of two columns within 1e-12 of the oracle's single-vector
    iteration with that teleport; at convergence (max L1 change < 1e-10)
    every column sums to 1, is non-negative, and is the fixed point
    beta A b + (1 - beta) e_s proportional to b."""
    o, g = stochastic_cfg2
    rng = np.random.default_rng(3)
    src = np.sort(rng.choice(N2, 64, replace=False)).astype(np.int32)
    say("cfg4 PPR: 4 iterations vs oracle")
    X, it, _ = g.personalized_page_rank(src, 0.85, 0.0, 4)
    assert it == 4
    for j in (0, 37):
        bo, ito, _ = o.page_rank_source(int(src[j]), 0.85, 0.0, 4)
        assert ito == 4 and rel_max(X[:, j], bo) < 1e-12
    del X
    say("cfg4 PPR: to 1e-10")
    X, it, err = g.personalized_page_rank(src, 0.85, 1e-10, 10000)
    assert err < 1e-10
    say(f"  {it} iterations")
    sums = X.sum(axis=0)
    assert np.max(np.abs(sums - 1.0)) < 1e-12 and X.min() >= 0.0
    A = csr_scipy(g)
    for j in (5, 63):
        v = 0.85 * (A @ X[:, j])
        v[src[j]] += 0.15
        assert np.abs(v / v.sum() - X[:, j]).sum() < 1e-9


# ---------------------------------------------------------------------------
# cfg5: the learned preconditioner
# ---------------------------------------------------------------------------

@pytest.fixture(scope="module")
def net():
    n = T.PrecondNet()
    assert learned.load_weights(n), "committed weights missing"
    return n


@pytest.mark.parametrize("side,seeds", [(64, (100000, 100001, 100002, 100003)), (128, (100000,))])
def test_cfg5_learned_gain(net, side, seeds):
    """The committed network on held-out variable-coefficient Poisson
    problems: learned mixed and FP64 PCG reach a true relative residual of
    1e-8 in at most 0.6x the iterations of Jacobi-PCG, without falling
    back."""
    factors = (T.Matrix(T.CSR), T.Matrix(T.CSR))
    for sd in seeds:
        say(f"cfg5 learned {side}^3 seed {sd}")
        A, rhs = learned.poisson_problem(side, sd)
        A.tune_layout()
        s = learned.Solver(A, rhs, net, factors)
        rj = s.run("jacobi_mixed")
        for m in ("learned_mixed", "learned_fp64"):
            rl = s.run(m)
            say(f"  {m}: {rl['iterations']} vs Jacobi {rj['iterations']}, "
                f"res {rl['true_rel_res']:.2e}, {rl['ms']:.2f} ms vs {rj['ms']:.2f} ms")
            assert not rl["fallback"]
            assert rl["true_rel_res"] <= 1e-8
            assert rl["iterations"] <= 0.6 * rj["iterations"]


def test_cfg5_fallback_to_jacobi(net):
    """A factor that cannot precondition makes the learned solve fall back
    to Jacobi-PCG, whose result it then returns bit for bit: G = 0 (r.z = 0,
    a breakdown at the first test) and a network whose parameters are NaN
    (a non-finite factor)."""
    A, rhs = learned.poisson_problem(24, 100000)
    s = learned.Solver(A, rhs, net)
    ref = s.run("jacobi_mixed")
    x_ref = s.x.to_numpy(count=s.n)
    G = T.Matrix(T.CSR).fill_poisson3d(24, 24, 24)
    rp, ci, va = G.arrays()
    geo = G.geometry()
    geom = {f: getattr(geo, f) for f in T.GEOM_FIELDS}
    geom["nnz"] = geo.nnz
    Z = T.Matrix.from_arrays(T.CSR, geom, rp, ci, np.zeros_like(va))
    r = s.run_factored("learned_mixed", Z, None)
    assert r["fallback"] and r["iterations"] == ref["iterations"]
    assert np.array_equal(s.x.to_numpy(count=s.n), x_ref)
    bad = T.PrecondNet()
    bad.set_params(np.full(T.PrecondNet.param_count(), np.nan, np.float32))
    s2 = learned.Solver(A, rhs, bad)
    r2 = s2.run("learned_mixed")
    assert r2["fallback"] and r2["iterations"] == ref["iterations"]
    assert np.array_equal(s2.x.to_numpy(count=s2.n), x_ref)
