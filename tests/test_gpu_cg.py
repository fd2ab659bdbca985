"""The fused CG engine on the GPU (tbsla_hip_cg / tbsla_hip_cg_batched).

* Preconditioned iterates (Jacobi, M = G G^T) within a relative 1e-12 of the
  oracle's PCG (orc_pcg) for the first iterations; the same iteration count
  at convergence within two.
* hipGraph batches and direct launches give identical bits; repeated solves
  give identical bits.
* A batch of systems gives, for each system, the bits of solving it alone.
* Mixed precision: the true FP64 residual meets the tolerance.
* Poisson 16^3 against a sparse direct solve.
"""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

import oracle as O
import tbsla_amd as T

pytestmark = pytest.mark.gpu


def rel_max(a, b):
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


def to_gpu(o):
    rp, ci, va = o.arrays()
    return T.Matrix.from_arrays(T.CSR, o.geom, rp, ci, va)


def scipy_of(m):
    rp, ci, va = m.arrays()
    n = len(rp) - 1
    return sp.csr_matrix((va, ci, rp), shape=(n, n))


def lower_factor(o):
    """A lower-triangular G on tril(A)'s pattern with a positive diagonal:
    G = tril(A) scaled by D^-1 (a crude SPD-preserving factor)."""
    A = scipy_of(o)
    d = A.diagonal()
    L = sp.tril(A).tocsr()
    G = sp.diags(1.0 / np.sqrt(d)) @ (sp.diags(1.0 / d) @ L) * 0.5 + sp.diags(0.5 / np.sqrt(d))
    G = sp.csr_matrix(G)
    G.sort_indices()
    return O.csr_from_arrays(G.shape[0], G.shape[1], G.indptr, G.indices, G.data)


@pytest.fixture(scope="module")
def sym():
    o = O.sym_cqmat(6000, 12, 0.1, 0, value_seed=3)
    return o, to_gpu(o)


def test_jacobi_early_iterates(sym):
    o, g = sym
    n = o.geom["ln_row"]
    rhs = 2 * O.unit_vector(21, n) - 1
    for k in (1, 2, 5, 12):
        xo, ito, ro = O.pcg(o, rhs, max_iterations=k, tolerance=1e-30, kind=1)
        xg, itg, rg = g.cg(rhs, max_iterations=k, tolerance=1e-30, relative=True,
                           preconditioner=T.PRECOND_JACOBI)
        assert itg == ito == k
        assert rel_max(xg, xo) < 1e-12, (k, rel_max(xg, xo))
        assert abs(rg - ro) <= 1e-10 * ro


def test_factored_early_iterates_and_given_transpose(sym):
    o, g = sym
    n = o.geom["ln_row"]
    Go = lower_factor(o)
    G = to_gpu(Go)
    Gt = G.transpose()
    rhs = 2 * O.unit_vector(5, n) - 1
    for k in (1, 3, 8):
        xo, ito, _ = O.pcg(o, rhs, max_iterations=k, tolerance=1e-30, kind=2, G=Go)
        xg, itg, _ = g.cg(rhs, max_iterations=k, tolerance=1e-30, relative=True,
                          preconditioner=T.PRECOND_FACTORED, factor=G)
        xt, itt, _ = g.cg(rhs, max_iterations=k, tolerance=1e-30, relative=True,
                          preconditioner=T.PRECOND_FACTORED, factor=G, factor_t=Gt)
        assert itg == ito == itt == k
        assert rel_max(xg, xo) < 1e-12, (k, rel_max(xg, xo))
        assert np.array_equal(xg, xt)


@pytest.fixture(scope="module")
def poisson():
    k = O.lognormal_field(16, 16, 16, 0.0, 1.0, 4.0, 7)
    o = O.poisson3d(16, 16, 16, kappa=k)
    return o, to_gpu(o)


@pytest.mark.parametrize("precision", [T.CG_FP64, T.CG_MIXED])
def test_factored_diagonals_and_rows_identical(poisson, sym, precision):
    """A stencil's lower factor (4 diagonals of offset <= 0) is applied from
    its diagonals by default; that form sums G^T's and G's rows in the order
    of the CSR form, so both give bitwise the same iterates (with residual
    replacement every 7 in the mixed form), and they follow the oracle's
    G G^T PCG (FP64 early iterates within 1e-12).  A factor off a few
    diagonals (the symmetrised cqmat's) runs the CSR form under AUTO and is
    refused under DIAGONALS."""
    o, g = poisson
    n = o.geom["ln_row"]
    Go = lower_factor(o)
    G = to_gpu(Go)
    rhs = 2 * O.unit_vector(19, n) - 1
    if precision == T.CG_FP64:
        for k in (1, 3, 8):
            xo, ito, _ = O.pcg(o, rhs, max_iterations=k, tolerance=1e-30, kind=2, G=Go)
            xg, itg, _ = g.cg(rhs, max_iterations=k, tolerance=1e-30, relative=True,
                              preconditioner=T.PRECOND_FACTORED, factor=G,
                              factor_form=T.FACTOR_DIAGONALS)
            assert itg == ito == k
            assert rel_max(xg, xo) < 1e-12, (k, rel_max(xg, xo))
    rep = 7 if precision == T.CG_MIXED else 0
    runs = [g.cg(rhs, tolerance=1e-10, relative=True, precision=precision, replace_every=rep,
                 preconditioner=T.PRECOND_FACTORED, factor=G, factor_form=form)
            for form in (T.FACTOR_ROWS, T.FACTOR_DIAGONALS, T.FACTOR_AUTO)]
    for x, it, rn in runs[1:]:
        assert it == runs[0][1] and rn == runs[0][2]
        assert np.array_equal(x, runs[0][0])
    xo, ito, _ = O.pcg(o, rhs, tolerance=1e-10, kind=2, G=Go)
    assert abs(runs[0][1] - ito) <= 2
    A = scipy_of(o)
    assert np.linalg.norm(rhs - A @ runs[0][0]) < 1e-10 * np.linalg.norm(rhs) * 1.0001
    so, sg = sym
    Gs = to_gpu(lower_factor(so))
    rs = 2 * O.unit_vector(20, so.geom["ln_row"]) - 1
    a = sg.cg(rs, tolerance=1e-10, relative=True, precision=precision, replace_every=rep,
              preconditioner=T.PRECOND_FACTORED, factor=Gs, factor_form=T.FACTOR_AUTO)
    b = sg.cg(rs, tolerance=1e-10, relative=True, precision=precision, replace_every=rep,
              preconditioner=T.PRECOND_FACTORED, factor=Gs, factor_form=T.FACTOR_ROWS)
    assert a[1] == b[1] and np.array_equal(a[0], b[0])
    with pytest.raises(T.TbslaHipError) as e:
        sg.cg(rs, preconditioner=T.PRECOND_FACTORED, factor=Gs, factor_form=T.FACTOR_DIAGONALS)
    assert e.value.status == T.ERR_UNSUPPORTED
    with pytest.raises(T.TbslaHipError):
        g.cg(rhs, preconditioner=T.PRECOND_FACTORED, factor=G, factor_form=3)


@pytest.mark.parametrize("precision", [T.CG_FP64, T.CG_MIXED])
def test_cache_of_destroyed_handle_reused(precision):
    """A destroyed handle's CG state (stream, vectors, captured graphs) serves
    the next handles: a graph of the same launch shape is replayed as it is
    (its kernels read the new matrix's pointers from the argument block),
    others are instantiated.  Every solve still follows the oracle."""
    import gc
    rep = 10 if precision == T.CG_MIXED else 0
    for side, seed in ((12, 1), (12, 2), (14, 3), (12, 4)):
        k = O.lognormal_field(side, side, side, 0.0, 1.0, 4.0, seed)
        o = O.poisson3d(side, side, side, kappa=k)
        g = to_gpu(o)
        n = o.geom["ln_row"]
        rhs = 2 * O.unit_vector(30 + seed, n) - 1
        if precision == T.CG_FP64:
            xo, ito, _ = O.pcg(o, rhs, max_iterations=6, tolerance=1e-30, kind=1)
            xg, itg, _ = g.cg(rhs, max_iterations=6, tolerance=1e-30, relative=True,
                              preconditioner=T.PRECOND_JACOBI)
            assert itg == ito == 6 and rel_max(xg, xo) < 1e-12
        x, it, rn = g.cg(rhs, tolerance=1e-10, relative=True, precision=precision,
                         replace_every=rep, preconditioner=T.PRECOND_JACOBI)
        xo, ito, _ = O.pcg(o, rhs, tolerance=1e-10, kind=1)
        assert abs(it - ito) <= 2
        A = scipy_of(o)
        assert np.linalg.norm(rhs - A @ x) < 1e-10 * np.linalg.norm(rhs) * 1.0001
        del g
        gc.collect()


@pytest.mark.parametrize("pre", [T.PRECOND_NONE, T.PRECOND_JACOBI, T.PRECOND_FACTORED])
def test_converged_iterations_like_oracle(sym, pre):
    o, g = sym
    n = o.geom["ln_row"]
    rhs = 2 * O.unit_vector(8, n) - 1
    Go = lower_factor(o) if pre == T.PRECOND_FACTORED else None
    G = to_gpu(Go) if Go is not None else None
    xo, ito, _ = O.pcg(o, rhs, tolerance=1e-10, kind=pre, G=Go)
    xg, itg, rg = g.cg(rhs, tolerance=1e-10, relative=True, preconditioner=pre, factor=G)
    assert abs(itg - ito) <= 2, (itg, ito)
    assert rel_max(xg, xo) < 1e-8
    A = scipy_of(o)
    assert np.linalg.norm(rhs - A @ xg) < 1e-10 * np.linalg.norm(rhs) * 1.0001


@pytest.mark.parametrize("pre", [T.PRECOND_NONE, T.PRECOND_JACOBI])
def test_graph_and_direct_launches_identical(sym, pre):
    o, g = sym
    n = o.geom["ln_row"]
    rhs = 2 * O.unit_vector(9, n) - 1
    a = g.cg(rhs, tolerance=1e-11, relative=True, preconditioner=pre, graph_batch=0)
    b = g.cg(rhs, tolerance=1e-11, relative=True, preconditioner=pre, graph_batch=-1)
    c = g.cg(rhs, tolerance=1e-11, relative=True, preconditioner=pre, graph_batch=6)
    d = g.cg(rhs, tolerance=1e-11, relative=True, preconditioner=pre, graph_batch=0)
    assert a[1] == b[1] == c[1] == d[1]
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[0], c[0]) and np.array_equal(a[0], d[0])


def test_max_iterations_count(sym):
    o, g = sym
    n = o.geom["ln_row"]
    rhs = 2 * O.unit_vector(2, n) - 1
    for k in (0, 1, 17, 33):
        _, it, _ = g.cg(rhs, max_iterations=k, tolerance=1e-30, relative=True)
        _, ito, _ = O.pcg(o, rhs, max_iterations=k, tolerance=1e-30)
        assert it == ito == k


def make_batch(sizes, seed=0):
    parts = [T.Matrix(T.CSR).fill_sym_cqmat(n, c, 0.1, 1, value_seed=seed + i)
             for i, (n, c) in enumerate(sizes)]
    offs = np.concatenate([[0], np.cumsum([n for n, _ in sizes])]).astype(np.int32)
    return parts, T.Matrix.block_diag(parts), offs


@pytest.mark.parametrize("pre", [T.PRECOND_NONE, T.PRECOND_JACOBI])
def test_batched_equals_single(pre):
    sizes = [(4096, 8), (777, 31), (20000, 16), (1, 3), (9000, 20)]
    parts, bd, offs = make_batch(sizes)
    rhs = 2 * O.unit_vector(4, int(offs[-1])) - 1
    # same lanes per row on both sides (the default picks from the mean row length)
    opt = bd.cg_options(tolerance=1e-10, relative=True, preconditioner=pre, lanes_per_row=8)
    x, res = bd.cg_batched(offs, rhs, opt)
    for k, p in enumerate(parts):
        # the batch's default of 256 workgroups per system (one per 512 rows, 256 to 1024, alone)
        xs, its, rs = p.cg(rhs[offs[k]:offs[k + 1]], tolerance=1e-10, relative=True,
                           preconditioner=pre, lanes_per_row=8, max_blocks_per_system=256)
        assert res[k].status == T.CG_CONVERGED
        assert res[k].iterations == its
        assert np.array_equal(x[offs[k]:offs[k + 1]], xs)
        assert res[k].res_norm == rs
        assert res[k].time_us > 0
    # systems that need more iterations finish later
    its = np.array([r.iterations for r in res])
    t = np.array([r.time_us for r in res])
    assert t[np.argmax(its)] >= t[np.argmin(its)]


def test_batched_inactive_and_mixed():
    sizes = [(3000, 10), (5000, 24), (2500, 8)]
    parts, bd, offs = make_batch(sizes, seed=7)
    rhs = 2 * O.unit_vector(6, int(offs[-1])) - 1
    opt = bd.cg_options(tolerance=1e-8, relative=True, precision=T.CG_MIXED, replace_every=50)
    x, res = bd.cg_batched(offs, rhs, opt, active=[1, 0, 1])
    assert res[1].status == T.CG_INACTIVE
    assert np.all(x[offs[1]:offs[2]] == 0.0)
    for k in (0, 2):
        assert res[k].status == T.CG_CONVERGED
        A = scipy_of(parts[k])
        b = rhs[offs[k]:offs[k + 1]]
        assert np.linalg.norm(b - A @ x[offs[k]:offs[k + 1]]) < 1e-8 * np.linalg.norm(b)


def test_breakdown_is_reported(sym):
    # a non-finite right-hand side makes r.r non-finite at the first test
    o, g = sym
    rhs = np.ones(o.geom["ln_row"])
    rhs[17] = np.nan
    with pytest.raises(T.TbslaHipError) as e:
        g.cg(rhs, tolerance=1e-12, relative=True, preconditioner=T.PRECOND_JACOBI,
             max_iterations=50)
    assert e.value.status == T.ERR_BREAKDOWN


def test_poisson_direct_solve():
    dims = (16, 16, 16)
    dk = T.lognormal_field(*dims, seed=2)
    g = T.Matrix(T.CSR).fill_poisson3d(*dims, kappa=dk)
    A = scipy_of(g).tocsc()
    n = A.shape[0]
    rhs = T.uniform_vector(n, 3, -1.0, 1.0).to_numpy(count=n)
    exact = spla.spsolve(A, rhs)
    for pre, prec in [(T.PRECOND_NONE, T.CG_FP64), (T.PRECOND_JACOBI, T.CG_FP64),
                      (T.PRECOND_JACOBI, T.CG_MIXED)]:
        x, it, rn = g.cg(rhs, tolerance=1e-10, relative=True, preconditioner=pre,
                         precision=prec, replace_every=50 if prec == T.CG_MIXED else 0)
        assert np.linalg.norm(rhs - A @ x) < 1e-10 * np.linalg.norm(rhs) * 1.0001
        assert rel_max(x, exact) < 1e-6, (pre, prec, rel_max(x, exact))


def test_mixed_sees_values_changed_in_place():
    """The mixed form's FP32 copy of A is kept between solves and remade when
    the values change: after normalize_rows halves every row (A/2), the
    mixed solution doubles; a second handle of the same size reusing the
    first one's CG state converts its own values."""
    import gc
    dims = (14, 14, 14)
    g = T.Matrix(T.CSR).fill_poisson3d(*dims, kappa=T.lognormal_field(*dims, seed=5))
    A = scipy_of(g)
    n = A.shape[0]
    rhs = T.uniform_vector(n, 7, -1.0, 1.0).to_numpy(count=n)
    opts = dict(tolerance=1e-10, relative=True, preconditioner=T.PRECOND_JACOBI,
                precision=T.CG_MIXED, replace_every=20)
    x1, _, _ = g.cg(rhs, **opts)
    assert np.linalg.norm(rhs - A @ x1) < 1e-10 * np.linalg.norm(rhs) * 1.0001
    g.normalize_rows(np.full(n, 2.0))
    x2, _, _ = g.cg(rhs, **opts)
    assert np.linalg.norm(rhs - (A / 2) @ x2) < 1e-10 * np.linalg.norm(rhs) * 1.0001
    assert rel_max(x2, 2 * x1) < 1e-7
    del g
    gc.collect()
    h = T.Matrix(T.CSR).fill_poisson3d(*dims, kappa=T.lognormal_field(*dims, seed=6))
    B = scipy_of(h)
    x3, _, _ = h.cg(rhs, **opts)
    assert np.linalg.norm(rhs - B @ x3) < 1e-10 * np.linalg.norm(rhs) * 1.0001


@pytest.mark.parametrize("prec", [T.CG_FP64, T.CG_MIXED])
def test_values_written_through_device_arrays(prec):
    """Values rewritten in place through tbsla_hip_device_arrays and announced
    with tbsla_hip_values_changed reach every derived copy: CG on the diagonal
    layout (its own copy of the values, and their FP32 copy in the mixed
    form) then solves the new matrix exactly as CG on a fresh handle imported
    with the new values in the same layout, bit for bit."""
    dims = (16, 14, 12)
    g = T.Matrix(T.CSR).fill_poisson3d(*dims, kappa=T.lognormal_field(*dims, seed=11))
    n = g.geometry().ln_row
    rhs = T.uniform_vector(n, 12, -1.0, 1.0).to_numpy(count=n)
    opts = dict(tolerance=1e-10, relative=True, preconditioner=T.PRECOND_JACOBI,
                precision=prec, replace_every=20 if prec == T.CG_MIXED else 0)
    g.set_csr_layout(T.CSR_DIAG, 65536)
    g.cg(rhs, **opts)  # FP32 copies of the old values made and cached
    rp, ci, va = g.arrays()
    # scale row i by s_i and column j by s_j: stays symmetric positive definite
    s = 1.0 + O.unit_vector(13, n)
    rows = np.repeat(np.arange(n), np.diff(rp))
    new = va * s[rows] * s[ci]
    _, _, dv = g.device_arrays()
    T.check(T.lib().tbsla_hip_memcpy_htod(dv, new.ctypes.data, new.nbytes, None))
    g.values_changed()
    geo = g.geometry()
    geom = {f: getattr(geo, f) for f in T.GEOM_FIELDS}
    geom["nnz"] = geo.nnz
    ref = T.Matrix.from_arrays(T.CSR, geom, rp, ci, new)
    ref.set_csr_layout(T.CSR_DIAG, 65536)
    xg, itg, rg = g.cg(rhs, **opts)
    xr, itr, rr = ref.cg(rhs, **opts)
    assert itg == itr and rg == rr and np.array_equal(xg, xr)
    A = scipy_of(ref)
    assert np.linalg.norm(rhs - A @ xg) < 1e-10 * np.linalg.norm(rhs) * 1.0001


@pytest.mark.parametrize("pre,prec", [(T.PRECOND_NONE, T.CG_FP64), (T.PRECOND_JACOBI, T.CG_FP64),
                                      (T.PRECOND_JACOBI, T.CG_MIXED),
                                      (T.PRECOND_FACTORED, T.CG_FP64),
                                      (T.PRECOND_FACTORED, T.CG_MIXED)])
def test_diagonal_layout_cg(pre, prec):
    """A 7-point variable-coefficient Poisson block in the diagonal layout
    (every entry on one of its 7 stored diagonals): CG reads the diagonals.
    Early iterates within a relative 1e-12 of the oracle's PCG (FP64), the
    converged solution's true residual below the tolerance and the same
    iteration count as the row layout within two; repeated solves identical;
    after normalize_rows (A/2) the layout is rebuilt and the solution
    doubles."""
    dims = (24, 20, 18)
    kap = T.lognormal_field(*dims, seed=8)
    rows = T.Matrix(T.CSR).fill_poisson3d(*dims, kappa=kap)
    g = T.Matrix(T.CSR).fill_poisson3d(*dims, kappa=kap)
    g.set_csr_layout(T.CSR_DIAG, 65536, 8)
    lay = g.get_csr_layout()
    assert lay["layout"] == T.CSR_DIAG and lay["remote"] == 0
    A = scipy_of(g)
    n = A.shape[0]
    rhs = T.uniform_vector(n, 4, -1.0, 1.0).to_numpy(count=n)
    extra = {}
    kind = {T.PRECOND_NONE: 0, T.PRECOND_JACOBI: 1, T.PRECOND_FACTORED: 2}[pre]
    Go = None
    if pre == T.PRECOND_FACTORED:
        rp, ci, va = g.arrays()
        Go = lower_factor(O.csr_from_arrays(n, n, rp, ci, va))
        extra = dict(factor=to_gpu(Go))
    mixed = prec == T.CG_MIXED
    opts = dict(relative=True, preconditioner=pre, precision=prec,
                replace_every=50 if mixed else 0, **extra)
    if not mixed:
        rp, ci, va = g.arrays()
        o = O.csr_from_arrays(n, n, rp, ci, va)
        for k in (1, 4):
            xo, _, _ = O.pcg(o, rhs, max_iterations=k, tolerance=1e-30, kind=kind, G=Go)
            xg, itg, _ = g.cg(rhs, max_iterations=k, tolerance=1e-30, **opts)
            assert itg == k and rel_max(xg, xo) < 1e-12, (k, rel_max(xg, xo))
    x1, it1, _ = g.cg(rhs, tolerance=1e-10, **opts)
    x2, it2, _ = g.cg(rhs, tolerance=1e-10, **opts)
    xr, itr, _ = rows.cg(rhs, tolerance=1e-10, **opts)
    assert np.array_equal(x1, x2) and it1 == it2
    assert abs(it1 - itr) <= 2, (it1, itr)
    assert np.linalg.norm(rhs - A @ x1) < 1e-10 * np.linalg.norm(rhs) * 1.0001
    assert rel_max(x1, xr) < 1e-7
    g.normalize_rows(np.full(n, 2.0))
    assert g.get_csr_layout()["layout"] == T.CSR_DIAG
    x3, _, _ = g.cg(rhs, tolerance=1e-10, **opts)
    assert np.linalg.norm(rhs - (A / 2) @ x3) < 1e-10 * np.linalg.norm(rhs) * 1.0001
    assert rel_max(x3, 2 * x1) < 1e-7


def test_diagonal_layout_mixed_copy_not_reused_across_handles():
    """The mixed form's FP32 copy of the diagonals belongs to one layout
    build: a second handle of the same size, whose diagonals may land at the
    addresses the destroyed first handle's held, converts its own values."""
    import gc
    dims = (20, 20, 20)
    opts = dict(tolerance=1e-10, relative=True, preconditioner=T.PRECOND_JACOBI,
                precision=T.CG_MIXED, replace_every=50)
    n = 20 ** 3
    rhs = T.uniform_vector(n, 9, -1.0, 1.0).to_numpy(count=n)
    for seed in (None, 11, 12):
        kap = None if seed is None else T.lognormal_field(*dims, seed=seed)
        g = (T.Matrix(T.CSR).fill_poisson3d(*dims) if kap is None
             else T.Matrix(T.CSR).fill_poisson3d(*dims, kappa=kap))
        g.set_csr_layout(T.CSR_DIAG, 65536, 8)
        A = scipy_of(g)
        x, it, _ = g.cg(rhs, **opts)
        assert it < 2000
        assert np.linalg.norm(rhs - A @ x) < 1e-10 * np.linalg.norm(rhs) * 1.0001
        del g
        gc.collect()


@pytest.mark.parametrize("precision,pre", [
    (T.CG_FP64, T.PRECOND_NONE), (T.CG_MIXED, T.PRECOND_NONE),
    (T.CG_FP64, T.PRECOND_JACOBI), (T.CG_MIXED, T.PRECOND_JACOBI),
])
def test_two_pass_k1_same_bits(sym, precision, pre):
    """K1 in two passes (p' written, then gathered alone) gives the iterates
    of the one-kernel form bit for bit, at every lane count."""
    o, g = sym
    n = o.geom["ln_row"]
    rhs = 2 * O.unit_vector(23, n) - 1
    rep = 50 if precision == T.CG_MIXED else 0
    for L in (1, 4, 16):
        out = [g.cg(rhs, tolerance=1e-10, relative=True, precision=precision,
                    preconditioner=pre, replace_every=rep, lanes_per_row=L, spmv_passes=k)
               for k in (1, 2)]
        (x1, it1, r1), (x2, it2, r2) = out
        assert it1 == it2 and r1 == r2
        assert np.array_equal(x1, x2)
    if precision != T.CG_FP64:
        return
    # early iterates of the two-pass form against the oracle's PCG
    for k in (1, 5):
        x, _, _ = g.cg(rhs, max_iterations=k, tolerance=1e-30, relative=True,
                       preconditioner=pre, spmv_passes=2)
        xo, _, _ = O.pcg(o, rhs, max_iterations=k, tolerance=1e-30,
                         kind=1 if pre == T.PRECOND_JACOBI else 0)
        assert rel_max(x, xo) < 1e-12


@pytest.mark.parametrize("precision", [T.CG_FP64, T.CG_MIXED])
def test_two_pass_k1_batched_same_bits(precision):
    """A batch with K1 in one and in two passes: the same bits, in FP64 and
    in mixed precision (with residual replacement)."""
    sizes = [(4096, 8), (20000, 16), (9000, 20)]
    parts, bd, offs = make_batch(sizes, seed=3)
    rhs = 2 * O.unit_vector(8, int(offs[-1])) - 1
    xs = []
    for k in (1, 2):
        opt = bd.cg_options(tolerance=1e-10, relative=True, lanes_per_row=8, spmv_passes=k,
                            precision=precision, replace_every=50 if precision else 0)
        x, res = bd.cg_batched(offs, rhs, opt)
        xs.append((x, [(r.iterations, r.res_norm, r.status) for r in res]))
    assert xs[0][1] == xs[1][1]
    assert np.array_equal(xs[0][0], xs[1][0])


def test_spmv_passes_rejected_out_of_range(sym):
    o, g = sym
    with pytest.raises(T.TbslaHipError):
        g.cg(np.ones(o.geom["ln_row"]), spmv_passes=3)


@pytest.mark.parametrize("precision", [T.CG_FP64, T.CG_MIXED])
def test_two_pass_k1_factored_same_bits(sym, precision):
    """M = G G^T with K1 in one and in two passes: the same bits (z comes from
    the preconditioner's buffer in both)."""
    o, g = sym
    n = o.geom["ln_row"]
    G = to_gpu(lower_factor(o))
    rhs = 2 * O.unit_vector(29, n) - 1
    rep = 50 if precision == T.CG_MIXED else 0
    out = [g.cg(rhs, tolerance=1e-10, relative=True, precision=precision, replace_every=rep,
                preconditioner=T.PRECOND_FACTORED, factor=G, spmv_passes=k) for k in (1, 2)]
    (x1, it1, r1), (x2, it2, r2) = out
    assert it1 == it2 and r1 == r2
    assert np.array_equal(x1, x2)
