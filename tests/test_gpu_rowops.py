"""Row operations on the GPU against the oracle (Matrix.hpp:46-51).

* get_row_sums, normalize_rows, get_row_max_abs: bit-exact (sums in the
  reference's serial order, IEEE division);
* apply_exponential: the device exp/pow within a relative 1e-14 of the C
  library's (a few ulp), zeros untouched, structure identical.
Blocks carry signed values with exact zeros (imported from the oracle's
arrays) on several block grids, so the per-format indexing (local rows for
CSR/DENSE, global rows for ELL/SCOO) and CSR's zero-sum safeguard are all
exercised.
"""
import json
import os
import subprocess

import numpy as np
import pytest

import oracle as O
import tbsla_amd as T

pytestmark = pytest.mark.gpu

CASES = [(3000, 3000, 12, 0.2, (0, 0, 1, 1)), (5000, 4000, 20, 0.5, (1, 2, 3, 4)),
         (20000, 20000, 32, 0.1, (0, 0, 1, 1)), (777, 1500, 7, 0.3, (2, 0, 3, 2))]


def signed(n, seed):
    rng = np.random.default_rng(seed)
    v = rng.uniform(-2.0, 2.0, n)
    v[rng.random(n) < 0.125] = 0.0
    return v


def oracle_block(fmt, n, m, c, q, grid, seed):
    o = O.fill_cqmat(fmt, n, m, c, q, 1, *grid)
    cur = o.arrays()[-1].reshape(-1)
    o.set_values(cur * signed(cur.size, seed))
    return o


def to_gpu(o):
    g, arr = o.geom, o.arrays()
    if o.fmt == O.CSR:
        return T.Matrix.from_arrays(T.CSR, g, *arr)
    if o.fmt == O.ELL:
        return T.Matrix.from_arrays(T.ELL, g, arr[0].reshape(-1), None, arr[1].reshape(-1))
    if o.fmt in (O.COO, O.SCOO):
        return T.Matrix.from_arrays(o.fmt, g, *arr)
    return T.Matrix.from_arrays(T.DENSE, g, None, None, arr[0].reshape(-1))


@pytest.mark.parametrize("fmt", [T.CSR, T.ELL, T.SCOO, T.DENSE])
def test_row_sums_and_normalize_rows_bit_exact(fmt):
    for k, (n, m, c, q, grid) in enumerate(CASES):
        if fmt == T.DENSE and n * m > 2e7:
            continue
        print(f"row sums fmt={fmt} n={n} m={m} grid={grid}", flush=True)
        o = oracle_block(fmt, n, m, c, q, grid, 10 + k)
        gm = to_gpu(o)
        nr = o.geom["n_row"]
        s0 = np.random.default_rng(k).uniform(-1, 1, nr) if fmt == T.SCOO else np.zeros(nr)
        so = O.Matrix.get_row_sums(o, s0)
        sg = gm.get_row_sums(s0)
        assert np.array_equal(sg, so), (fmt, n, grid)
        s1 = so.copy()
        s1[::7] = 0.0  # CSR sets those rows to 0, the others divide by zero
        o.normalize_rows(s1)
        gm.normalize_rows(s1)
        for a, b in zip(gm.arrays(), o.arrays()):
            assert np.array_equal(a, b, equal_nan=True), (fmt, n, grid)


def test_row_ops_after_fill_at_scale():
    # 2^20 rows of the cqmat workload shape, values in [0, 1): row sums, then
    # normalised rows sum to 1 and the SpMV of a ones vector returns them
    n = 1 << 20
    o = O.fill_cqmat(O.CSR, n, n, 32, 0.1, 1, value_seed=5)
    g = T.Matrix(T.CSR).fill_cqmat(n, n, 32, 0.1, 1, value_seed=5)
    s = g.get_row_sums()
    assert np.array_equal(s, o.get_row_sums())
    g.normalize_rows(s)
    y = g.spmv(np.ones(n))
    assert np.max(np.abs(y[s != 0] - 1.0)) < 1e-13
    assert np.all(y[s == 0] == 0.0)


def test_max_abs_and_exponential():
    for k, (n, m, c, q, grid) in enumerate(CASES):
        print(f"exponential n={n} m={m} grid={grid}", flush=True)
        o = oracle_block(O.CSR, n, m, c, q, grid, 50 + k)
        gm = to_gpu(o)
        mo, mg = o.get_row_max_abs(), gm.get_row_max_abs()
        assert np.array_equal(mg, mo)
        v = o.arrays()[-1].copy()
        for base, shift in ((-1, True), (0, False), (2, True), (3, False)):
            o.set_values(v)
            gm = to_gpu(o)
            o.apply_exponential(mo if shift else None, base)
            gm.apply_exponential(mg if shift else None, base)
            a, b = gm.arrays(), o.arrays()
            assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
            nz = v != 0
            assert np.array_equal(a[2][~nz], v[~nz])
            rel = np.abs(a[2][nz] - b[2][nz]) / np.abs(b[2][nz])
            assert rel.max() < 1e-14, (base, shift, rel.max())


@pytest.mark.parametrize("fmt", [T.ELL, T.SCOO, T.DENSE, T.COO])
def test_max_abs_and_exponential_noop_formats(fmt):
    gm = T.Matrix(fmt).fill_cqmat(500, 500, 8, 0.2, 1, value_seed=1)
    before = gm.arrays()
    assert np.array_equal(gm.get_row_max_abs(), np.zeros(500))
    gm.apply_exponential(np.ones(500), -1)
    gm.apply_exponential(None, 2)
    for a, b in zip(gm.arrays(), before):
        assert np.array_equal(a, b)


def test_coo_row_sums_noop():
    gm = T.Matrix(T.COO).fill_cqmat(500, 500, 8, 0.2, 1, value_seed=1)
    before = gm.arrays()
    assert np.array_equal(gm.get_row_sums(np.full(500, 3.0)), np.full(500, 3.0))
    gm.normalize_rows(np.zeros(500))
    for a, b in zip(gm.arrays(), before):
        assert np.array_equal(a, b)


def test_normalize_rows_refreshes_split_layout():
    # the split layout holds a copy of the values: SpMV after normalize_rows
    # must see the new values
    n = 50000
    o = O.fill_cqmat(O.CSR, n, n, 24, 0.3, 1, value_seed=3)
    g = T.Matrix(T.CSR).fill_cqmat(n, n, 24, 0.3, 1, value_seed=3)
    g.set_csr_layout(T.CSR_SPLIT, 64, 8)
    x = np.random.default_rng(0).uniform(-1, 1, n)
    g.spmv(x)
    s = g.get_row_sums()
    g.normalize_rows(s)
    o.normalize_rows(o.get_row_sums())
    assert g.get_csr_layout()["layout"] == T.CSR_SPLIT
    y = g.spmv(x)
    yo = o.spmv(x)
    assert np.max(np.abs(y - yo)) <= 1e-12 * np.max(np.abs(yo))
    g.apply_exponential(g.get_row_max_abs(), -1)
    o.apply_exponential(o.get_row_max_abs(), -1)
    y = g.spmv(x)
    yo = o.spmv(x)
    assert np.max(np.abs(y - yo)) <= 1e-12 * np.max(np.abs(yo))


# ---- the attention step's fused row softmax ---------------------------------

SOFTMAX_CASES = [
    (20000, 20000, 10, 1.0, (0, 0, 1, 1)),   # 4 lanes x 4 entries per row
    (20000, 20000, 24, 0.3, (0, 0, 1, 1)),   # 8 lanes x 4
    (20000, 20000, 32, 0.1, (0, 0, 1, 1)),   # 16 lanes x 4
    (6000, 6000, 64, 0.5, (0, 0, 1, 1)),     # 32 lanes x 4
    (3000, 3000, 150, 0.5, (0, 0, 1, 1)),    # 64 lanes x 2: two trips per row
    (9000, 7000, 20, 0.2, (1, 2, 3, 4)),     # an off-diagonal block of a grid
    (20000, 20000, 1, 0.5, (0, 0, 1, 1)),    # one lane per row
]


def separate_softmax(g, base):
    m = g.get_row_max_abs()
    g.apply_exponential(m, base)
    s = g.get_row_sums()
    g.normalize_rows(s)
    return m, s


@pytest.mark.parametrize("base", [-1, 2])
def test_fused_softmax_equals_the_four_row_ops(base):
    for k, (n, m, c, q, grid) in enumerate(SOFTMAX_CASES):
        print(f"softmax base={base} n={n} c={c} grid={grid}", flush=True)
        o = oracle_block(O.CSR, n, m, c, q, grid, 70 + k)
        ga, gb = to_gpu(o), to_gpu(o)
        ma, sa = separate_softmax(ga, base)
        mb, sb = gb.softmax_rows(base)
        assert np.array_equal(ma, mb) and np.array_equal(sa, sb)
        assert np.array_equal(ga.arrays()[2], gb.arrays()[2])
        # and the oracle's four calls, to the exponential's tolerance
        mo = o.get_row_max_abs()
        o.apply_exponential(mo, base)
        so = o.get_row_sums()
        o.normalize_rows(so)
        assert np.array_equal(mb, mo)
        vo, vb = o.arrays()[2], gb.arrays()[2]
        nz = vo != 0
        assert np.array_equal(vb[~nz], vo[~nz])
        assert np.max(np.abs(vb[nz] - vo[nz]) / np.abs(vo[nz])) < 1e-13
        assert np.array_equal(sb == 0, so == 0)  # rows whose values are all zero
        rows = so != 0
        assert np.max(np.abs(sb[rows] - so[rows]) / np.abs(so[rows])) < 1e-13


def test_fused_softmax_mixed_row_lengths():
    # rows of 0..200 entries (8 lanes x 2 per trip): long rows take several
    # trips, empty rows none
    rng = np.random.default_rng(3)
    n = 5000
    lens = np.where(rng.random(n) < 0.9, rng.integers(0, 12, n), rng.integers(0, 200, n))
    rp = np.zeros(n + 1, np.int32)
    rp[1:] = np.cumsum(lens)
    cols = rng.integers(0, n, int(rp[-1])).astype(np.int32)
    vals = signed(int(rp[-1]), 4)
    geom = dict(n_row=n, n_col=n, f_row=0, f_col=0, ln_row=n, ln_col=n, pr=0, pc=0, NR=1, NC=1,
                nnz=int(rp[-1]))
    ga = T.Matrix.from_arrays(T.CSR, geom, rp, cols, vals)
    gb = T.Matrix.from_arrays(T.CSR, geom, rp, cols, vals)
    ma, sa = separate_softmax(ga, -1)
    mb, sb = gb.softmax_rows(-1)
    assert np.array_equal(ma, mb) and np.array_equal(sa, sb)
    assert np.array_equal(ga.arrays()[2], gb.arrays()[2])


def test_attention_app():
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    exe = os.path.join(root, "tbsla_amd", "bin", "tbsla_attention_hip")
    assert os.path.exists(exe)
    outs = []
    for extra in ([], ["--separate"]):
        r = subprocess.run([exe, "--matrix_dim", "20000", "--NNZ", "10", "--cols_B", "512",
                            "--nb_multiplication", "2", *extra], capture_output=True, text=True,
                           timeout=120)
        assert r.returncode == 0, r.stderr
        rec = json.loads(r.stdout.strip().splitlines()[-1])
        for key in ("parameters", "timings", "gflops"):
            assert key in rec
        assert rec["parameters"]["cols_B"] == 512 and rec["parameters"]["world_size"] == 1
        assert rec["check"] < 1e-13, rec["check"]
        assert rec["gflops"]["softmax_operations"] > 0 and rec["gflops"]["multiplication"] > 0
        outs.append(rec)
    assert outs[0]["fused"] is True and outs[1]["fused"] is False
