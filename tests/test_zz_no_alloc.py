"""tbsla_hip_Ax_noalloc (the allocation-free Ax) against tbsla_hip_Ax through
the C ABI: bit for bit, on two calls in a row whose second block is smaller
(the per-GPU scratch keeps the first call's contents), for row-sorted COO
and SCOO triplets (the path that uses the scratch), unsorted COO triplets
(atomics: compared within rounding) and a CSR block (no scratch)."""
import numpy as np
import pytest

import oracle as O
import tbsla_amd as T

pytestmark = pytest.mark.gpu


def ax_both(m, y0, x):
    L = T.lib()
    dx = T.DeviceBuffer.from_numpy(x)
    a, b = T.DeviceBuffer.from_numpy(y0), T.DeviceBuffer.from_numpy(y0)
    T.check(L.tbsla_hip_Ax(m.h, a.ptr, dx.ptr, 0, None))
    T.check(L.tbsla_hip_Ax_noalloc(m.h, b.ptr, dx.ptr, 0, None))
    n = len(y0)
    return a.to_numpy(count=n), b.to_numpy(count=n)


def block(fmt, n, c, q, seed, grid=(0, 0, 1, 1)):
    m = T.Matrix(fmt).fill_cqmat(n, n, c, q, seed, *grid, value_seed=seed + 3)
    g = m.geometry()
    nin = g.n_col if fmt == T.COO else g.ln_col
    nout = g.n_row if fmt == T.COO else g.ln_row
    return m, O.unit_vector(seed + 1, nin), 2 * O.unit_vector(seed + 2, nout) - 1


@pytest.mark.parametrize("fmt", [T.COO, T.SCOO, T.CSR])
def test_ax_noalloc_two_calls(fmt):
    # 1.2 M triplets (2,344 chunks of scratch), then 9,000 (18 chunks)
    for n, c, q, seed, grid in ((100_000, 12, 0.1, 1, (0, 0, 1, 1)),
                                (3_000, 3, 0.3, 5, (1, 0, 2, 2))):
        m, x, y0 = block(fmt, n, c, q, seed, grid)
        a, b = ax_both(m, y0, x)
        assert np.array_equal(a, b), (fmt, n)


def test_ax_noalloc_unsorted_triplets():
    # triplets imported out of row order: the atomic path (no scratch)
    o = O.fill_cqmat(O.COO, 20_000, 20_000, 8, 0.2, 2, value_seed=4)
    r, c, v = o.arrays()
    perm = np.random.default_rng(0).permutation(len(v))
    m = T.Matrix.from_arrays(T.COO, o.geom, r[perm], c[perm], v[perm], stored=len(v))
    x, y0 = O.unit_vector(7, 20_000), 2 * O.unit_vector(8, 20_000) - 1
    a, b = ax_both(m, y0, x)
    assert np.max(np.abs(a - b)) <= 1e-12 * np.max(np.abs(a))
