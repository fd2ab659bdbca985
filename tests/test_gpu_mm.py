"""Matrix Market import on the GPU (tbsla_hip_read_mm): the COO block holds
exactly the oracle reader's triplets, SpMV through it and through a CSR
conversion matches the oracle, and the converter app (tbsla_mm_to_binCOO's
counterpart) round-trips the file."""
import os
import subprocess

import numpy as np
import pytest

import oracle as O
import tbsla_amd as T
from test_mm import entries, write_mm

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.parametrize("kind", ["general", "symmetric"])
def test_read_mm_matches_oracle(tmp_path, kind):
    nr, nv = 50000, 400000
    r, c, v = entries(nr, nr, nv, 5, lower=kind == "symmetric")
    path = str(tmp_path / f"{kind}.mtx")
    write_mm(path, nr, nr, r, c, v, kind=kind)
    print(f"read_mm {kind} n={nr} nv={nv}", flush=True)
    rc, o = O.coo_read_mm(path)
    assert rc == 0
    g = T.Matrix.read_mm(path)
    geo = g.geometry()
    assert (geo.n_row, geo.n_col, geo.nnz) == (nr, nr, o.geom["nnz"])
    for a, b in zip(g.arrays(), o.arrays()):
        assert np.array_equal(a, b)
    x = O.unit_vector(3, nr)
    y, yo = g.spmv(x), o.spmv(x)
    assert np.max(np.abs(y - yo)) <= 1e-12 * np.max(np.abs(yo))
    csr = g.to(T.CSR)
    assert np.max(np.abs(csr.spmv(x) - yo)) <= 1e-12 * np.max(np.abs(yo))


def test_read_mm_errors(tmp_path):
    bad = str(tmp_path / "bad.mtx")
    write_mm(bad, 3, 3, [1], [1], [1.0], header="%%MatrixMarket matrix coordinate pattern general")
    with pytest.raises(T.TbslaHipError):
        T.Matrix.read_mm(bad)
    with pytest.raises(T.TbslaHipError):
        T.Matrix.read_mm(str(tmp_path / "missing.mtx"))
    ok = str(tmp_path / "ok.mtx")
    write_mm(ok, 3, 3, [1], [1], [1.0])
    with pytest.raises(T.TbslaHipError):  # readMM fills COO blocks only
        T.check(T.lib().tbsla_hip_read_mm(T.Matrix(T.CSR).h, os.fsencode(ok), None))


def test_mm_to_bin_app(tmp_path):
    exe = os.path.join(ROOT, "tbsla_amd", "bin", "tbsla_mm_to_bin_hip")
    assert os.path.exists(exe)
    r, c, v = entries(300, 300, 2000, 7, lower=True)
    mtx, binf = str(tmp_path / "a.mtx"), str(tmp_path / "a.bin")
    write_mm(mtx, 300, 300, r, c, v, kind="symmetric")
    out = subprocess.run([exe, mtx, binf], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    rc, o = O.coo_read_mm(mtx)
    assert f"nnz {o.geom['nnz']}" in out.stdout
    assert os.path.getsize(binf) > 0
