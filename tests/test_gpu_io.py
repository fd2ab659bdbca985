"""Binary images through the C ABI (tbsla_hip_read / tbsla_hip_write and
their file forms) against the bytes the compiled reference's Matrix::write
produced (tests/golden/write_*.bin):

* reading a reference image gives the oracle's block (geometry, arrays in
  the reference layout) and its SpMV;
* writing that block, and writing the same block filled on the GPU, gives
  the reference's bytes, byte for byte -- except where the reference writes
  uninitialised memory (CSR entries past rowptr[ln_row] that fill_cqmat's
  nnz field still counts), which this library writes as zeros;
* the file forms honour a seek offset; a truncated image is refused.
"""
import numpy as np
import pytest

import oracle as O
import tbsla_amd as T
from binary_image import cases, masked, undefined_ranges

pytestmark = pytest.mark.gpu

CASES = list(cases())


@pytest.mark.parametrize("k", range(len(CASES)))
def test_read_reference_image(k):
    entry, img, (nr, nc, c, q, s, pr, pc, NR, NC) = CASES[k]
    fmt = O.FORMATS[entry["format"]]
    m = T.Matrix.read(fmt, img)
    o = O.fill_cqmat(fmt, nr, nc, c, q, s, pr, pc, NR, NC)
    geo = m.geometry()
    for f in O.GEOM_FIELDS:
        assert getattr(geo, f) == o.geom[f], f
    assert geo.nnz == o.geom["nnz"]
    for a, b in zip(m.arrays(), o.arrays()):
        assert np.array_equal(np.asarray(a).reshape(-1), np.asarray(b).reshape(-1))
    x = O.unit_vector(3, o.geom["n_col"] if fmt == O.COO else o.geom["ln_col"])
    y, yo = m.spmv(x), o.spmv(x)
    assert np.max(np.abs(y - yo)) <= 1e-12 * max(np.max(np.abs(yo)), 1.0)


@pytest.mark.parametrize("k", range(len(CASES)))
def test_write_matches_reference_bytes(k):
    entry, img, (nr, nc, c, q, s, pr, pc, NR, NC) = CASES[k]
    fmt = O.FORMATS[entry["format"]]
    undefined = undefined_ranges(entry, img)
    want = masked(img, undefined)
    back = T.Matrix.read(fmt, img).write()
    assert back == want
    filled = T.Matrix(fmt).fill_cqmat(nr, nc, c, q, s, pr, pc, NR, NC)
    assert filled.write() == want


def test_file_forms_and_truncation(tmp_path):
    entry, img, _ = CASES[0]
    fmt = O.FORMATS[entry["format"]]
    m = T.Matrix.read(fmt, img)
    path = tmp_path / "block.bin"
    T.check(T.lib().tbsla_hip_write_file(m.h, str(path).encode(), None))
    assert path.read_bytes() == m.write()
    shifted = tmp_path / "shifted.bin"
    shifted.write_bytes(b"\x07" * 13 + img)
    r = T.Matrix(fmt)
    T.check(T.lib().tbsla_hip_read_file(r.h, str(shifted).encode(), 13, None))
    assert r.write() == m.write()
    for cut in (3, 40, len(img) - 1):
        with pytest.raises(T.TbslaHipError) as e:
            T.Matrix.read(fmt, img[:cut])
        assert e.value.status == T.ERR_INVALID
