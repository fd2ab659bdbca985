"""The reference's binary matrix images (tests/golden/write_*.bin, written by
the compiled reference's Matrix::write) against the oracle's fills of the
same blocks: every header field and array, parsed in plain Python
(tests/binary_image.py), equals the oracle's geometry and arrays in the
reference layout.  This pins the images the GPU tests of read/write compare
with (tests/test_gpu_io.py)."""
import numpy as np
import pytest

import oracle as O
from binary_image import cases, parse

CASES = list(cases())


@pytest.mark.parametrize("k", range(len(CASES)))
def test_reference_image_equals_oracle(k):
    entry, img, (nr, nc, c, q, s, pr, pc, NR, NC) = CASES[k]
    fmt = entry["format"]
    p = parse(fmt, img)
    assert p["size"] == len(img) == entry["bytes"]
    o = O.fill_cqmat(O.FORMATS[fmt], nr, nc, c, q, s, pr, pc, NR, NC)
    g = o.geom
    for f in O.GEOM_FIELDS:
        assert p[f] == g[f], f
    assert p["nnz"] == g["nnz"] == entry["nnz"]
    arrs = o.arrays()
    if fmt == "CSR":
        rp, ci, va = arrs
        st = int(rp[-1])
        assert st == entry["stored"]
        assert np.array_equal(p["rowptr"], rp)
        assert len(p["values"]) == len(p["colidx"]) == g["nnz"]
        assert np.array_equal(p["colidx"][:st], ci) and np.array_equal(p["values"][:st], va)
    elif fmt == "ELL":
        cols, vals = arrs
        assert p["max_col"] == g["max_col"]
        assert np.array_equal(p["columns"], cols.reshape(-1))
        assert np.array_equal(p["values"], vals.reshape(-1))
    elif fmt in ("COO", "SCOO"):
        r, cidx, v = arrs
        assert np.array_equal(p["row"], r) and np.array_equal(p["col"], cidx)
        assert np.array_equal(p["values"], v)
    else:
        assert np.array_equal(p["values"], arrs[0].reshape(-1))
