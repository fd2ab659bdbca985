"""Writes the golden fixtures under tests/golden/ from the compiled reference.

Run in a container where the reference sources were present when
``make -C oracle`` ran (it needs oracle/_ref/libtbsla_ref.so):

    python tests/golden/make_golden.py

Each ``cqmat_<fmt>_<case>.npz`` holds one fill_cqmat block as the reference
library produced it (geometry, nnz field, arrays in the reference layout), an
input vector x (iota, so every product is an exact integer) and the
reference's ``spmv(x)``.  These are data -- inputs and expected outputs --
and pin the oracle wherever oracle/_ref is not built.

``write_<fmt>_<case>.bin`` holds the bytes the reference's ``Matrix::write``
produced for that block (MatrixCSR.cpp:236-261 and the per-format twins),
listed with their cases in ``write_index.json``.  The CSR case 0 block has
an nnz field larger than its stored entries (fill_cqmat's count pass): the
reference writes its uninitialised array tail there, so those bytes carry no
expected value and the index records where they start.
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import oracle as O  # noqa: E402

CASES = [
    # (n_row, n_col, c, q, seed_mult, pr, pc, NR, NC)
    (30, 30, 6, 0.1, 1, 0, 0, 1, 1),
    (30, 30, 6, 1.0, 0, 0, 0, 1, 1),
    (20, 30, 8, 0.5, 3, 1, 0, 2, 2),
    (30, 20, 12, 0.3, 2, 1, 1, 2, 2),
    (10, 5, 12, 0.5, 1, 2, 1, 3, 2),
    (400, 400, 10, 0.1, 0, 0, 0, 1, 1),
    (1000, 1000, 16, 0.05, 0, 1, 2, 4, 4),
]
# (format, case) pairs whose Matrix::write images are kept
WRITE_CASES = [("CSR", 6), ("CSR", 0), ("ELL", 6), ("COO", 6), ("SCOO", 6), ("DENSE", 3)]
NAMES = {O.CSR: ("rowptr", "colidx", "values"), O.ELL: ("columns", "values"),
         O.COO: ("row", "col", "values"), O.SCOO: ("row", "col", "values"),
         O.DENSE: ("values",)}


def main():
    if not O.ref_available():
        raise SystemExit("oracle/_ref/libtbsla_ref.so is not built")
    for fmt_name, fmt in O.FORMATS.items():
        for i, (nr, nc, c, q, s, pr, pc, NR, NC) in enumerate(CASES):
            if fmt == O.DENSE and nr * nc > 200000:
                continue
            r = O.ref_fill_cqmat(fmt, nr, nc, c, q, s, pr, pc, NR, NC)
            g = r.geom
            # stored triplet count for COO/SCOO is the nnz field after a fill
            arrays = r.arrays()
            if fmt == O.DENSE and not r._ptrs()[0]:
                continue  # empty DENSE block: the reference leaves nnz uninitialised
            x = np.arange(g["n_col"], dtype=np.float64)[g["f_col"]:] if fmt != O.COO else \
                np.arange(g["n_col"], dtype=np.float64)
            length = g["n_row"] if fmt == O.COO else g["ln_row"]
            y = r.spmv(x, length)
            out = dict(fmt=fmt, n_row=nr, n_col=nc, c=c, q=q, seed_mult=s, pr=pr, pc=pc, NR=NR,
                       NC=NC, geom=np.array([g[f] for f in O.GEOM_FIELDS], np.int64),
                       nnz=g["nnz"], names=np.array(NAMES[fmt]), x=x, y=y)
            for name, arr in zip(NAMES[fmt], arrays):
                out[name] = arr
            path = os.path.join(HERE, f"cqmat_{fmt_name.lower()}_{i}.npz")
            np.savez_compressed(path, **out)
            print(path)
    index = []
    for fmt_name, case in WRITE_CASES:
        fmt = O.FORMATS[fmt_name]
        nr, nc, c, q, s, pr, pc, NR, NC = CASES[case]
        r = O.ref_fill_cqmat(fmt, nr, nc, c, q, s, pr, pc, NR, NC)
        img = r.write()
        name = f"write_{fmt_name.lower()}_{case}.bin"
        with open(os.path.join(HERE, name), "wb") as f:
            f.write(img)
        g = r.geom
        stored = int(r.arrays()[0][-1]) if fmt == O.CSR else None
        index.append(dict(file=name, format=fmt_name, case=case, nnz=int(g["nnz"]),
                          stored=stored, bytes=len(img)))
        print(name, len(img))
    with open(os.path.join(HERE, "write_index.json"), "w") as f:
        json.dump(index, f, indent=1)


if __name__ == "__main__":
    main()
