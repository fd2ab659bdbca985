"""A plain-Python parser of the binary matrix image that
tbsla::cpp::Matrix::write produces (MatrixCSR.cpp:236-261,
MatrixELL.cpp:208-228, MatrixCOO.cpp:144-169, MatrixSCOO.cpp:166-191,
MatrixDENSE.cpp:88-105), for the tests: header n_row n_col ln_row ln_col
f_row f_col (int32), nnz (int64), pr pc NR NC (int32), ELL max_col (int32),
then each array as a size_t count and its elements."""
import json
import os
import struct

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
HEADER = ("n_row", "n_col", "ln_row", "ln_col", "f_row", "f_col")
TAIL = ("pr", "pc", "NR", "NC")
ARRAYS = {"CSR": (("values", "<f8"), ("rowptr", "<i4"), ("colidx", "<i4")),
          "COO": (("values", "<f8"), ("row", "<i4"), ("col", "<i4")),
          "SCOO": (("values", "<f8"), ("row", "<i4"), ("col", "<i4")),
          "ELL": (("values", "<f8"), ("columns", "<i4")),
          "DENSE": (("values", "<f8"),)}


def parse(fmt: str, img: bytes) -> dict:
    """Fields and arrays of an image; 'offsets' maps each array to the byte
    offset of its first element."""
    out, at = {}, 0
    for f in HEADER:
        out[f] = struct.unpack_from("<i", img, at)[0]
        at += 4
    out["nnz"] = struct.unpack_from("<q", img, at)[0]
    at += 8
    for f in TAIL:
        out[f] = struct.unpack_from("<i", img, at)[0]
        at += 4
    if fmt == "ELL":
        out["max_col"] = struct.unpack_from("<i", img, at)[0]
        at += 4
    out["offsets"] = {}
    for name, dt in ARRAYS[fmt]:
        n = struct.unpack_from("<Q", img, at)[0]
        at += 8
        out["offsets"][name] = at
        out[name] = np.frombuffer(img, dtype=dt, count=n, offset=at).copy()
        at += n * np.dtype(dt).itemsize
    out["size"] = at
    return out


def cases():
    """(index entry, image bytes, fill_cqmat arguments) of the reference's
    write() images kept under tests/golden/."""
    with open(os.path.join(GOLDEN, "write_index.json")) as f:
        index = json.load(f)
    from golden.make_golden import CASES  # noqa: E402 (tests/ is on sys.path)
    for e in index:
        with open(os.path.join(GOLDEN, e["file"]), "rb") as f:
            img = f.read()
        yield e, img, CASES[e["case"]]


def undefined_ranges(entry: dict, img: bytes):
    """Byte ranges of the image that the reference leaves undefined: the CSR
    values and colidx entries past rowptr[ln_row] that the nnz field still
    counts (never initialised by MatrixCSR::fill_cqmat)."""
    if entry["format"] != "CSR" or entry["stored"] == entry["nnz"]:
        return []
    p = parse("CSR", img)
    s = entry["stored"]
    return [(p["offsets"]["values"] + 8 * s, p["offsets"]["values"] + 8 * entry["nnz"]),
            (p["offsets"]["colidx"] + 4 * s, p["offsets"]["colidx"] + 4 * entry["nnz"])]


def masked(img: bytes, ranges) -> bytes:
    b = bytearray(img)
    for lo, hi in ranges:
        b[lo:hi] = bytes(hi - lo)
    return bytes(b)
