"""The C-ABI library loads and exports every symbol include/*.h declares
(no compute: this runs on machines without a GPU)."""
import ctypes
import glob
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "tbsla_amd", "libtbsla_hip.so")


def declared_symbols():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names |= set(re.findall(r"\b(tbsla_hip_\w+)\s*\(", text))
    return sorted(names)


def test_header_declares_entry_points():
    names = declared_symbols()
    assert "tbsla_hip_fill_cqmat" in names and "tbsla_hip_spmv" in names
    assert len(names) >= 25


def test_library_exports_every_declared_symbol():
    assert os.path.exists(LIB), "libtbsla_hip.so not built (run __graft_entry__.build())"
    lib = ctypes.CDLL(LIB)
    missing = [n for n in declared_symbols() if not hasattr(lib, n)]
    assert not missing, missing
    out = subprocess.run(["nm", "-D", "--defined-only", LIB], capture_output=True, text=True).stdout
    exported = set(re.findall(r"\sT\s(tbsla_hip_\w+)", out))
    assert set(declared_symbols()) <= exported


def test_python_binding_covers_header():
    import tbsla_amd
    bound = {name for name, _, _ in tbsla_amd.SIGNATURES}
    assert set(declared_symbols()) <= bound, set(declared_symbols()) - bound


def test_status_strings_and_bad_arguments():
    import tbsla_amd as T
    L = T.lib()
    assert L.tbsla_hip_status_string(0) == b"success"
    assert L.tbsla_hip_create(None, T.CSR) != 0
    h = ctypes.c_void_p()
    assert L.tbsla_hip_create(ctypes.byref(h), 99) == 1
    assert b"format" in L.tbsla_hip_last_error()
    assert L.tbsla_hip_create(ctypes.byref(h), T.CSR) == 0
    assert L.tbsla_hip_set_csr_schedule(h, 3, 0) == 1
    assert L.tbsla_hip_set_csr_schedule(h, 16, 3) == 1
    assert L.tbsla_hip_set_csr_schedule(h, 16, 4) == 0
    # vectorised CSR kernels: only the compiled (lanes, rows, vec, nt) sets
    for bad in [(32, 1, 4, 0), (16, 4, 2, 0), (8, 2, 3, 0), (8, 2, 2, 2), (2, 1, 2, 1), (64, 1, 2, 0)]:
        assert L.tbsla_hip_set_csr_kernel(h, *bad) == 1, bad
    assert L.tbsla_hip_set_csr_kernel(h, 8, 2, 4, 0) == 0
    got = [ctypes.c_int() for _ in range(4)]
    assert L.tbsla_hip_get_csr_kernel(h, *[ctypes.byref(v) for v in got]) == 0
    assert [v.value for v in got] == [8, 2, 4, 0]
    assert L.tbsla_hip_set_csr_schedule(h, 16, 4) == 0
    assert L.tbsla_hip_get_csr_kernel(h, *[ctypes.byref(v) for v in got]) == 0
    assert [v.value for v in got] == [16, 4, 1, 1]
    assert L.tbsla_hip_fill_cqmat(h, 10, 10, -1, 0.1, 1, 0, 0, 1, 1, -1, None) == 1
    assert L.tbsla_hip_fill_cqmat(h, 10, 10, 3, 0.1, 1, 2, 0, 1, 1, -1, None) == 1
    assert L.tbsla_hip_destroy(h) == 0


def test_compute_without_gpu_fails_loudly():
    import tbsla_amd as T
    if T.device_count() > 0:
        pytest.skip("a GPU is present")
    m = T.Matrix(T.CSR)
    with pytest.raises(T.TbslaHipError):
        m.fill_cqmat(100, 100, 5, 0.1)
