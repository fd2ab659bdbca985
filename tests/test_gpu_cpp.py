"""The C++ drop-in classes (include/tbsla/hip/Matrix.hpp) and the perf app
on the GPU: the reference's own cpp tests rewritten against tbsla::hip."""
import json
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_cpp_dropin_reference_tests():
    exe = os.path.join(ROOT, "tests", "cpp", "test_dropin")
    assert os.path.exists(exe), "tests/cpp/test_dropin not built (run __graft_entry__.build())"
    r = subprocess.run([exe], capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, r.stderr[-2000:]
    assert "finished without error" in r.stdout


@pytest.mark.parametrize("fmt", ["CSR", "ELL", "COO", "SCOO", "DENSE"])
def test_perf_app_json(fmt):
    exe = os.path.join(ROOT, "tbsla_amd", "bin", "tbsla_perf_hip")
    assert os.path.exists(exe)
    n = 4096 if fmt == "DENSE" else 200000
    r = subprocess.run([exe, "--format", fmt, "--matrix", "cqmat", "--NR", str(n), "--NC", str(n),
                        "--C", "12", "--Q", "0.1", "--S", "0", "--op", "spmv", "--iterations",
                        "20"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    rec = json.loads(r.stdout.strip().splitlines()[-1])
    for key in ("test", "matrix_format", "n_row", "n_col", "nnz", "time_op", "gflops", "lang"):
        assert key in rec
    assert rec["matrix_format"] == fmt and float(rec["gflops"]) > 0


def test_perf_app_drivers():
    exe = os.path.join(ROOT, "tbsla_amd", "bin", "tbsla_perf_hip")
    r = subprocess.run([exe, "--format", "CSR", "--matrix", "cqmat", "--NR", "100000", "--NC",
                        "100000", "--C", "10", "--Q", "0.1", "--S", "0", "--op", "page_rank",
                        "--epsilon", "1e-10"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    rec = json.loads(r.stdout.strip().splitlines()[-1])
    assert int(rec["nb_iterations"]) > 10
