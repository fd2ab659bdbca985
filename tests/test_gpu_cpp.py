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


# keys the reference MPI apps print (tbsla_perf_page_rank_mpi.cpp:448-494,
# tbsla_perf_cg_mpi.cpp:134-176)
MPI_KEYS = ("test", "matrix_format", "matrix_dim", "g_row", "g_col", "nnz", "nnz_min", "nnz_max",
            "nb_iterations", "time_app_in", "time_op", "time_gen_mat", "gnnz", "processes", "lang",
            "matrix_type", "compiler", "compile_options", "vectorization")


def run_app(name, *args, timeout=180):
    exe = os.path.join(ROOT, "tbsla_amd", "bin", name)
    assert os.path.exists(exe), f"{exe} not built"
    return subprocess.run([exe, *map(str, args)], capture_output=True, text=True, timeout=timeout)


def test_page_rank_app():
    """The PageRank app: the reference's flags and keys, 10 runs (the total
    iteration count is 10x one run's), median time and the reference's
    GFLOP formula; against the library's PageRank of the same matrix."""
    import numpy as np
    import tbsla_amd as T
    r = run_app("tbsla_perf_page_rank_hip", "--format", "CSR", "--matrix", "cqmat",
                "--matrix_dim", 200000, "--C", 12, "--Q", 0.1, "--S", 0, "--epsilon", 1e-10)
    assert r.returncode == 0, r.stderr
    rec = json.loads(r.stdout.strip().splitlines()[-1])
    for key in MPI_KEYS + ("gflops", "cqmat_c", "cqmat_q", "cqmat_s"):
        assert key in rec, key
    assert rec["test"] == "page_rank" and rec["matrix_dim"] == "200000"
    g = T.Matrix(T.CSR).fill_cqmat(200000, 200000, 12, 0.1, 0)
    g.make_stochastic()
    _, it, _ = g.page_rank(0.85, 1e-10)
    assert int(rec["nb_iterations"]) == 10 * it
    n, nnz, t = 200000, int(rec["nnz"]), float(rec["time_op"])
    assert np.isclose(float(rec["gflops"]), ((2 * nnz + 2 * n) * it + 2 * n) / (t * 1e9), rtol=1e-3)
    # a grid of more than one block needs more than one process
    r = run_app("tbsla_perf_page_rank_hip", "--format", "CSR", "--matrix", "cqmat", "--GR", 2)
    assert r.returncode == 99


def test_cg_app(tmp_path):
    """The CG app: the reference's keys.  In the reference's own setting
    (absolute |r| < 1e-10, CG.cpp:34, with its 1e-10 clamp of the divisors,
    :35) the iteration stalls once r.r falls below the clamp, so it runs to
    --max-iterations; with --relative on an SPD Poisson problem it converges.
    A binary image written by the library is read back through --matrix
    <name>."""
    import numpy as np
    import tbsla_amd as T
    r = run_app("tbsla_perf_cg_hip", "--format", "CSR", "--matrix", "poisson", "--side", 16,
                "--max-iterations", 300)
    assert r.returncode == 0, r.stderr
    rec = json.loads(r.stdout.strip().splitlines()[-1])
    for key in MPI_KEYS:
        assert key in rec, key
    assert rec["test"] == "conjugate_gradient" and int(rec["nb_iterations"]) <= 300
    r = run_app("tbsla_perf_cg_hip", "--format", "CSR", "--matrix", "poisson", "--side", 32,
                "--tolerance", "1e-8", "--relative")
    assert r.returncode == 0, r.stderr
    rec = json.loads(r.stdout.strip().splitlines()[-1])
    assert rec["status"] == "ok" and int(rec["nb_iterations"]) > 10
    assert float(rec["res_norm"]) < 1e-8 * np.sqrt(32 ** 3)  # |b| = sqrt(n): b is all ones
    g = T.Matrix(T.CSR).fill_sym_cqmat(50000, 8, 0.1, 1)
    T.check(T.lib().tbsla_hip_write_file(g.h, str(tmp_path / "spd.CSR").encode(), None))
    r = run_app("tbsla_perf_cg_hip", "--format", "CSR", "--matrix", "spd", "--matrix_folder",
                tmp_path, "--precision", "mixed", "--tolerance", "1e-8", "--relative")
    assert r.returncode == 0, r.stderr
    rec = json.loads(r.stdout.strip().splitlines()[-1])
    assert rec["status"] == "ok" and rec["matrix_dim"] == "50000"
    assert run_app("tbsla_perf_cg_hip", "--format", "ELL", "--matrix", "cqmat").returncode == 1


def test_cpp_subclass_of_reference_matrix():
    """tbsla::hip::CppMatrix (include/tbsla/hip/CppMatrix.hpp) is a
    tbsla::cpp::Matrix: GPU and reference CPU blocks driven through the
    reference's base class -- its virtual fills, spmv, Ax, write/read and
    its own compiled a_axpx_ / AAxpAxpx / make_stochastic (Matrix.cpp:5-38)
    -- give equal results for all five formats.  The program is built by
    oracle/Makefile where the reference's sources are present."""
    exe = os.path.join(ROOT, "oracle", "_ref", "test_cpp_subclass")
    if not os.path.exists(exe):
        pytest.skip("oracle/_ref/test_cpp_subclass not built (needs the reference's sources)")
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, (r.stdout[-3000:], r.stderr[-2000:])
    assert "test_cpp_subclass ok" in r.stdout
