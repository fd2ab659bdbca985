"""The CPU-side code under AddressSanitizer and UBSan (no GPU): the oracle
(oracle/tbsla_oracle.c, oracle/problems.c, built with gcc) and the library's
host code -- binary images (tbsla_amd/csrc/io.hip), the Matrix Market
parser (tbsla_amd/csrc/mmio.hip) and the C++ class's read/write
(include/tbsla/hip/Matrix.hpp) -- built with sanitizers on the host side
only, over a host-memory stand-in for the C-ABI calls they make
(tests/cpp/mock_capi.cpp).  Drivers: tests/cpp/sanitize_oracle.c and
tests/cpp/sanitize_host.cpp; any report fails the run."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ENV = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0",
           UBSAN_OPTIONS="print_stacktrace=1")
SAN = ["-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-fno-omit-frame-pointer"]


def run(cmd, **kw):
    r = subprocess.run(cmd, capture_output=True, text=True, **kw)
    assert r.returncode == 0, (cmd[0], r.stdout[-3000:], r.stderr[-3000:])
    return r


@pytest.mark.skipif(shutil.which("gcc") is None, reason="gcc not found")
def test_oracle_sanitized(tmp_path):
    exe = str(tmp_path / "sanitize_oracle")
    run(["gcc", "-O1", "-g", "-std=c11", *SAN, "-fopenmp", "-ffp-contract=off",
         "-I" + os.path.join(ROOT, "oracle"), os.path.join(ROOT, "oracle", "tbsla_oracle.c"),
         os.path.join(ROOT, "oracle", "problems.c"),
         os.path.join(ROOT, "tests", "cpp", "sanitize_oracle.c"), "-lm", "-o", exe], timeout=300)
    r = run([exe, str(tmp_path)], env=ENV, timeout=600)
    assert "sanitize_oracle ok" in r.stdout


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc not found")
def test_host_code_sanitized(tmp_path):
    exe = str(tmp_path / "sanitize_host")
    host_san = []
    for f in ("-fsanitize=address", "-fsanitize=undefined", "-fno-sanitize-recover=all"):
        host_san += ["-Xarch_host", f]  # host code only; device code is not instrumented
    run([HIPCC, "-O1", "-g", "-std=c++17", "--offload-arch=gfx950", *host_san,
         "-fno-omit-frame-pointer", "-Wno-unused-value", "-I" + os.path.join(ROOT, "include"),
         "-I" + os.path.join(ROOT, "tbsla_amd", "csrc"),
         os.path.join(ROOT, "tests", "cpp", "sanitize_host.cpp"),
         os.path.join(ROOT, "tests", "cpp", "mock_capi.cpp"),
         os.path.join(ROOT, "tbsla_amd", "csrc", "io.hip"),
         os.path.join(ROOT, "tbsla_amd", "csrc", "mmio.hip"), "-o", exe], timeout=600)
    images = []
    for name in sorted(os.listdir(GOLDEN)):
        if name.startswith("write_") and name.endswith(".bin"):
            images += [name.split("_")[1].upper(), os.path.join(GOLDEN, name)]
    r = run([exe, str(tmp_path), *images], env=ENV, timeout=600)
    assert "sanitize_host ok" in r.stdout
