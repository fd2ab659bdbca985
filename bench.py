"""Benchmark: FP64 CSR SpMV on the n = 2^22, c = 32, q = 0.1 cqmat (one GPU).

    python bench.py [--gpus 1] [--steps K] [--warmup W]

A step is one SpMV y = A x over the whole matrix with A and x resident in
HBM (values and x uniform in [0,1), structure seed 0).  The timed region is K
back-to-back launches between two HIP events on the launch stream.  `value`
is algorithmic bytes / time: 12 bytes per stored entry (FP64 value + int32
column), 4 (n+1) for the row pointers, 8 n each for x and y.

Printed: one JSON line with the roofline of the CSR kernel (HBM-bound,
against 8 TB/s) and the reference's own OpenMP CPU SpMV (oracle/_ref, the
reference library compiled from its sources) timed on this host.

The same line carries "cg": time to relative residual 1e-8 (vectors in HBM,
host clock around each solve, second solve of each kind; every matrix is
autotuned first, so a stencil's CG reads its diagonal layout):
  * 7-point Poisson 128^3 (k = 1) and the symmetrised cqmat (n = 2^21, c = 16,
    q = 0.05): plain FP64 CG against mixed precision (FP32 matrix, FP64 dots,
    true residual every 50), with the error of both against a sparse direct
    solve of the 32^3 Poisson;
  * held-out variable-coefficient Poisson (64^3 and 128^3, k log-normal,
    sigma 1, correlation 4 cells): plain FP64, Jacobi FP64 / mixed, and the
    learned preconditioner (committed weights) mixed / FP64.
"batched_cg": 512 symmetrised cqmat systems of uneven size in one batched,
graph-captured solve (tbsla_amd/batched.py), p50/p99 per-system time and the
autotuned schedule; "ppr_spmm": personalised PageRank with 64 teleports on
the n = 2^22 stochastic cqmat, GB/s per fused SpMM iteration.
"""
import argparse
import json
import os
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "SpMV GB/s (% HBM peak); CG time-to-1e-8, learned vs Jacobi vs plain"
HBM_PEAK_GBS = 8000.0
MALL_BYTES = 256 << 20  # MI355X Infinity Cache: a working set below it is not an HBM figure
N, C, Q = 4194304, 32, 0.1


CAL_BYTES = 256 << 20
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "csr_spmv_cqmat_n4194304_c32_traffic.txt")


def measured_traffic(kernel: str):
    """HBM bytes per launch from the committed rocprofv3 --pmc summary of this
    workload (FETCH_SIZE + WRITE_SIZE, scaled by the calibration copy): the
    `traffic_bytes_per_launch:` line that follows the `kernel:` line naming
    the kernel measured now; else None."""
    if not os.path.exists(TRAFFIC_FILE):
        return None
    current = None
    with open(TRAFFIC_FILE) as f:
        for line in f:
            if ":" not in line or line.startswith(" "):
                continue
            k, v = (t.strip() for t in line.split(":", 1))
            if k == "kernel":
                current = v
            elif k == "traffic_bytes_per_launch" and current == kernel:
                return int(float(v))
    return None


def step_done(what: str) -> None:
    """One progress line on stderr as each section ends (stdout keeps the JSON line)."""
    print(f"[bench] {what} done", file=sys.stderr, flush=True)


def csr_kernel_name(k: dict, lay: dict) -> str:
    """The kernels of one SpMV as rocprofv3 names them: the row kernel with
    its template arguments (k_csr_rows<vec, lanes, rows, nontemporal,
    split>), or the diagonal kernel, plus the remote kernels of the split
    and diagonal layouts (the diagonal layout's LDS-slice form: the slice
    products and the diagonal kernel that sums them per row)."""
    import tbsla_amd as T
    layout = lay["layout"]
    if layout == T.CSR_DIAG:
        if lay.get("lds_block_rows", 0) > 0:
            return "k_slice_products + k_dia_rows_remote"
        if lay["remote"] == 0:
            return "k_dia_rows"
        return "k_dia_rows + k_remote_products + k_remote_combine"
    split = layout == T.CSR_SPLIT
    name = "k_csr_rows<%d, %d, %d, %d, %s>" % (k["vec"], k["lanes"], k["rows"], k["nontemporal"],
                                               "true" if split else "false")
    return name + " + k_remote_products + k_remote_combine" if split else name


def algorithmic_bytes_csr(stored, n_rows, n_cols):
    return 12 * stored + 4 * (n_rows + 1) + 8 * n_cols + 8 * n_rows


def host_cpu():
    """(model name, logical CPUs this process may run on)."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return model, len(os.sched_getaffinity(0))


def physical_cores():
    """Distinct physical cores (package, core id) among this process's CPUs,
    or None when the topology is not readable."""
    cores = set()
    try:
        for cpu in os.sched_getaffinity(0):
            base = f"/sys/devices/system/cpu/cpu{cpu}/topology/"
            with open(base + "physical_package_id") as f:
                pkg = f.read().strip()
            with open(base + "core_id") as f:
                core = f.read().strip()
            cores.add((pkg, core))
    except OSError:
        return None
    return len(cores) or None


def cpu_threads(requested=None):
    """Threads for the CPU baseline: the CPUs of this process's affinity,
    capped by OMP_NUM_THREADS (the host's CPU share) when it is set."""
    _, avail = host_cpu()
    cap = requested or int(os.environ.get("OMP_NUM_THREADS", avail))
    return max(1, min(avail, cap))


def cpu_baseline(n, c, q, stored, iters, reps, threads):
    """The reference's CPU SpMV (MatrixCSR::spmv, OpenMP static rows) on this
    host: oracle/_ref/ref_bench, built from the reference sources; where that
    is not built, the oracle's restatement of the same loop (orc_csr_Ax_omp,
    "port").  Median over reps of the mean of iters calls."""
    import oracle
    model, avail = host_cpu()
    exe = oracle.REF_BENCH
    phys = physical_cores()
    host = (f"{model}, {threads} threads on {avail} logical CPUs in this process's affinity"
            + (f" ({phys} physical cores)" if phys else ""))
    if os.path.exists(exe):
        env = dict(os.environ, OMP_NUM_THREADS=str(threads), OMP_PROC_BIND="close",
                   OMP_PLACES="cores")
        out = subprocess.run([exe, "CSR", str(n), str(c), str(q), "0", "0", "0", "1", "1",
                              str(iters), str(reps)], capture_output=True, text=True, env=env,
                             timeout=600)
        if out.returncode == 0:
            rec = json.loads(out.stdout.strip().splitlines()[-1])
            t = rec["time_op"]
            gbs = algorithmic_bytes_csr(stored, n, n) / t / 1e9
            return {"value": round(gbs, 2), "unit": "GB/s", "cores": rec["omp_threads"],
                    "kind": "reference", "cpu": host,
                    "sample": f"reference MatrixCSR::spmv (OpenMP static) on the full n={n} "
                              f"c={c} q={q} matrix, median over {reps} repeats of the mean of "
                              f"{iters} calls ({t * 1e3:.2f} ms per SpMV)"}
    import numpy as np
    oracle.lib().orc_set_threads(threads)
    o = oracle.fill_cqmat(oracle.CSR, n, n, c, q, 0, value_seed=0)
    x = oracle.unit_vector(1, n)
    r = np.zeros(n)
    o.csr_Ax_omp(r, x)
    times = []
    for _ in range(reps):
        t0 = time.perf_counter()
        for _ in range(iters):
            o.csr_Ax_omp(r, x)
        times.append((time.perf_counter() - t0) / iters)
    t = float(np.median(times))
    return {"value": round(algorithmic_bytes_csr(stored, n, n) / t / 1e9, 2), "unit": "GB/s",
            "cores": threads, "kind": "port", "cpu": host,
            "sample": f"oracle orc_csr_Ax_omp (the reference's OpenMP static row loop, "
                      f"restated) on the full n={n} c={c} q={q} matrix, median over {reps} "
                      f"repeats of the mean of {iters} calls ({t * 1e3:.2f} ms per SpMV)"}


def cg_section(args):
    """CG time-to-1e-8 figures (see the module docstring)."""
    import numpy as np
    import tbsla_amd as T
    from tbsla_amd import learned as LP
    out = {}

    def timed(A, rhs, **kw):
        n = A.geometry().ln_row
        x = T.DeviceBuffer(n * 8)
        o = A.cg_options(20000, 1e-8, True, **kw)
        A.cg_device(rhs, x, o)  # work vectors, graphs
        t0 = time.perf_counter()
        it, rn = A.cg_device(rhs, x, o)
        ms = (time.perf_counter() - t0) * 1e3
        y = T.DeviceBuffer(n * 8)
        T.check(T.lib().tbsla_hip_spmv(A.h, y.ptr, x.ptr, 0, None))
        b = rhs.to_numpy(count=n)
        rel = float(np.linalg.norm(b - y.to_numpy(count=n)) / np.linalg.norm(b))
        return {"ms": round(ms, 3), "iterations": it, "true_rel_res": rel}, x

    plain = dict(precision=T.CG_FP64)
    mixed = dict(precision=T.CG_MIXED, replace_every=50)
    side = 128
    A = T.Matrix(T.CSR).fill_poisson3d(side, side, side)
    A.tune_layout()  # the autotuned layout (diagonal) is what the CG engine reads
    rhs = T.uniform_vector(side ** 3, 3, -1.0, 1.0)
    out["poisson128"] = {"n": side ** 3, "nnz": A.geometry().stored,
                         "layout": A.get_csr_layout()["layout"],
                         "plain_fp64": timed(A, rhs, **plain)[0],
                         "mixed": timed(A, rhs, **mixed)[0]}
    del A
    S = T.Matrix(T.CSR).fill_sym_cqmat(1 << 21, 16, 0.05, 0, value_seed=0)
    S.tune_layout()
    rhs = T.uniform_vector(1 << 21, 4, -1.0, 1.0)
    out["sym_cqmat_2p21"] = {"n": 1 << 21, "nnz": S.geometry().stored,
                             "plain_fp64": timed(S, rhs, **plain)[0],
                             "mixed": timed(S, rhs, **mixed)[0]}
    del S
    # error against a sparse direct solve (32^3)
    try:
        import scipy.sparse as sp
        import scipy.sparse.linalg as spla
        A = T.Matrix(T.CSR).fill_poisson3d(32, 32, 32)
        A.tune_layout()
        rhs = T.uniform_vector(32 ** 3, 5, -1.0, 1.0)
        rp, ci, va = A.arrays()
        exact = spla.spsolve(sp.csr_matrix((va, ci, rp)).tocsc(), rhs.to_numpy(count=32 ** 3))
        errs = {}
        for name, kw in (("plain_fp64", plain), ("mixed", mixed)):
            _, x = timed(A, rhs, **kw)
            xv = x.to_numpy(count=32 ** 3)
            errs[name] = float(np.max(np.abs(xv - exact)) / np.max(np.abs(exact)))
        out["poisson32_error_vs_direct"] = errs
    except ImportError:
        pass
    # learned preconditioner on held-out variable-coefficient problems
    net = T.PrecondNet()
    have = LP.load_weights(net)
    for side_v, count in ((64, args.cg_problems), (128, max(1, args.cg_problems // 4))):
        rows = LP.evaluate(net if have else None, side_v, range(100_000, 100_000 + count))
        out[f"varcoef{side_v}"] = LP.summarise(rows)
    out["learned_weights"] = os.path.relpath(LP.WEIGHTS, ROOT) if have else None
    return out


def ppr_section():
    """Personalised PageRank on the n = 2^22, c = 32 stochastic cqmat with 64
    one-hot teleports: one fused SpMM + update per iteration.  Algorithmic
    bytes per iteration: A (12 stored + 4 (n+1)) + X read + X' written
    (2 * 8 * n * 64)."""
    import numpy as np
    import tbsla_amd as T
    n = 1 << 22
    m = T.Matrix(T.CSR).fill_cqmat(n, n, 32, 0.1, 0, value_seed=0)
    s = T.DeviceBuffer(n * 8)
    T.check(T.lib().tbsla_hip_make_stochastic(m.h, s.ptr, None))
    g = m.geometry()
    src = np.random.default_rng(0).choice(n, 64, replace=False).astype(np.int32)
    X = T.DeviceBuffer(n * 64 * 8)
    m.personalized_page_rank(src, 0.85, 1e-30, 2, out=X)
    iters = 20
    t0 = time.perf_counter()
    _, it, _ = m.personalized_page_rank(src, 0.85, 1e-30, iters, out=X)
    ms = (time.perf_counter() - t0) * 1e3
    nbytes = 12 * g.stored + 4 * (n + 1) + 2 * 8 * n * 64
    per = ms / it
    return {"n": n, "k": 64, "iterations": it, "ms_per_iteration": round(per, 4),
            "bytes_per_iteration": nbytes, "GBps": round(nbytes / (per * 1e-3) / 1e9, 1),
            "note": "host clock around 20 iterations incl. per-iteration column sums"}


def formats_section(iters: int = 20):
    """The other cfg2 measurements: SpMV of the ELL, COO, SCOO blocks and of
    the 16384^2 DENSE slice (block (0,0) of a 256 x 256 grid), the CSR blocks
    (0,0) and (0,1) of a 4 x 4 grid, PageRank on the stochastic cfg2 CSR
    (beta 0.85, epsilon 1e-10), the Poisson 128^3 SpMV on the row layout and
    on the autotuner's choice, and a 1 GiB device copy.  GB/s over each
    format's algorithmic bytes (stored entries, x and y once); HIP events
    around `iters` back-to-back SpMVs."""
    import tbsla_amd as T
    out = {}
    dx = T.DeviceBuffer(N * 8)
    T.check(T.lib().tbsla_hip_fill_unit_vector(dx.ptr, N, 1, None))
    dy = T.DeviceBuffer(N * 8)

    def timed(m, nbytes):
        m.time_spmv(dy, dx, 2)
        us = m.time_spmv(dy, dx, iters) / iters * 1e3
        return {"us": round(us, 1), "GBps": round(nbytes / (us * 1e-6) / 1e9, 1),
                "frac": round(nbytes / (us * 1e-6) / 1e9 / HBM_PEAK_GBS, 4)}

    for name, fmt in (("ELL", T.ELL), ("COO", T.COO), ("SCOO", T.SCOO)):
        m = T.Matrix(fmt).fill_cqmat(N, N, C, Q, 0, value_seed=0)
        g = m.geometry()
        if fmt == T.ELL:
            nbytes = 12 * g.ln_row * g.max_col + 8 * g.ln_col + 8 * g.ln_row
        else:
            # row-sorted triplets are read through a row pointer derived once
            # from the row array (kept with the block): the CSR bytes of the
            # same entries, 12 per entry, not the format's 16
            rows = g.n_row if fmt == T.COO else g.ln_row
            nbytes = 12 * g.stored + 4 * (rows + 1) + 8 * g.ln_col + 8 * rows
        extra = {}
        if fmt != T.ELL:
            extra = {"bytes_format": 16 * g.stored + 8 * g.ln_col + 8 * g.ln_row}
        # ELL's column-major kernel or COO / SCOO's row kernel, or the split /
        # diagonal layout when one is at least 3 % faster
        m.autotune(dy, dx, 5)
        extra["layout"] = m.get_csr_layout()["layout"]
        out[name] = {"bytes": nbytes, **extra, **timed(m, nbytes)}
        del m
    m = T.Matrix(T.DENSE).fill_cqmat(N, N, C, Q, 0, 0, 0, 256, 256, value_seed=0)
    g = m.geometry()
    nbytes = 8 * g.ln_row * g.ln_col + 8 * g.ln_col + 8 * g.ln_row
    out["DENSE_16384_slice"] = {"bytes": nbytes, **timed(m, nbytes)}
    del m
    for pc in (0, 1):
        m = T.Matrix(T.CSR).fill_cqmat(N, N, C, Q, 0, 0, pc, 4, 4, value_seed=0)
        g = m.geometry()
        nbytes = algorithmic_bytes_csr(g.stored, g.ln_row, g.ln_col)
        m.autotune(dy, dx, 5)
        out[f"CSR_grid4x4_block0{pc}"] = {"stored": g.stored, "bytes": nbytes,
                                          "kernel": m.get_csr_kernel(),
                                          "layout": m.get_csr_layout()["layout"],
                                          "mall_resident": nbytes < MALL_BYTES,
                                          **timed(m, nbytes)}
        del m
    # PageRank on the column-stochastic cfg2 CSR
    m = T.Matrix(T.CSR).fill_cqmat(N, N, C, Q, 0, value_seed=0)
    ds = T.DeviceBuffer(N * 8)
    T.check(T.lib().tbsla_hip_make_stochastic(m.h, ds.ptr, None))
    m.autotune(dy, dx, 5)
    import ctypes
    it, err = ctypes.c_int(), ctypes.c_double()
    T.check(T.lib().tbsla_hip_page_rank(m.h, 0.85, 1e-10, 3, dy.ptr, ctypes.byref(it),
                                        ctypes.byref(err), None))
    t0 = time.perf_counter()
    T.check(T.lib().tbsla_hip_page_rank(m.h, 0.85, 1e-10, 10000, dy.ptr, ctypes.byref(it),
                                        ctypes.byref(err), None))
    ms = (time.perf_counter() - t0) * 1e3
    g = m.geometry()
    spmv_bytes = algorithmic_bytes_csr(g.stored, g.ln_row, g.ln_col)
    out["page_rank"] = {"iterations": it.value, "max_error": err.value, "ms": round(ms, 2),
                        "ms_per_iteration": round(ms / max(it.value, 1), 4),
                        "spmv_GBps_equivalent": round(spmv_bytes * it.value / (ms * 1e-3) / 1e9, 1),
                        "note": "host clock around the whole solve (iterations, vector passes, "
                                "final normalisation), vectors resident"}
    del m, ds
    # the 7-point Poisson 128^3 stencil: row layout against the autotuner's
    # choice (the diagonal layout stores its 7 diagonals without indices)
    side = 128
    n3 = side ** 3
    A = T.Matrix(T.CSR).fill_poisson3d(side, side, side)
    g = A.geometry()
    nbytes = algorithmic_bytes_csr(g.stored, g.ln_row, g.ln_col)
    x3, y3 = T.DeviceBuffer(n3 * 8), T.DeviceBuffer(n3 * 8)
    T.check(T.lib().tbsla_hip_fill_unit_vector(x3.ptr, n3, 2, None))
    A.autotune(y3, x3, 10)
    lay = A.get_csr_layout()

    dia_bytes = 7 * 8 * n3 + 16 * n3

    def timed3(own):
        A.time_spmv(y3, x3, 2)
        us = A.time_spmv(y3, x3, iters) / iters * 1e3
        return {"us": round(us, 1), "bytes": own, "GBps": round(own / (us * 1e-6) / 1e9, 1),
                "frac": round(own / (us * 1e-6) / 1e9 / HBM_PEAK_GBS, 4),
                "csr_equivalent_GBps": round(nbytes / (us * 1e-6) / 1e9, 1),
                "mall_resident": own < MALL_BYTES}
    tuned = timed3(dia_bytes if lay["layout"] == T.CSR_DIAG else nbytes)
    A.set_csr_layout(T.CSR_ROWS)
    out["poisson128_spmv"] = {"stored": g.stored, "bytes_csr": nbytes,
                              "bytes_diagonal_layout": dia_bytes,
                              "autotuned": {"layout": lay["layout"], **tuned},
                              "row_layout": timed3(nbytes),
                              "note": "frac over the bytes the timed layout moves (the diagonal "
                                      "layout: 7 slots of 8 bytes, x and y, per row); both fit "
                                      "in the 256 MB MALL, so neither is an HBM figure"}
    del A, x3, y3
    # a plain device copy for reference (read + write)
    B = 1 << 30
    a, b = T.DeviceBuffer(B), T.DeviceBuffer(B)
    T.check(T.lib().tbsla_hip_copy_kernel(b.ptr, a.ptr, B, None))
    T.check(T.lib().tbsla_hip_stream_synchronize(None))
    t0 = time.perf_counter()
    for _ in range(10):
        T.check(T.lib().tbsla_hip_copy_kernel(b.ptr, a.ptr, B, None))
    T.check(T.lib().tbsla_hip_stream_synchronize(None))
    us = (time.perf_counter() - t0) / 10 * 1e6
    out["copy_1GiB"] = {"us": round(us, 1), "GBps_moved": round(2 * B / (us * 1e-6) / 1e9, 1)}
    # the read-stream ceiling: one pass over 2 GiB (8x the MALL), 16-byte loads
    del b
    R = 2 << 30
    big = T.DeviceBuffer(R)
    sink = T.DeviceBuffer(8192 * 8)
    T.check(T.lib().tbsla_hip_fill_unit_vector(big.ptr, R // 8, 3, None))
    T.check(T.lib().tbsla_hip_read_kernel(big.ptr, R, sink.ptr, None))
    T.check(T.lib().tbsla_hip_stream_synchronize(None))
    t0 = time.perf_counter()
    for _ in range(10):
        T.check(T.lib().tbsla_hip_read_kernel(big.ptr, R, sink.ptr, None))
    T.check(T.lib().tbsla_hip_stream_synchronize(None))
    us = (time.perf_counter() - t0) / 10 * 1e6
    out["read_2GiB"] = {"us": round(us, 1), "GBps_read": round(R / (us * 1e-6) / 1e9, 1),
                        "frac": round(R / (us * 1e-6) / 1e9 / HBM_PEAK_GBS, 4),
                        "note": "attainable read stream: the ceiling for read-dominated kernels"}
    return out


def attention_section(n: int = 1 << 20, nnz_per_row: int = 10, cols_b: int = 512):
    """The sparse attention step (tbsla_attention_mpi.cpp:179-227) on one
    block: n x n CSR with nnz_per_row entries per row at random columns
    (cqmat, q = 1), values U[0,1); the row softmax fused into one kernel and
    as the four separate row operations, then C = A B with B n x cols_b.
    Softmax bytes: values read and written once plus the row pointers
    (16 nnz + 4 (n + 1)); SpMM bytes: A once, one B row per entry, C written
    (12 nnz + 4 (n + 1) + 8 nnz cols_b + 8 n cols_b).  Host clock around
    repeated calls, each of which ends with a stream synchronisation."""
    import tbsla_amd as T
    m = T.Matrix(T.CSR).fill_cqmat(n, n, nnz_per_row, 1.0, 0, value_seed=1)
    g = m.geometry()
    nnz = g.stored
    dmax, ds = T.DeviceBuffer(n * 8), T.DeviceBuffer(n * 8)
    L = T.lib()

    def clock(fn, reps):
        fn()
        T.check(L.tbsla_hip_stream_synchronize(None))
        t0 = time.perf_counter()
        for _ in range(reps):
            fn()
        T.check(L.tbsla_hip_stream_synchronize(None))
        return (time.perf_counter() - t0) / reps * 1e6

    fused_us = clock(lambda: T.check(L.tbsla_hip_softmax_rows(m.h, -1, dmax.ptr, ds.ptr, None)), 20)

    def separate():
        T.check(L.tbsla_hip_get_row_max_abs(m.h, dmax.ptr, None))
        T.check(L.tbsla_hip_apply_exponential(m.h, dmax.ptr, -1, None))
        T.check(L.tbsla_hip_get_row_sums(m.h, ds.ptr, None))
        T.check(L.tbsla_hip_normalize_rows(m.h, ds.ptr, None))
    separate_us = clock(separate, 20)
    soft_bytes = 16 * nnz + 4 * (n + 1)
    B, Cm = T.DeviceBuffer(n * cols_b * 8), T.DeviceBuffer(n * cols_b * 8)
    T.check(L.tbsla_hip_fill_unit_vector(B.ptr, n * cols_b, 2, None))
    spmm_us = clock(lambda: T.check(L.tbsla_hip_spmm(m.h, B.ptr, Cm.ptr, cols_b, None)), 3)
    spmm_bytes = 12 * nnz + 4 * (n + 1) + 8 * nnz * cols_b + 8 * n * cols_b
    return {"n": n, "nnz_per_row": nnz_per_row, "cols_B": cols_b, "nnz": nnz,
            "softmax_fused_us": round(fused_us, 1), "softmax_separate_us": round(separate_us, 1),
            "softmax_fused_GBps": round(soft_bytes / (fused_us * 1e-6) / 1e9, 1),
            "softmax_gflops": round(6.0 * nnz / (fused_us * 1e-6) / 1e9, 1),
            "spmm_us": round(spmm_us, 1),
            "spmm_GBps": round(spmm_bytes / (spmm_us * 1e-6) / 1e9, 1),
            "spmm_gflops": round(2.0 * nnz * cols_b / (spmm_us * 1e-6) / 1e9, 1)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=None,
                    help="CPU baseline threads (default: this process's CPUs, capped by "
                         "OMP_NUM_THREADS)")
    ap.add_argument("--no-cg", action="store_true", help="skip the CG time-to-solution section")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the default CSR schedule (row layout) instead of the autotuned one")
    ap.add_argument("--csr-kernel", default=None, metavar="VEC,LANES,ROWS,NT",
                    help="use this CSR row kernel (row layout) instead of the autotuned one")
    ap.add_argument("--csr-layout", default=None, metavar="split|diag,SLICES",
                    help="a fixed derived layout instead of the autotuner's choice (with "
                         "--csr-kernel, that row schedule)")
    ap.add_argument("--no-formats", action="store_true",
                    help="skip the other formats, grid blocks, PageRank, attention and copy "
                         "figures")
    ap.add_argument("--no-batched", action="store_true",
                    help="skip the batched CG and batched PageRank sections")
    ap.add_argument("--no-ppr", action="store_true",
                    help="skip the personalised-PageRank (SpMM) section")
    ap.add_argument("--batched-systems", type=int, default=512)
    ap.add_argument("--cg-problems", type=int, default=8,
                    help="held-out 64^3 problems in the CG section (128^3: a quarter)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (one block of the 2-D grid per GPU); "
              f"--gpus {args.gpus} is not supported", file=sys.stderr)
        sys.exit(2)

    import tbsla_amd as T
    m = T.Matrix(T.CSR)
    t0 = time.time()
    m.fill_cqmat(N, N, C, Q, 0, value_seed=0)
    t_fill = time.time() - t0
    g = m.geometry()
    dx = T.DeviceBuffer(N * 8)
    T.check(T.lib().tbsla_hip_fill_unit_vector(dx.ptr, N, 1, None))
    dy = T.DeviceBuffer(N * 8)
    # schedule and layout from the autotuner (untimed)
    t0 = time.time()
    if args.csr_kernel or args.csr_layout:
        if args.csr_kernel:
            vec, lanes, rows, nt = (int(t) for t in args.csr_kernel.split(","))
            m.set_csr_kernel(lanes, rows, vec, nt)
        if args.csr_layout:
            name, slices = args.csr_layout.split(",")
            kind = {"split": T.CSR_SPLIT, "diag": T.CSR_DIAG}[name]
            m.set_csr_layout(kind, 65536 if kind == T.CSR_DIAG else 64, int(slices))
    elif not args.no_autotune:
        m.autotune(dy, dx, 5)
    t_tune = time.time() - t0
    layout = m.get_csr_layout()
    if args.warmup:
        m.time_spmv(dy, dx, args.warmup)
    ms = m.time_spmv(dy, dx, args.steps)
    per = ms / args.steps
    nbytes = algorithmic_bytes_csr(g.stored, g.ln_row, g.ln_col)
    gbs = nbytes / (per * 1e-3) / 1e9
    roof = {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(gbs / HBM_PEAK_GBS, 4), "traffic": None,
            "kernel": csr_kernel_name(m.get_csr_kernel(), layout),
            "bytes_per_launch": nbytes,
            "avg_launch_us": round(per * 1e3, 2)}
    # calibration copy of known size: FETCH_SIZE / WRITE_SIZE of this dispatch
    # give the scale factor applied to the SpMV kernel's counters
    cal = T.DeviceBuffer(CAL_BYTES), T.DeviceBuffer(CAL_BYTES)
    T.check(T.lib().tbsla_hip_copy_kernel(cal[1].ptr, cal[0].ptr, CAL_BYTES, None))
    T.check(T.lib().tbsla_hip_stream_synchronize(None))
    del cal
    traffic = measured_traffic(roof["kernel"])
    if traffic is not None:
        roof["traffic"] = traffic
    step_done("csr spmv (%.1f GB/s)" % gbs)
    cpu = None if args.no_cpu else cpu_baseline(N, C, Q, g.stored, 100, 10,
                                                cpu_threads(args.cpu_threads))
    if cpu is not None:
        step_done("cpu baseline")
    sched = m.get_csr_kernel()
    del m, dx, dy
    cg = None if args.no_cg else cg_section(args)
    if cg is not None:
        step_done("cg")
    extra = {}
    if not args.no_formats:
        extra["cfg2_other"] = formats_section()
        step_done("formats")
        extra["attention"] = attention_section()
        step_done("attention")
    if not args.no_batched:
        from tbsla_amd import batched
        extra["batched_cg"] = batched.bench(args.batched_systems)
        step_done("batched cg")
        extra["batched_page_rank"] = batched.page_rank_bench(args.batched_systems)
        step_done("batched page rank")
    if not args.no_ppr:
        extra["ppr_spmm"] = ppr_section()
        step_done("personalised page rank")
    line = {"metric": METRIC, "value": round(gbs, 1), "unit": "GB/s", "n_gpus": 1,
            "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(per, 5),
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
            "data": "synthetic",
            "config": {"workload": "csr_spmv_cqmat", "format": "CSR", "n": N, "c": C, "q": Q,
                       "seed_mult": 0, "values": "U[0,1) seed 0", "x": "U[0,1) seed 1",
                       "stored_nnz": g.stored, "nnz_field": g.nnz, "csr_schedule": sched,
                       "csr_layout": layout, "fill_s": round(t_fill, 3),
                       "autotune_s": round(t_tune, 2)},
            "roofline": roof, "cpu_baseline": cpu, "cg": cg, **extra}
    print(json.dumps(line))


if __name__ == "__main__":
    main()
