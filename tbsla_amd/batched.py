"""Batched CG over many independent SPD systems of uneven size.

Workload: `count` symmetrised cqmat systems (tbsla_hip_fill_sym_cqmat) with
n = floor(2^U(12, 18)) (log-uniform), c uniform in 8..32, q = 0.1, value seed
= system index; right-hand sides uniform in [-1, 1).  All systems are the
diagonal blocks of one CSR block (tbsla_hip_block_diag) and are solved by
one kernel sequence per iteration, captured in hipGraphs
(tbsla_hip_cg_batched): FP32 matrix values with FP32 SpMV accumulation, FP64
vectors and dot products, the true FP64 residual every 50 iterations and
before a system is accepted.

The schedule (lanes per matrix row in the iteration's SpMV) is picked by
``autotune``: each candidate runs a few iterations of the real batch and the
fastest is kept.  Per-system solve time is the device wall clock from the
first iteration to the iteration at which that system's test passed.

    python -m tbsla_amd.batched --count 512

``page_rank_bench`` runs the same instance sizes as column-stochastic cqmat
blocks (c uniform in 8..32, q = 0.1) through tbsla_hip_page_rank_batched:
damping 0.85, each instance until its L1 change is below 1e-10.
"""
from __future__ import annotations

import argparse
import json
import time

import numpy as np

import tbsla_amd as T


def instance_sizes(count: int, seed: int = 0):
    rng = np.random.default_rng(seed)
    n = np.floor(2.0 ** rng.uniform(12, 18, count)).astype(np.int64)
    c = rng.integers(8, 33, count)
    return n, c


def build(count: int, seed: int = 0, q: float = 0.1):
    """(block-diagonal matrix, row offsets, rhs on the device, sizes, row lengths)."""
    n, c = instance_sizes(count, seed)
    parts = [T.Matrix(T.CSR).fill_sym_cqmat(int(n[k]), int(c[k]), q, 1, value_seed=k)
             for k in range(count)]
    bd = T.Matrix.block_diag(parts)
    del parts
    offs = np.concatenate([[0], np.cumsum(n)]).astype(np.int32)
    rhs = T.uniform_vector(int(offs[-1]), 17, -1.0, 1.0)
    return bd, offs, rhs, n, c


def page_rank_bench(count: int = 512, seed: int = 0, beta: float = 0.85,
                    epsilon: float = 1e-10) -> dict:
    """Batched PageRank over `count` stochastic cqmat instances of uneven size;
    host clock around the call (which ends synchronised), after one warm-up
    call.  GB/s counts one SpMV of the whole block per iteration of the
    slowest instance."""
    n, c = instance_sizes(count, seed)
    t0 = time.perf_counter()
    parts = []
    for k in range(count):
        g = T.Matrix(T.CSR).fill_cqmat(int(n[k]), int(n[k]), int(c[k]), 0.1, 0, value_seed=k)
        g.make_stochastic()
        parts.append(g)
    bd = T.Matrix.block_diag(parts)
    del parts
    build_s = time.perf_counter() - t0
    offs = np.concatenate([[0], np.cumsum(n)]).astype(np.int32)
    g = bd.geometry()
    # the block's SpMV row kernel from the autotuner (untimed)
    dx, dy = T.DeviceBuffer(g.ln_row * 8), T.DeviceBuffer(g.ln_row * 8)
    T.check(T.lib().tbsla_hip_fill_unit_vector(dx.ptr, g.ln_row, 1, None))
    bd.autotune(dy, dx, 3)
    del dx, dy
    bd.page_rank_batched(offs, beta, epsilon, 10000)
    t0 = time.perf_counter()
    _, its, errs, tus = bd.page_rank_batched(offs, beta, epsilon, 10000, times=True)
    ms = (time.perf_counter() - t0) * 1e3
    spmv_bytes = 12 * g.stored + 4 * (g.ln_row + 1) + 16 * g.ln_row
    return {"instances": count, "rows": int(offs[-1]), "nnz": g.stored,
            "csr_kernel": bd.get_csr_kernel(), "csr_layout": bd.get_csr_layout()["layout"],
            "build_s": round(build_s, 2), "batch_ms": round(ms, 3),
            "instances_per_s": round(count / (ms * 1e-3), 1),
            "per_instance_us_p50": round(float(np.percentile(tus, 50)), 1),
            "per_instance_us_p99": round(float(np.percentile(tus, 99)), 1),
            "per_instance_us_max": round(float(tus.max()), 1),
            "iterations_p50": float(np.median(its)), "iterations_max": int(its.max()),
            "max_error": float(errs.max()),
            "ms_per_iteration": round(ms / max(int(its.max()), 1), 4),
            "spmv_GBps_equivalent": round(spmv_bytes * int(its.max()) / (ms * 1e-3) / 1e9, 1)}


def options(A: T.Matrix, lanes: int, max_iterations: int = 5000, tol: float = 1e-8):
    return A.cg_options(max_iterations, tol, True, T.CG_MIXED, T.PRECOND_NONE, 50,
                        lanes_per_row=lanes)


def run(A, offs, rhs, opt, x=None):
    k = len(offs) - 1
    res = (T.CGResult * k)()
    x = x if x is not None else T.DeviceBuffer(int(offs[-1]) * 8)
    o = np.ascontiguousarray(offs, np.int32)
    t0 = time.perf_counter()
    T.check(T.lib().tbsla_hip_cg_batched(A.h, k, o.ctypes.data_as(T._ip), None, rhs.ptr, x.ptr,
                                         T.C.byref(opt), res, None))
    return (time.perf_counter() - t0) * 1e3, list(res), x


def autotune(A, offs, rhs, candidates=(2, 4, 8, 16, 32), iterations: int = 20):
    """Time `iterations` iterations of the batch per candidate; the fastest wins."""
    times = {}
    x = T.DeviceBuffer(int(offs[-1]) * 8)
    for L in candidates:
        opt = options(A, L, max_iterations=iterations, tol=1e-30)
        run(A, offs, rhs, opt, x)  # graphs, work vectors
        times[L] = min(run(A, offs, rhs, opt, x)[0] for _ in range(2))
    best = min(times, key=times.get)
    return best, {str(k): round(v, 3) for k, v in times.items()}


def bench(count: int = 512, seed: int = 0, tol: float = 1e-8, log=None) -> dict:
    t0 = time.perf_counter()
    A, offs, rhs, n, c = build(count, seed)
    build_s = time.perf_counter() - t0
    g = A.geometry()
    best, tune = autotune(A, offs, rhs)
    opt = options(A, best, tol=tol)
    x = T.DeviceBuffer(int(offs[-1]) * 8)
    run(A, offs, rhs, opt, x)
    ms, res, x = run(A, offs, rhs, opt, x)
    t = np.array([r.time_us for r in res])
    it = np.array([r.iterations for r in res])
    conv = np.array([r.status == T.CG_CONVERGED for r in res])
    # true residual per system
    y = T.DeviceBuffer(int(offs[-1]) * 8)
    T.check(T.lib().tbsla_hip_spmv(A.h, y.ptr, x.ptr, 0, None))
    b, ax = rhs.to_numpy(count=int(offs[-1])), y.to_numpy(count=int(offs[-1]))
    rel = np.array([np.linalg.norm(b[offs[k]:offs[k + 1]] - ax[offs[k]:offs[k + 1]]) /
                    np.linalg.norm(b[offs[k]:offs[k + 1]]) for k in range(count)])
    out = {"systems": count, "rows": int(offs[-1]), "nnz": g.stored, "build_s": round(build_s, 2),
           "n_min": int(n.min()), "n_max": int(n.max()), "c_range": [int(c.min()), int(c.max())],
           "schedule_lanes_per_row": best, "autotune_ms_20_iterations": tune,
           "batch_ms": round(ms, 3), "systems_per_s": round(count / (ms * 1e-3), 1),
           "per_system_us_p50": round(float(np.percentile(t, 50)), 1),
           "per_system_us_p99": round(float(np.percentile(t, 99)), 1),
           "per_system_us_max": round(float(t.max()), 1),
           "iterations_p50": float(np.median(it)),
           "iterations_p99": float(np.percentile(it, 99)), "iterations_max": int(it.max()),
           "converged": int(conv.sum()), "max_true_rel_res": float(rel.max())}
    if log:
        log(json.dumps(out))
    return out


def main(argv=None):
    ap = argparse.ArgumentParser(prog="python -m tbsla_amd.batched")
    ap.add_argument("--count", type=int, default=512)
    ap.add_argument("--seed", type=int, default=0)
    args = ap.parse_args(argv)
    bench(args.count, args.seed, log=lambda s: print(s, flush=True))


if __name__ == "__main__":
    main()
