"""Training and evaluation of the learned factored preconditioner.

Problems: -div(k grad u) = f on an s^3 cell grid (7-point, Dirichlet), with
k = exp(z), z a unit-variance Gaussian field smoothed over 4 cells
(tbsla_hip_fill_lognormal_field), and f uniform in [-1, 1).  Problem `seed`
uses field seed `seed` and right-hand side seed `seed + 1_000_000`.

Training (``train``) runs the network's own GPU forward/backward kernels and
Adam (tbsla_hip_precond_net_train_step), one problem per step.  Evaluation
(``solve``) times CG to a relative residual of 1e-8 with vectors resident in
HBM, for:

  plain_fp64     no preconditioner, FP64
  jacobi_fp64    D^-1, FP64
  jacobi_mixed   D^-1, FP32 matrix / FP64 dots, true residual every 50
  learned_mixed  M = G G^T from the network (inference time included),
                 FP32 matrix and factor / FP64 dots, true residual every 50;
                 Jacobi-PCG takes over when it breaks down
  learned_fp64   the same in FP64

    python -m tbsla_amd.learned train --side 64 --problems 2000 --out W.f32
    python -m tbsla_amd.learned eval  --side 64 --first 100000 --count 20 --weights W.f32
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

import tbsla_amd as T

WEIGHTS = os.path.join(T.PKG, "weights", "precond_net_poisson64.f32")
RHS_SEED_OFFSET = 1_000_000
METHODS = ("plain_fp64", "jacobi_fp64", "jacobi_mixed", "learned_mixed", "learned_fp64")


def poisson_problem(side: int, seed: int, sigma: float = 1.0, corr_len: float = 4.0):
    """(A, rhs on the device) for problem `seed`."""
    dk = T.lognormal_field(side, side, side, 0.0, sigma, corr_len, seed)
    A = T.Matrix(T.CSR).fill_poisson3d(side, side, side, kappa=dk)
    rhs = T.uniform_vector(side ** 3, seed + RHS_SEED_OFFSET, -1.0, 1.0)
    return A, rhs


def load_weights(net: T.PrecondNet, path: str = WEIGHTS) -> bool:
    if not os.path.exists(path):
        return False
    theta = np.fromfile(path, dtype="<f4")
    if theta.size != T.PrecondNet.param_count():
        raise ValueError(f"{path}: {theta.size} parameters, expected {T.PrecondNet.param_count()}")
    net.set_params(theta)
    return True


def save_weights(net: T.PrecondNet, path: str) -> None:
    os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
    net.get_params().astype("<f4").tofile(path)


def train(net: T.PrecondNet, seeds, side: int = 64, probes: int = 8, lr: float = 1e-3,
          clip: float = 1.0, epochs: int = 1, lr_final: float = 1e-4, log=None):
    """One Adam step per problem and epoch (cosine learning-rate decay from lr
    to lr_final); returns the losses."""
    seeds = list(seeds)
    total = len(seeds) * epochs
    losses = []
    step = 0
    t0 = time.time()
    for ep in range(epochs):
        for sd in seeds:
            A, _ = poisson_problem(side, sd)
            frac = step / max(1, total - 1)
            cur = lr_final + 0.5 * (lr - lr_final) * (1 + np.cos(np.pi * frac))
            losses.append(net.train_step(A, probes, 7_000_000 + step, float(cur), clip))
            step += 1
            if log and (step % 100 == 0 or step == total):
                log(f"step {step}/{total} loss {np.mean(losses[-100:]):.5f} "
                    f"lr {cur:.2e} {time.time() - t0:.1f}s")
    return losses


class Solver:
    """CG solves on one matrix with device vectors; times with the host
    clock around the library calls (which end synchronised)."""

    def __init__(self, A: T.Matrix, rhs: T.DeviceBuffer, net: T.PrecondNet | None = None,
                 factors: "tuple[T.Matrix, T.Matrix] | None" = None):
        """factors: the (G, G^T) handles the network writes into; passing the
        same pair for every problem of one size reuses their arrays."""
        self.A, self.rhs, self.net = A, rhs, net
        self.n = A.geometry().ln_row
        self.x = T.DeviceBuffer(self.n * 8)
        self.G, self.GT = factors if factors is not None else (T.Matrix(T.CSR), T.Matrix(T.CSR))

    def run(self, method: str, tol: float = 1e-8, max_iterations: int = 20000) -> dict:
        if method.startswith("learned"):
            return self.run_factored(method, None, None, tol, max_iterations)
        mixed = method.endswith("mixed")
        prec = T.CG_MIXED if mixed else T.CG_FP64
        pre = T.PRECOND_JACOBI if method.startswith("jacobi") else T.PRECOND_NONE
        t0 = time.perf_counter()
        o = self.A.cg_options(max_iterations, tol, True, prec, pre, 50 if mixed else 0)
        it, _ = self.A.cg_device(self.rhs, self.x, o)
        return self._result(method, t0, it, False)

    def run_factored(self, method: str, G: "T.Matrix | None", GT: "T.Matrix | None",
                     tol: float = 1e-8, max_iterations: int = 20000) -> dict:
        """M = G G^T PCG; G = None: the network's factor of A (inference
        timed with the solve).  Jacobi-PCG takes over when the factored
        iteration breaks down (r.z <= 0 or not finite, or no progress in 100
        iterations); the true FP64 residual judges the result."""
        A = self.A
        mixed = method.endswith("mixed")
        prec = T.CG_MIXED if mixed else T.CG_FP64
        rep = 50 if mixed else 0
        fallback = False
        t0 = time.perf_counter()
        if G is None:
            # G only: on the stencil's pattern the CG engine applies G and G^T
            # from G's diagonals (a factor off them gets G^T built per call)
            T.check(T.lib().tbsla_hip_precond_net_apply(self.net.h, A.h, self.G.h, None, None))
            G, GT = self.G, None
        o = A.cg_options(max_iterations, tol, True, prec, T.PRECOND_FACTORED, rep,
                         factor=G, factor_t=GT, stall_window=100)
        try:
            it, _ = A.cg_device(self.rhs, self.x, o)
        except T.TbslaHipError as e:
            if e.status != T.ERR_BREAKDOWN:
                raise
            fallback = True
            o = A.cg_options(max_iterations, tol, True, prec, T.PRECOND_JACOBI, rep)
            it, _ = A.cg_device(self.rhs, self.x, o)
        return self._result(method, t0, it, fallback)

    def _result(self, method: str, t0: float, it: int, fallback: bool) -> dict:
        ms = (time.perf_counter() - t0) * 1e3
        # accepted on the true FP64 residual
        y = T.DeviceBuffer(self.n * 8)
        T.check(T.lib().tbsla_hip_spmv(self.A.h, y.ptr, self.x.ptr, 0, None))
        b = self.rhs.to_numpy(count=self.n)
        r = b - y.to_numpy(count=self.n)
        return {"method": method, "ms": ms, "iterations": it,
                "true_rel_res": float(np.linalg.norm(r) / np.linalg.norm(b)),
                "fallback": fallback}


def evaluate(net, side: int, seeds, methods=METHODS, tol: float = 1e-8, warm: bool = True,
             log=None) -> list[dict]:
    rows = []
    factors = (T.Matrix(T.CSR), T.Matrix(T.CSR))  # one pair for all problems of this size
    for k, sd in enumerate(seeds):
        A, rhs = poisson_problem(side, sd)
        # the 7-point matrix in the diagonal layout: the CG engine then reads
        # its 7 stored diagonals (no column indices).  The SpMV autotuner's
        # 3 % rule can leave a 64^3 matrix on its rows, where K1 costs 11.7
        # instead of 9.5 us.
        A.set_csr_layout(T.CSR_DIAG, 65536, 8)
        s = Solver(A, rhs, net, factors)
        for m in methods:
            if net is None and m.startswith("learned"):
                continue
            if warm and k == 0:
                s.run(m, tol)  # first call allocates work vectors and captures graphs
            r = s.run(m, tol)
            r["seed"] = sd
            rows.append(r)
            if log:
                log(json.dumps(r))
    return rows


def summarise(rows: list[dict]) -> dict:
    out = {}
    for m in METHODS:
        rs = [r for r in rows if r["method"] == m]
        if not rs:
            continue
        ms = np.array([r["ms"] for r in rs])
        out[m] = {"ms_median": round(float(np.median(ms)), 3), "ms_mean": round(float(ms.mean()), 3),
                  "iterations_median": float(np.median([r["iterations"] for r in rs])),
                  "max_true_rel_res": float(max(r["true_rel_res"] for r in rs)),
                  "fallbacks": int(sum(r["fallback"] for r in rs)), "problems": len(rs)}
    return out


def main(argv=None):
    ap = argparse.ArgumentParser(prog="python -m tbsla_amd.learned")
    sub = ap.add_subparsers(dest="cmd", required=True)
    tr = sub.add_parser("train")
    tr.add_argument("--side", type=int, default=64)
    tr.add_argument("--problems", type=int, default=2000)
    tr.add_argument("--epochs", type=int, default=1)
    tr.add_argument("--probes", type=int, default=8)
    tr.add_argument("--lr", type=float, default=1e-3)
    tr.add_argument("--seed", type=int, default=0)
    tr.add_argument("--out", default=WEIGHTS)
    ev = sub.add_parser("eval")
    ev.add_argument("--side", type=int, default=64)
    ev.add_argument("--first", type=int, default=100_000)
    ev.add_argument("--count", type=int, default=20)
    ev.add_argument("--weights", default=WEIGHTS)
    ev.add_argument("--methods", default=",".join(METHODS),
                    help="comma-separated subset of " + ",".join(METHODS))
    args = ap.parse_args(argv)

    def log(s):
        print(s, flush=True)
    if args.cmd == "train":
        net = T.PrecondNet(seed=args.seed)
        losses = train(net, range(args.problems), args.side, args.probes, args.lr,
                       epochs=args.epochs, log=log)
        save_weights(net, args.out)
        log(json.dumps({"first100": float(np.mean(losses[:100])),
                        "last100": float(np.mean(losses[-100:])), "out": args.out}))
    else:
        net = T.PrecondNet()
        if not load_weights(net, args.weights):
            sys.exit(f"no weights at {args.weights}")
        methods = tuple(m for m in args.methods.split(",") if m)
        unknown = [m for m in methods if m not in METHODS]
        if unknown:
            sys.exit(f"unknown methods {unknown}")
        rows = evaluate(net, args.side, range(args.first, args.first + args.count), methods,
                        log=log)
        log(json.dumps(summarise(rows)))


if __name__ == "__main__":
    main()
