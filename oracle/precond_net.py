"""CPU restatement of the learned factored preconditioner (test infrastructure).

The network maps an SPD CSR matrix A to a lower-triangular G on tril(A)'s
pattern, used as M = G G^T in CG.  It has no counterpart in the reference
(whose CG, src/mpi/matrix/CG.cpp:32-73, is unpreconditioned); the GPU
implementation is tbsla_amd/csrc/precond_net.hip.  Everything here is float64;
``bf16=True`` rounds exactly the tensors the GPU feeds to its BF16 MFMA
products (round to nearest even), so GPU results can be compared to ~1e-5.

Definition
----------
s_i = A_ii^-1/2; A^_ij = s_i A_ij s_j; edges e_ij = A^_ij (j != i).
Node features x_i = [1, sum_j |e_ij|, max_j |e_ij|, deg_i / 8].
Layer 0: h1 = relu([x, X~] W0 + b0),   X~_i = sum_j e_ij x_j          (8 -> 64)
Layer l: U = h_l W_l (64 x 128);  h_{l+1,i} = relu(U_i[:64] + sum_j e_ij U_j[64:] + b_l)
Head:    P = h3 V (64 x 128); for j < i on the pattern:
         s_ij = relu(P_i[:64] + P_j[64:] + e_ij ve + c),  g^_ij = wo . s_ij + bo
         g^_ii = exp(wd . h3_i + bd)
Output:  G_ij = s_i g^_ij  (so G^T A G = G^^T A^ G^).
Loss:    L = |Z - G^^T A^ G^ Z|_F^2 / (n k) over k Rademacher probes Z.

With wo = bo = wd = bd = 0 the network gives G^ = I, i.e. Jacobi.
"""
from __future__ import annotations

import numpy as np

H = 64
F = 4
SHAPES = [("W0", (2 * F, H)), ("b0", (H,)), ("W1", (H, 2 * H)), ("b1", (H,)),
          ("W2", (H, 2 * H)), ("b2", (H,)), ("V", (H, 2 * H)), ("ve", (H,)), ("c", (H,)),
          ("wo", (H,)), ("bo", ()), ("wd", (H,)), ("bd", ())]
N_PARAMS = sum(int(np.prod(s)) for _, s in SHAPES)


def unflatten(theta: np.ndarray) -> dict:
    p, o = {}, 0
    for name, shp in SHAPES:
        k = int(np.prod(shp))
        p[name] = np.asarray(theta[o:o + k], np.float64).reshape(shp)
        o += k
    return p


def flatten(p: dict) -> np.ndarray:
    return np.concatenate([np.asarray(p[name], np.float64).ravel() for name, _ in SHAPES])


def bf16(a: np.ndarray) -> np.ndarray:
    """Round float values to bfloat16 (nearest even) and back, via float32."""
    f = np.asarray(a, np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


class Graph:
    """The scaled matrix's structure: CSR of A (columns sorted, diagonal
    present), off-diagonal edges, and the lower pattern of G."""

    def __init__(self, rowptr, colidx, values):
        self.rp = np.asarray(rowptr, np.int64)
        self.ci = np.asarray(colidx, np.int64)
        v = np.asarray(values, np.float64)
        n = len(self.rp) - 1
        self.n = n
        rows = np.repeat(np.arange(n), np.diff(self.rp))
        diag = np.zeros(n)
        on = rows == self.ci
        np.add.at(diag, rows[on], v[on])
        self.s = 1.0 / np.sqrt(diag)
        off = ~on
        self.er, self.ec = rows[off], self.ci[off]
        self.ev = v[off] * (self.s[self.er] * self.s[self.ec])
        lo = self.ec < self.er
        self.lr, self.lc, self.le = self.er[lo], self.ec[lo], self.ev[lo]
        # A^ as a dense-free operator
        import scipy.sparse as sp
        self.Ahat = sp.csr_matrix((np.concatenate([self.ev, np.ones(n)]),
                                   (np.concatenate([self.er, np.arange(n)]),
                                    np.concatenate([self.ec, np.arange(n)]))), shape=(n, n))
        self.E = sp.csr_matrix((self.ev, (self.er, self.ec)), shape=(n, n))

    def features(self):
        n = self.n
        x = np.zeros((n, F))
        x[:, 0] = 1.0
        np.add.at(x[:, 1], self.er, np.abs(self.ev))
        mx = np.zeros(n)
        np.maximum.at(mx, self.er, np.abs(self.ev))
        x[:, 2] = mx
        deg = np.bincount(self.er, minlength=n).astype(np.float64)
        x[:, 3] = deg / 8.0
        return x


def rademacher(n: int, k: int, seed: int) -> np.ndarray:
    """Z[i, p] = -1 if V(seed, i*k + p) < 0.5 else +1 (row-major n x k)."""
    from oracle import unit_vector
    u = unit_vector(seed, n * k).reshape(n, k)
    return np.where(u < 0.5, -1.0, 1.0)


def forward(p: dict, g: Graph, bf: bool = False, keep: bool = False):
    """G^ as (diag, lower values in g.lr/g.lc order) and the activations."""
    r = (lambda a: bf16(a)) if bf else (lambda a: a)
    x = g.features()
    xa = g.E @ x
    z0 = np.concatenate([x, xa], axis=1)
    pre1 = z0 @ p["W0"] + p["b0"]
    h1 = np.maximum(pre1, 0.0)
    U1 = r(h1) @ r(p["W1"])
    pre2 = U1[:, :H] + g.E @ U1[:, H:] + p["b1"]
    h2 = np.maximum(pre2, 0.0)
    U2 = r(h2) @ r(p["W2"])
    pre3 = U2[:, :H] + g.E @ U2[:, H:] + p["b2"]
    h3 = np.maximum(pre3, 0.0)
    P = r(h3) @ r(p["V"])
    S = P[g.lr, :H] + P[g.lc, H:] + g.le[:, None] * p["ve"] + p["c"]
    Sr = np.maximum(S, 0.0)
    glow = Sr @ p["wo"] + p["bo"]
    gdiag = np.exp(h3 @ p["wd"] + p["bd"])
    acts = dict(x=x, z0=z0, h1=h1, U1=U1, h2=h2, U2=U2, h3=h3, P=P, S=S, Sr=Sr) if keep else None
    return gdiag, glow, acts


def ghat_matrix(g: Graph, gdiag, glow):
    import scipy.sparse as sp
    n = g.n
    return sp.csr_matrix((np.concatenate([gdiag, glow]),
                          (np.concatenate([np.arange(n), g.lr]),
                           np.concatenate([np.arange(n), g.lc]))), shape=(n, n))


def G_csr(g: Graph, gdiag, glow):
    """G = D^-1/2 G^ as a CSR (scipy) matrix, columns sorted."""
    import scipy.sparse as sp
    Gh = ghat_matrix(g, gdiag, glow)
    Gm = sp.diags(g.s) @ Gh
    Gm = sp.csr_matrix(Gm)
    Gm.sort_indices()
    return Gm


def loss_and_ghat_grad(g: Graph, gdiag, glow, Z):
    """L and dL/dg^ (diag, lower) for probes Z."""
    n, k = Z.shape
    Gh = ghat_matrix(g, gdiag, glow)
    U = Gh @ Z
    Vv = g.Ahat @ U
    W = Gh.T @ Vv
    R = Z - W
    L = float((R * R).sum() / (n * k))
    Y = g.Ahat @ (Gh @ R)
    c = -2.0 / (n * k)
    dgd = c * ((Vv * R).sum(axis=1) + (Y * Z).sum(axis=1))
    dgl = c * ((Vv[g.lr] * R[g.lc]).sum(axis=1) + (Y[g.lr] * Z[g.lc]).sum(axis=1))
    return L, dgd, dgl


def backward(p: dict, g: Graph, acts: dict, gdiag, dgd, dgl, bf: bool = False) -> dict:
    r = (lambda a: bf16(a)) if bf else (lambda a: a)
    d = {name: np.zeros(shp) for name, shp in SHAPES}
    h3, P, Sr = acts["h3"], acts["P"], acts["Sr"]
    # diagonal head
    dq = dgd * gdiag
    d["wd"] = h3.T @ dq
    d["bd"] = np.array(dq.sum())
    dh3 = np.outer(dq, p["wd"])
    # edge head
    d["wo"] = Sr.T @ dgl
    d["bo"] = np.array(dgl.sum())
    dS = np.outer(dgl, p["wo"]) * (acts["S"] > 0)
    d["ve"] = (dS * g.le[:, None]).sum(axis=0)
    d["c"] = dS.sum(axis=0)
    dP = np.zeros_like(P)
    np.add.at(dP[:, :H], g.lr, dS)
    np.add.at(dP[:, H:], g.lc, dS)
    d["V"] = r(h3).T @ r(dP)
    dh3 = dh3 + r(dP) @ r(p["V"]).T
    # layer 2 -> 3
    dpre3 = dh3 * (h3 > 0)
    d["b2"] = dpre3.sum(axis=0)
    dU2 = np.concatenate([dpre3, g.E.T @ dpre3], axis=1)
    d["W2"] = r(acts["h2"]).T @ r(dU2)
    dh2 = r(dU2) @ r(p["W2"]).T
    dpre2 = dh2 * (acts["h2"] > 0)
    d["b1"] = dpre2.sum(axis=0)
    dU1 = np.concatenate([dpre2, g.E.T @ dpre2], axis=1)
    d["W1"] = r(acts["h1"]).T @ r(dU1)
    dh1 = r(dU1) @ r(p["W1"]).T
    dpre1 = dh1 * (acts["h1"] > 0)
    d["b0"] = dpre1.sum(axis=0)
    d["W0"] = acts["z0"].T @ dpre1
    return d


def loss_grad(theta: np.ndarray, g: Graph, Z: np.ndarray, bf: bool = False):
    """L and dL/dtheta (flattened in SHAPES order)."""
    p = unflatten(theta)
    gdiag, glow, acts = forward(p, g, bf, keep=True)
    L, dgd, dgl = loss_and_ghat_grad(g, gdiag, glow, Z)
    d = backward(p, g, acts, gdiag, dgd, dgl, bf)
    return L, flatten(d)


def init_params(seed: int) -> np.ndarray:
    """The initial parameters tbsla_hip_precond_net_create draws (same order
    and generator): normal draws by Box-Muller on V(seed, .), scaled per
    tensor; the output heads start at zero (G^ = I)."""
    from oracle import unit_vector
    scale = {"W0": np.sqrt(2.0 / (2 * F)), "b0": 0.0, "W1": np.sqrt(2.0 / H), "b1": 0.0,
             "W2": np.sqrt(2.0 / H), "b2": 0.0, "V": np.sqrt(1.0 / H), "ve": 0.1, "c": 0.0,
             "wo": 0.0, "bo": 0.0, "wd": 0.0, "bd": 0.0}
    u = unit_vector(seed, 2 * N_PARAMS)
    z = np.sqrt(-2.0 * np.log(1.0 - u[0::2])) * np.cos(6.283185307179586 * u[1::2])
    out, o = [], 0
    for name, shp in SHAPES:
        k = int(np.prod(shp))
        out.append(z[o:o + k] * scale[name])
        o += k
    return np.concatenate(out)
