"""CPU oracle for TBSLA's sparse hot path -- test infrastructure only.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import this package; it is the oracle, never the product.

``liboracle.so`` (``tbsla_oracle.c``) is a plain-C restatement of the
reference routines; every C function cites the reference file:line it
follows.  ``ref`` (``oracle/_ref/libtbsla_ref.so``) is the reference's own
``src/cpp`` library compiled from its sources by ``oracle/Makefile`` -- it
exists only where the reference sources were present at build time, and is
used to pin the restatement.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libtbsla_ref.so")
REF_BENCH = os.path.join(HERE, "_ref", "ref_bench")

COO, SCOO, CSR, ELL, DENSE = 0, 1, 2, 3, 4
FORMATS = {"COO": COO, "SCOO": SCOO, "CSR": CSR, "ELL": ELL, "DENSE": DENSE}
GEOM_FIELDS = ("n_row", "n_col", "f_row", "f_col", "ln_row", "ln_col", "pr", "pc", "NR", "NC")


class Geom(C.Structure):
    _fields_ = [(f, C.c_int) for f in GEOM_FIELDS] + [("nnz", C.c_long)]


class CCsr(C.Structure):
    _fields_ = [("g", Geom), ("stored", C.c_long), ("rowptr", C.POINTER(C.c_int)),
                ("colidx", C.POINTER(C.c_int)), ("values", C.POINTER(C.c_double))]


class CEll(C.Structure):
    _fields_ = [("g", Geom), ("max_col", C.c_int), ("columns", C.POINTER(C.c_int)),
                ("values", C.POINTER(C.c_double))]


class CCoo(C.Structure):
    _fields_ = [("g", Geom), ("len", C.c_long), ("row", C.POINTER(C.c_int)),
                ("col", C.POINTER(C.c_int)), ("values", C.POINTER(C.c_double))]


class CDense(C.Structure):
    _fields_ = [("g", Geom), ("values", C.POINTER(C.c_double))]


_STRUCT = {COO: CCoo, SCOO: CCoo, CSR: CCsr, ELL: CEll, DENSE: CDense}
_FREE = {COO: "orc_coo_free", SCOO: "orc_coo_free", CSR: "orc_csr_free", ELL: "orc_ell_free",
         DENSE: "orc_dense_free"}
_FILL_CQ = {COO: "orc_coo_fill_cqmat", SCOO: "orc_scoo_fill_cqmat", CSR: "orc_csr_fill_cqmat",
            ELL: "orc_ell_fill_cqmat", DENSE: "orc_dense_fill_cqmat"}
_FILL_CD = {COO: "orc_coo_fill_cdiag", SCOO: "orc_scoo_fill_cdiag", CSR: "orc_csr_fill_cdiag",
            ELL: "orc_ell_fill_cdiag", DENSE: "orc_dense_fill_cdiag"}
_FROM_COO = {SCOO: "orc_scoo_from_coo", CSR: "orc_csr_from_coo", ELL: "orc_ell_from_coo",
             DENSE: "orc_dense_from_coo"}

_lib = None
_ref = None


def build(force: bool = False) -> None:
    """Compile liboracle.so (and oracle/_ref when the reference sources exist)."""
    if force or not os.path.exists(LIB_PATH):
        subprocess.run(["make", "-C", HERE, "-s"], check=True)


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = C.CDLL(LIB_PATH)
        _lib.orc_rand_r.argtypes = [C.POINTER(C.c_uint)]
        _lib.orc_rand_r.restype = C.c_int
        _lib.orc_cqmat_value.argtypes = [C.c_uint64, C.c_uint64, C.c_uint64, C.c_uint64,
                                         C.c_double, C.c_uint, C.POINTER(C.c_int64),
                                         C.POINTER(C.c_int64)]
        _lib.orc_cdiag_value.argtypes = [C.c_uint64] * 5 + [C.POINTER(C.c_int64)] * 2
        _lib.orc_lnv.argtypes = [C.c_long, C.c_int, C.c_int]
        _lib.orc_lnv.restype = C.c_long
        _lib.orc_pflv.argtypes = [C.c_long, C.c_int, C.c_int]
        _lib.orc_pflv.restype = C.c_long
        _lib.orc_unit_value.argtypes = [C.c_uint64, C.c_uint64]
        _lib.orc_unit_value.restype = C.c_double
        _lib.orc_cqmat_gnv.argtypes = [C.c_int] * 3
        _lib.orc_cqmat_gnv.restype = C.c_long
        for fmt, name in _FILL_CQ.items():
            f = getattr(_lib, name)
            f.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_double, C.c_uint,
                          C.c_int, C.c_int, C.c_int, C.c_int, C.c_int64]
        for fmt, name in _FILL_CD.items():
            getattr(_lib, name).argtypes = [C.c_void_p] + [C.c_int] * 7
        for name in _FROM_COO.values():
            getattr(_lib, name).argtypes = [C.c_void_p, C.c_void_p]
        for name in set(_FREE.values()):
            getattr(_lib, name).argtypes = [C.c_void_p]
        dp = C.POINTER(C.c_double)
        _lib.orc_Ax.argtypes = [C.c_int, C.c_void_p, dp, dp, C.c_int]
        _lib.orc_spmv.argtypes = [C.c_int, C.c_void_p, dp, dp, C.c_int]
        _lib.orc_a_axpx_.argtypes = [C.c_int, C.c_void_p, dp, dp, C.c_int]
        _lib.orc_AAxpAx.argtypes = [C.c_int, C.c_void_p, dp, dp, dp, C.c_int]
        _lib.orc_AAxpAxpx.argtypes = [C.c_int, C.c_void_p, dp, dp, dp, C.c_int]
        _lib.orc_result_len.argtypes = [C.c_int, C.c_void_p]
        _lib.orc_csr_Ax_omp.argtypes = [C.c_void_p, dp, dp]
        _lib.orc_set_threads.argtypes = [C.c_int]
        _lib.orc_get_col_sums.argtypes = [C.c_int, C.c_void_p, dp]
        _lib.orc_normalize_cols.argtypes = [C.c_int, C.c_void_p, dp]
        _lib.orc_make_stochastic.argtypes = [C.c_int, C.c_void_p, dp]
        _lib.orc_coo_read_mm.argtypes = [C.c_void_p, C.c_char_p]
        _lib.orc_coo_read_mm.restype = C.c_int
        _lib.orc_get_row_sums.argtypes = [C.c_int, C.c_void_p, dp]
        _lib.orc_normalize_rows.argtypes = [C.c_int, C.c_void_p, dp]
        _lib.orc_get_row_max_abs.argtypes = [C.c_int, C.c_void_p, dp]
        _lib.orc_apply_exponential.argtypes = [C.c_int, C.c_void_p, dp, C.c_int]
        _lib.orc_page_rank.argtypes = [C.c_int, C.c_void_p, C.c_double, C.c_double, C.c_int, dp,
                                       dp]
        _lib.orc_page_rank_source.argtypes = [C.c_int, C.c_void_p, C.c_double, C.c_double,
                                              C.c_int, C.c_int, dp, dp]
        _lib.orc_page_rank_max_norm.argtypes = [C.c_int, C.c_void_p, C.c_double, C.c_double,
                                                C.c_int, dp, dp]
        _lib.orc_cg.argtypes = [C.c_int, C.c_void_p, dp, dp, C.c_int, C.c_double, C.c_int, dp]
        _lib.orc_cg_clamp.argtypes = [C.c_int, C.c_void_p, dp, dp, C.c_int, C.c_double, C.c_int,
                                      C.c_double, dp]
        _lib.orc_csr_spmm.argtypes = [C.c_void_p, dp, dp, C.c_int]
        _lib.orc_poisson3d.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, dp]
        _lib.orc_uniform_vector.argtypes = [dp, C.c_long, C.c_uint64, C.c_double, C.c_double]
        _lib.orc_lognormal_field.argtypes = [dp, C.c_int, C.c_int, C.c_int, C.c_double,
                                             C.c_double, C.c_double, C.c_uint64]
        _lib.orc_sym_cqmat.argtypes = [C.c_void_p, C.c_void_p]
        _lib.orc_csr_transpose.argtypes = [C.c_void_p, C.c_void_p]
        _lib.orc_pcg.argtypes = [C.c_void_p, dp, dp, C.c_int, C.c_double, C.c_int, C.c_double,
                                 C.c_int, C.c_void_p, dp]
    return _lib


def _dp(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _arr(ptr, n, dtype):
    if n <= 0 or not ptr:
        return np.zeros(0, dtype=dtype)
    return np.ctypeslib.as_array(ptr, shape=(n,)).astype(dtype, copy=True)


# ---------------------------------------------------------------------------
# scalar generators
# ---------------------------------------------------------------------------

def rand_r(seed: int) -> tuple[int, int]:
    """glibc rand_r restated; returns (value, new_seed)."""
    s = C.c_uint(seed & 0xFFFFFFFF)
    v = lib().orc_rand_r(C.byref(s))
    return v, s.value


def cqmat_value(k, nr, nc, c, q, seed_mult):
    r, col = C.c_int64(), C.c_int64()
    lib().orc_cqmat_value(k, nr, nc, c, q, seed_mult, C.byref(r), C.byref(col))
    return r.value, col.value


def cdiag_value(i, nv, nr, nc, cdiag):
    r, col = C.c_int64(), C.c_int64()
    lib().orc_cdiag_value(i, nv, nr, nc, cdiag, C.byref(r), C.byref(col))
    return r.value, col.value


def lnv(size, pos, npos):
    return lib().orc_lnv(size, pos, npos)


def pflv(size, pos, npos):
    return lib().orc_pflv(size, pos, npos)


def unit_value(seed, k):
    return lib().orc_unit_value(seed, k)


def cqmat_gnv(n_row, n_col, c):
    return lib().orc_cqmat_gnv(n_row, n_col, c)


# ---------------------------------------------------------------------------
# matrices
# ---------------------------------------------------------------------------

class Matrix:
    """One oracle matrix block; owns its C struct.

    ``arrays`` returns numpy copies in the reference's layout:
    CSR (rowptr, colidx, values) with the stored entries only; ELL
    (columns, values) row-major ``ln_row x max_col``; COO/SCOO
    (row, col, values); DENSE (values,) row-major ``ln_row x ln_col``.
    """

    def __init__(self, fmt: int):
        self.fmt = fmt
        self.s = _STRUCT[fmt]()

    def __del__(self):
        try:
            getattr(lib(), _FREE[self.fmt])(C.byref(self.s))
        except Exception:
            pass

    @property
    def ptr(self):
        return C.cast(C.byref(self.s), C.c_void_p)

    @property
    def geom(self) -> dict:
        g = self.s.g
        d = {f: getattr(g, f) for f in GEOM_FIELDS}
        d["nnz"] = g.nnz
        if self.fmt == ELL:
            d["max_col"] = self.s.max_col
        return d

    def arrays(self):
        s, g = self.s, self.s.g
        if self.fmt == CSR:
            rp = _arr(s.rowptr, g.ln_row + 1, np.int32)
            n = int(rp[-1]) if len(rp) else 0
            return rp, _arr(s.colidx, n, np.int32), _arr(s.values, n, np.float64)
        if self.fmt == ELL:
            rows = g.ln_row if s.values else 0
            n = rows * s.max_col
            return (_arr(s.columns, n, np.int32).reshape(rows, s.max_col),
                    _arr(s.values, n, np.float64).reshape(rows, s.max_col))
        if self.fmt in (COO, SCOO):
            n = s.len
            return _arr(s.row, n, np.int32), _arr(s.col, n, np.int32), _arr(s.values, n, np.float64)
        n = g.ln_row * g.ln_col if s.values else 0
        return (_arr(s.values, n, np.float64).reshape(g.ln_row, g.ln_col) if n else
                np.zeros((g.ln_row, g.ln_col)),)

    def set_values(self, v) -> None:
        """Overwrite the value array in place (same layout as arrays())."""
        cur = self.arrays()[-1]
        v = np.ascontiguousarray(v, dtype=np.float64).reshape(-1)
        assert v.size == cur.size, (v.size, cur.size)
        if not self.s.values:  # a block without arrays reads as zeros
            assert not np.any(v)
            return
        if v.size:
            np.ctypeslib.as_array(self.s.values, shape=(v.size,))[:] = v

    @property
    def result_len(self) -> int:
        return lib().orc_result_len(self.fmt, self.ptr)

    def spmv(self, v, vect_incr: int = 0) -> np.ndarray:
        v = np.ascontiguousarray(v, dtype=np.float64)
        r = np.zeros(max(self.result_len, 1))
        lib().orc_spmv(self.fmt, self.ptr, _dp(r), _dp(v), vect_incr)
        return r[: self.result_len]

    def Ax(self, r, v, vect_incr: int = 0) -> np.ndarray:
        r = np.array(r, dtype=np.float64, copy=True)
        v = np.ascontiguousarray(v, dtype=np.float64)
        lib().orc_Ax(self.fmt, self.ptr, _dp(r), _dp(v), vect_incr)
        return r

    def a_axpx_(self, v, vect_incr: int = 0) -> np.ndarray:
        v = np.ascontiguousarray(v, dtype=np.float64)
        r = np.zeros(max(self.result_len, 1))
        lib().orc_a_axpx_(self.fmt, self.ptr, _dp(r), _dp(v), vect_incr)
        return r[: self.result_len]

    def AAxpAx(self, r, v, buf, vect_incr: int = 0):
        r = np.array(r, dtype=np.float64, copy=True)
        buf = np.array(buf, dtype=np.float64, copy=True)
        v = np.ascontiguousarray(v, dtype=np.float64)
        lib().orc_AAxpAx(self.fmt, self.ptr, _dp(r), _dp(v), _dp(buf), vect_incr)
        return r, buf

    def AAxpAxpx(self, r, v, buf, vect_incr: int = 0):
        r = np.array(r, dtype=np.float64, copy=True)
        buf = np.array(buf, dtype=np.float64, copy=True)
        v = np.ascontiguousarray(v, dtype=np.float64)
        lib().orc_AAxpAxpx(self.fmt, self.ptr, _dp(r), _dp(v), _dp(buf), vect_incr)
        return r, buf

    def make_stochastic(self) -> np.ndarray:
        s = np.zeros(max(self.s.g.ln_col, 1))
        lib().orc_make_stochastic(self.fmt, self.ptr, _dp(s))
        return s[: self.s.g.ln_col]

    def get_col_sums(self) -> np.ndarray:
        s = np.zeros(max(self.s.g.ln_col, 1))
        lib().orc_get_col_sums(self.fmt, self.ptr, _dp(s))
        return s[: self.s.g.ln_col]

    def row_sums_len(self) -> int:
        g = self.s.g
        return g.n_row if self.fmt in (ELL, SCOO) else g.ln_row

    def get_row_sums(self, s=None) -> np.ndarray:
        """SCOO accumulates into ``s`` (zeros if None); a copy is returned."""
        n = self.row_sums_len()
        out = np.zeros(max(n, 1)) if s is None else np.array(s, dtype=np.float64, copy=True)
        assert out.size >= n
        lib().orc_get_row_sums(self.fmt, self.ptr, _dp(out))
        return out[:n] if s is None else out

    def normalize_rows(self, s) -> None:
        s = np.ascontiguousarray(s, dtype=np.float64)
        if s.size == 0:
            s = np.zeros(1)
        lib().orc_normalize_rows(self.fmt, self.ptr, _dp(s))

    def get_row_max_abs(self) -> np.ndarray:
        n = self.s.g.ln_row
        m = np.zeros(max(n, 1))
        lib().orc_get_row_max_abs(self.fmt, self.ptr, _dp(m))
        return m[:n]

    def apply_exponential(self, max_abs=None, base: int = -1) -> None:
        if max_abs is None:
            lib().orc_apply_exponential(self.fmt, self.ptr, None, base)
        else:
            m = np.ascontiguousarray(max_abs, dtype=np.float64)
            lib().orc_apply_exponential(self.fmt, self.ptr, _dp(m), base)

    def page_rank(self, beta=0.85, epsilon=1e-10, max_iterations=10000):
        b = np.zeros(max(self.s.g.ln_col, 1))
        err = C.c_double()
        it = lib().orc_page_rank(self.fmt, self.ptr, beta, epsilon, max_iterations, _dp(b),
                                 C.byref(err))
        return b[: self.s.g.ln_col], it, err.value

    def page_rank_max_norm(self, beta=0.85, epsilon=1e-10, max_iterations=10000):
        """MatrixCOO's max-normalised PageRank (src/mpi/matrix/MatrixCOO.cpp:108-175)."""
        n = self.s.g.n_col
        b = np.zeros(max(n, 1))
        err = C.c_double()
        it = lib().orc_page_rank_max_norm(self.fmt, self.ptr, beta, epsilon, max_iterations,
                                          _dp(b), C.byref(err))
        return b[:n], it, err.value

    def page_rank_source(self, source, beta=0.85, epsilon=1e-10, max_iterations=10000):
        """Personalised PageRank: teleport to `source` only."""
        b = np.zeros(max(self.s.g.ln_col, 1))
        err = C.c_double()
        it = lib().orc_page_rank_source(self.fmt, self.ptr, beta, epsilon, max_iterations, source,
                                        _dp(b), C.byref(err))
        return b[: self.s.g.ln_col], it, err.value

    def cg(self, rhs, max_iterations=10000, tolerance=1e-10, relative=False, x0=None,
           near_zero=0.0):
        """near_zero: divisor clamp; 0 = the reference's 1e-10 for the
        absolute test, none for the relative one; < 0 none."""
        rhs = np.ascontiguousarray(rhs, dtype=np.float64)
        x = np.zeros_like(rhs) if x0 is None else np.array(x0, dtype=np.float64, copy=True)
        rn = C.c_double()
        nz = near_zero if near_zero > 0 else (1e-300 if (near_zero < 0 or relative) else 1e-10)
        it = lib().orc_cg_clamp(self.fmt, self.ptr, _dp(rhs), _dp(x), max_iterations, tolerance,
                                int(relative), nz, C.byref(rn))
        return x, it, rn.value

    def spmm(self, B: np.ndarray) -> np.ndarray:
        assert self.fmt == CSR
        B = np.ascontiguousarray(B, dtype=np.float64)
        k = B.shape[1]
        Cm = np.zeros((self.s.g.ln_row, k))
        lib().orc_csr_spmm(self.ptr, _dp(B), _dp(Cm), k)
        return Cm

    def csr_Ax_omp(self, r, v):
        assert self.fmt == CSR
        lib().orc_csr_Ax_omp(self.ptr, _dp(r), _dp(v))

    def to(self, fmt: int) -> "Matrix":
        """Conversion constructor <fmt>(const MatrixCOO&); self must be COO."""
        assert self.fmt == COO
        m = Matrix(fmt)
        getattr(lib(), _FROM_COO[fmt])(m.ptr, self.ptr)
        return m


def fill_cqmat(fmt, n_row, n_col, c, q, seed_mult=1, pr=0, pc=0, NR=1, NC=1,
               value_seed=-1) -> Matrix:
    m = Matrix(fmt)
    getattr(lib(), _FILL_CQ[fmt])(m.ptr, n_row, n_col, c, q, seed_mult, pr, pc, NR, NC,
                                  value_seed)
    return m


def coo_read_mm(path) -> tuple[int, "Matrix"]:
    """MatrixCOO::readMM (symmetric overwrite fixed): (status, COO matrix);
    status 0, -1 cannot open, -2 header rejected, -3 malformed."""
    m = Matrix(COO)
    rc = lib().orc_coo_read_mm(m.ptr, os.fsencode(path))
    return rc, m


def ref_read_mm(path) -> tuple[int, "RefMatrix"]:
    """The compiled reference's MatrixCOO::readMM (0, or -2 on its
    MatrixFormatReadException)."""
    m = RefMatrix(COO)
    rc = ref().ref_read_mm(m.h, os.fsencode(path))
    return rc, m


def fill_cdiag(fmt, n_row, n_col, cdiag, pr=0, pc=0, NR=1, NC=1) -> Matrix:
    m = Matrix(fmt)
    getattr(lib(), _FILL_CD[fmt])(m.ptr, n_row, n_col, cdiag, pr, pc, NR, NC)
    return m


def csr_from_arrays(n_row, n_col, rowptr, colidx, values) -> Matrix:
    """Single-block oracle CSR holding copies of the given arrays."""
    rowptr = np.ascontiguousarray(rowptr, dtype=np.int32)
    colidx = np.ascontiguousarray(colidx, dtype=np.int32)
    values = np.ascontiguousarray(values, dtype=np.float64)
    m = Matrix(CSR)
    lib().orc_csr_alloc.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_long]
    lib().orc_csr_alloc(m.ptr, n_row, n_col, len(colidx))
    C.memmove(m.s.rowptr, rowptr.ctypes.data, rowptr.nbytes)
    if len(colidx):
        C.memmove(m.s.colidx, colidx.ctypes.data, colidx.nbytes)
        C.memmove(m.s.values, values.ctypes.data, values.nbytes)
    return m


def csr_from_dense(A: np.ndarray) -> Matrix:
    rows, cols = np.nonzero(A)
    rowptr = np.zeros(A.shape[0] + 1, np.int64)
    np.add.at(rowptr, rows + 1, 1)
    return csr_from_arrays(A.shape[0], A.shape[1], np.cumsum(rowptr), cols, A[rows, cols])


def poisson3d(nx, ny, nz, kappa=None) -> Matrix:
    """7-point -div(k grad u) with Dirichlet boundaries (orc_poisson3d)."""
    m = Matrix(CSR)
    k = None if kappa is None else np.ascontiguousarray(kappa, np.float64)
    lib().orc_poisson3d(m.ptr, nx, ny, nz, None if k is None else _dp(k))
    return m


def uniform_vector(n, seed, lo=0.0, hi=1.0) -> np.ndarray:
    out = np.zeros(n)
    lib().orc_uniform_vector(_dp(out), n, seed, lo, hi)
    return out


def lognormal_field(nx, ny, nz, mu=0.0, sigma=1.0, corr_len=4.0, seed=0) -> np.ndarray:
    out = np.zeros(nx * ny * nz)
    lib().orc_lognormal_field(_dp(out), nx, ny, nz, mu, sigma, corr_len, seed)
    return out


def sym_cqmat(n, c, q, seed_mult=1, value_seed=0) -> Matrix:
    """A = (B + B^T)/2, diagonal = off-diagonal |row| sum + 1 (orc_sym_cqmat)."""
    B = fill_cqmat(CSR, n, n, c, q, seed_mult, value_seed=value_seed)
    A = Matrix(CSR)
    lib().orc_sym_cqmat(A.ptr, B.ptr)
    return A


def transpose(A: Matrix) -> Matrix:
    T = Matrix(CSR)
    lib().orc_csr_transpose(T.ptr, A.ptr)
    return T


def pcg(A: Matrix, rhs, max_iterations=10000, tolerance=1e-8, relative=True, near_zero=1e-300,
        kind=0, G: Matrix | None = None):
    """Preconditioned CG from x = 0 (kind 0 none, 1 Jacobi, 2 M = G G^T)."""
    rhs = np.ascontiguousarray(rhs, np.float64)
    x = np.zeros_like(rhs)
    rn = C.c_double()
    it = lib().orc_pcg(A.ptr, _dp(rhs), _dp(x), max_iterations, tolerance, int(relative),
                       near_zero, kind, None if G is None else G.ptr, C.byref(rn))
    return x, it, rn.value


def coo_from_arrays(n_row, n_col, row, col, values) -> Matrix:
    row = np.ascontiguousarray(row, dtype=np.int32)
    col = np.ascontiguousarray(col, dtype=np.int32)
    values = np.ascontiguousarray(values, dtype=np.float64)
    m = Matrix(COO)
    lib().orc_coo_alloc.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_long]
    lib().orc_coo_alloc(m.ptr, n_row, n_col, len(row))
    if len(row):
        C.memmove(m.s.row, row.ctypes.data, row.nbytes)
        C.memmove(m.s.col, col.ctypes.data, col.nbytes)
        C.memmove(m.s.values, values.ctypes.data, values.nbytes)
    return m


def unit_vector(seed: int, n: int) -> np.ndarray:
    """x_i = V(seed, i), the project value generator, vectorised in numpy
    (same integer steps as orc_unit_value)."""
    M = np.uint64(0xFFFFFFFFFFFFFFFF)

    def mix(z):
        z = (z + np.uint64(0x9E3779B97F4A7C15)) & M
        z = ((z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)) & M
        z = ((z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)) & M
        return z ^ (z >> np.uint64(31))

    with np.errstate(over="ignore"):
        s = mix(np.array([seed], dtype=np.uint64))[0]
        h = mix(np.arange(n, dtype=np.uint64) ^ s)
    return (h >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


# ---------------------------------------------------------------------------
# compiled reference (pins the oracle; present only where it was built)
# ---------------------------------------------------------------------------

def ref_available() -> bool:
    return os.path.exists(REF_PATH)


def ref() -> C.CDLL:
    global _ref
    if _ref is None:
        _ref = C.CDLL(REF_PATH)
        vp, ip, dp = C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_double)
        _ref.ref_new.restype = vp
        _ref.ref_new.argtypes = [C.c_int]
        _ref.ref_delete.argtypes = [vp]
        _ref.ref_fill_cqmat.argtypes = [vp, C.c_int, C.c_int, C.c_int, C.c_double, C.c_uint,
                                        C.c_int, C.c_int, C.c_int, C.c_int]
        _ref.ref_fill_cdiag.argtypes = [vp] + [C.c_int] * 7
        _ref.ref_geom.argtypes = [vp, ip, C.POINTER(C.c_long)]
        _ref.ref_arrays.argtypes = [C.c_int, vp, C.POINTER(dp), C.POINTER(ip), C.POINTER(ip), ip]
        _ref.ref_from_coo.restype = vp
        _ref.ref_from_coo.argtypes = [C.c_int, vp]
        _ref.ref_spmv.argtypes = [vp, dp, dp, C.c_int, C.c_int]
        _ref.ref_Ax.argtypes = [vp, dp, dp, C.c_int]
        _ref.ref_a_axpx.argtypes = [vp, dp, dp, C.c_int, C.c_int]
        _ref.ref_AAxpAx.argtypes = [vp, dp, dp, dp, C.c_int]
        _ref.ref_AAxpAxpx.argtypes = [vp, dp, dp, dp, C.c_int]
        _ref.ref_make_stochastic.argtypes = [vp, dp]
        _ref.ref_read_mm.argtypes = [vp, C.c_char_p]
        _ref.ref_read_mm.restype = C.c_int
        _ref.ref_get_row_sums.argtypes = [vp, dp]
        _ref.ref_normalize_rows.argtypes = [vp, dp]
        _ref.ref_get_row_max_abs.argtypes = [vp, dp]
        _ref.ref_apply_exponential.argtypes = [vp, dp, C.c_int]
        _ref.ref_cqmat_value.argtypes = [C.c_ulong] * 4 + [C.c_double, C.c_uint,
                                                           C.POINTER(C.c_long),
                                                           C.POINTER(C.c_long)]
        _ref.ref_cdiag_value.argtypes = [C.c_ulong] * 5 + [C.POINTER(C.c_long)] * 2
        _ref.ref_write.restype = C.c_long
        _ref.ref_write.argtypes = [vp, C.POINTER(C.POINTER(C.c_char))]
        _ref.ref_free.argtypes = [vp]
        _ref.ref_set_threads.argtypes = [C.c_int]
        _ref.ref_set_threads(1)
    return _ref


class RefMatrix:
    """A reference tbsla::cpp::Matrix object (compiled reference)."""

    def __init__(self, fmt: int, handle=None):
        self.fmt = fmt
        self.h = handle if handle is not None else ref().ref_new(fmt)

    def __del__(self):
        try:
            ref().ref_delete(self.h)
        except Exception:
            pass

    @property
    def geom(self) -> dict:
        g = (C.c_int * 10)()
        nnz = C.c_long()
        ref().ref_geom(self.h, g, C.byref(nnz))
        d = dict(zip(GEOM_FIELDS, list(g)))
        d["nnz"] = nnz.value
        if self.fmt == ELL:
            d["max_col"] = self._ptrs()[3]
        return d

    def _ptrs(self):
        v = C.POINTER(C.c_double)()
        a = C.POINTER(C.c_int)()
        b = C.POINTER(C.c_int)()
        mc = C.c_int()
        ref().ref_arrays(self.fmt, self.h, C.byref(v), C.byref(a), C.byref(b), C.byref(mc))
        return v, a, b, mc.value

    def arrays(self, stored: int | None = None):
        """Same layout as Matrix.arrays.  For CSR only rowptr[ln_row] entries
        are read (the count-pass nnz field can exceed what was written); for
        SCOO/COO ``stored`` defaults to the nnz field."""
        g = self.geom
        v, a, b, mc = self._ptrs()
        if self.fmt == CSR:
            if not a:
                return np.zeros(g["ln_row"] + 1, np.int32), np.zeros(0, np.int32), np.zeros(0)
            rp = _arr(a, g["ln_row"] + 1, np.int32)
            n = int(rp[-1])
            return rp, _arr(b, n, np.int32), _arr(v, n, np.float64)
        if self.fmt == ELL:
            rows = g["ln_row"] if v else 0
            n = rows * mc
            return (_arr(a, n, np.int32).reshape(rows, mc), _arr(v, n, np.float64).reshape(rows, mc))
        if self.fmt in (COO, SCOO):
            n = g["nnz"] if stored is None else stored
            if not v:
                n = 0
            return _arr(a, n, np.int32), _arr(b, n, np.int32), _arr(v, n, np.float64)
        n = g["ln_row"] * g["ln_col"] if v else 0
        return (_arr(v, n, np.float64).reshape(g["ln_row"], g["ln_col"]) if n else
                np.zeros((g["ln_row"], g["ln_col"])),)

    def to(self, fmt: int) -> "RefMatrix":
        assert self.fmt == COO
        return RefMatrix(fmt, ref().ref_from_coo(fmt, self.h))

    def spmv(self, v, length: int, vect_incr: int = 0) -> np.ndarray:
        v = np.ascontiguousarray(v, dtype=np.float64)
        r = np.zeros(max(length, 1))
        ref().ref_spmv(self.h, _dp(v), _dp(r), length, vect_incr)
        return r[:length]

    def a_axpx_(self, v, length: int, vect_incr: int = 0) -> np.ndarray:
        v = np.ascontiguousarray(v, dtype=np.float64)
        r = np.zeros(max(length, 1))
        ref().ref_a_axpx(self.h, _dp(v), _dp(r), length, vect_incr)
        return r[:length]

    def make_stochastic(self, ln_col: int) -> np.ndarray:
        s = np.zeros(max(ln_col, 1))
        ref().ref_make_stochastic(self.h, _dp(s))
        return s[:ln_col]

    def set_values(self, v) -> None:
        """Overwrite the object's value array in place (layout of arrays())."""
        cur = self.arrays()[-1]
        v = np.ascontiguousarray(v, dtype=np.float64).reshape(-1)
        assert v.size == cur.size, (v.size, cur.size)
        if v.size:
            np.ctypeslib.as_array(self._ptrs()[0], shape=(v.size,))[:] = v

    def get_row_sums(self, s: np.ndarray) -> np.ndarray:
        out = np.array(s, dtype=np.float64, copy=True)
        ref().ref_get_row_sums(self.h, _dp(out))
        return out

    def normalize_rows(self, s: np.ndarray) -> None:
        s = np.array(s, dtype=np.float64, copy=True)
        ref().ref_normalize_rows(self.h, _dp(s))

    def get_row_max_abs(self, n: int) -> np.ndarray:
        m = np.zeros(max(n, 1))
        ref().ref_get_row_max_abs(self.h, _dp(m))
        return m[:n]

    def apply_exponential(self, max_abs=None, base: int = -1) -> None:
        if max_abs is None:
            ref().ref_apply_exponential(self.h, None, base)
        else:
            m = np.array(max_abs, dtype=np.float64, copy=True)
            ref().ref_apply_exponential(self.h, _dp(m), base)

    def write(self) -> bytes:
        p = C.POINTER(C.c_char)()
        n = ref().ref_write(self.h, C.byref(p))
        data = C.string_at(p, n)
        ref().ref_free(p)
        return data


def ref_fill_cqmat(fmt, n_row, n_col, c, q, seed_mult=1, pr=0, pc=0, NR=1, NC=1) -> RefMatrix:
    m = RefMatrix(fmt)
    ref().ref_fill_cqmat(m.h, n_row, n_col, c, q, seed_mult, pr, pc, NR, NC)
    return m


def ref_fill_cdiag(fmt, n_row, n_col, cdiag, pr=0, pc=0, NR=1, NC=1) -> RefMatrix:
    m = RefMatrix(fmt)
    ref().ref_fill_cdiag(m.h, n_row, n_col, cdiag, pr, pc, NR, NC)
    return m


def ref_cqmat_value(k, nr, nc, c, q, seed_mult):
    r, col = C.c_long(), C.c_long()
    ref().ref_cqmat_value(k, nr, nc, c, q, seed_mult, C.byref(r), C.byref(col))
    return r.value, col.value


def ref_cdiag_value(i, nv, nr, nc, cdiag):
    r, col = C.c_long(), C.c_long()
    ref().ref_cdiag_value(i, nv, nr, nc, cdiag, C.byref(r), C.byref(col))
    return r.value, col.value
