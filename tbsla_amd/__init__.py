"""tbsla_amd -- MI355X (gfx950) implementation of TBSLA's sparse hot path.

The product is ``libtbsla_hip.so`` (HIP kernels behind the C ABI declared in
``include/tbsla_hip.h``).  This package loads it with ctypes and mirrors the
reference's ``tbsla::cpp::Matrix`` API (src/cpp/include/tbsla/cpp/Matrix.hpp)
for Python callers and tests: ``Matrix.fill_cqmat``, ``spmv``, ``Ax``,
``a_axpx_``, ... with the same argument meaning.

There is no CPU fallback: if the library is missing, or no GPU is visible,
every compute call raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libtbsla_hip.so")

COO, SCOO, CSR, ELL, DENSE = 0, 1, 2, 3, 4
FORMATS = {"COO": COO, "SCOO": SCOO, "CSR": CSR, "ELL": ELL, "DENSE": DENSE}
GEOM_FIELDS = ("n_row", "n_col", "f_row", "f_col", "ln_row", "ln_col", "pr", "pc", "NR", "NC")

# every entry point of include/tbsla_hip.h: (name, restype, argtypes)
_vp, _dp, _ip = C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_int)


class CGOptions(C.Structure):
    _fields_ = [("max_iterations", C.c_int), ("tolerance", C.c_double), ("relative", C.c_int),
                ("precision", C.c_int), ("preconditioner", C.c_int), ("replace_every", C.c_int),
                ("use_initial_guess", C.c_int), ("near_zero", C.c_double),
                ("factor", C.c_void_p), ("factor_t", C.c_void_p), ("graph_batch", C.c_int),
                ("lanes_per_row", C.c_int), ("stall_window", C.c_int), ("factor_form", C.c_int),
                ("max_blocks_per_system", C.c_int), ("spmv_passes", C.c_int)]


class CGResult(C.Structure):
    _fields_ = [("iterations", C.c_int), ("status", C.c_int), ("res_norm", C.c_double),
                ("time_us", C.c_double)]


CG_FP64, CG_MIXED = 0, 1
PRECOND_NONE, PRECOND_JACOBI, PRECOND_FACTORED = 0, 1, 2
CSR_ROWS, CSR_SPLIT, CSR_DIAG = 0, 1, 2
FACTOR_AUTO, FACTOR_ROWS, FACTOR_DIAGONALS = 0, 1, 2
CG_CONVERGED, CG_MAX_ITERATIONS, CG_BREAKDOWN, CG_INACTIVE = 0, 1, 2, 3
ERR_INVALID, ERR_HIP, ERR_ALLOC, ERR_UNSUPPORTED, ERR_RANGE, ERR_BREAKDOWN = 1, 2, 3, 4, 5, 6


class Geometry(C.Structure):
    _fields_ = ([(f, C.c_int) for f in GEOM_FIELDS] +
                [("nnz", C.c_longlong), ("stored", C.c_longlong), ("max_col", C.c_int),
                 ("format", C.c_int)])


_GP = C.POINTER(Geometry)
SIGNATURES = [
    ("tbsla_hip_status_string", C.c_char_p, [C.c_int]),
    ("tbsla_hip_last_error", C.c_char_p, []),
    ("tbsla_hip_version", C.c_int, []),
    ("tbsla_hip_malloc", C.c_int, [C.POINTER(C.c_void_p), C.c_size_t]),
    ("tbsla_hip_free", C.c_int, [_vp]),
    ("tbsla_hip_memcpy_htod", C.c_int, [_vp, _vp, C.c_size_t, _vp]),
    ("tbsla_hip_memcpy_dtoh", C.c_int, [_vp, _vp, C.c_size_t, _vp]),
    ("tbsla_hip_memcpy_dtod", C.c_int, [_vp, _vp, C.c_size_t, _vp]),
    ("tbsla_hip_stream_synchronize", C.c_int, [_vp]),
    ("tbsla_hip_copy_kernel", C.c_int, [_vp, _vp, C.c_size_t, _vp]),
    ("tbsla_hip_read_kernel", C.c_int, [_vp, C.c_size_t, _vp, _vp]),
    ("tbsla_hip_device_count", C.c_int, [_ip]),
    ("tbsla_hip_create", C.c_int, [C.POINTER(C.c_void_p), C.c_int]),
    ("tbsla_hip_destroy", C.c_int, [_vp]),
    ("tbsla_hip_get_geometry", C.c_int, [_vp, _GP]),
    ("tbsla_hip_fill_cdiag", C.c_int, [_vp] + [C.c_int] * 7 + [_vp]),
    ("tbsla_hip_fill_cqmat", C.c_int, [_vp, C.c_int, C.c_int, C.c_int, C.c_double, C.c_uint,
                                       C.c_int, C.c_int, C.c_int, C.c_int, C.c_longlong, _vp]),
    ("tbsla_hip_fill_unit_vector", C.c_int, [_vp, C.c_longlong, C.c_longlong, _vp]),
    ("tbsla_hip_fill_uniform_vector", C.c_int, [_vp, C.c_longlong, C.c_ulonglong, C.c_double,
                                                C.c_double, _vp]),
    ("tbsla_hip_fill_poisson3d", C.c_int, [_vp, C.c_int, C.c_int, C.c_int, _vp, _vp]),
    ("tbsla_hip_fill_lognormal_field", C.c_int, [_vp, C.c_int, C.c_int, C.c_int, C.c_double,
                                                 C.c_double, C.c_double, C.c_ulonglong, _vp]),
    ("tbsla_hip_fill_sym_cqmat", C.c_int, [_vp, C.c_int, C.c_int, C.c_double, C.c_uint,
                                           C.c_longlong, _vp]),
    ("tbsla_hip_transpose", C.c_int, [_vp, _vp, _vp]),
    ("tbsla_hip_block_diag", C.c_int, [_vp, _vp, C.c_int, _vp]),
    ("tbsla_hip_export", C.c_int, [_vp, _vp, _vp, _vp, _vp]),
    ("tbsla_hip_import", C.c_int, [_vp, C.c_int, _GP, _vp, _vp, _vp, _vp]),
    ("tbsla_hip_convert", C.c_int, [_vp, C.c_int, _vp, _vp]),
    ("tbsla_hip_device_arrays", C.c_int, [_vp, C.POINTER(C.c_void_p), C.POINTER(C.c_void_p),
                                          C.POINTER(C.c_void_p)]),
    ("tbsla_hip_values_changed", C.c_int, [_vp, _vp]),
    ("tbsla_hip_write", C.c_int, [_vp, _vp, C.c_size_t, C.POINTER(C.c_size_t), _vp]),
    ("tbsla_hip_read", C.c_int, [_vp, _vp, C.c_size_t, _vp]),
    ("tbsla_hip_write_file", C.c_int, [_vp, C.c_char_p, _vp]),
    ("tbsla_hip_read_file", C.c_int, [_vp, C.c_char_p, C.c_longlong, _vp]),
    ("tbsla_hip_Ax", C.c_int, [_vp, _vp, _vp, C.c_int, _vp]),
    ("tbsla_hip_Ax_noalloc", C.c_int, [_vp, _vp, _vp, C.c_int, _vp]),
    ("tbsla_hip_spmv", C.c_int, [_vp, _vp, _vp, C.c_int, _vp]),
    ("tbsla_hip_a_axpx", C.c_int, [_vp, _vp, _vp, _vp, C.c_int, _vp]),
    ("tbsla_hip_AAxpAx", C.c_int, [_vp, _vp, _vp, _vp, C.c_int, _vp]),
    ("tbsla_hip_AAxpAxpx", C.c_int, [_vp, _vp, _vp, _vp, C.c_int, _vp]),
    ("tbsla_hip_time_spmv", C.c_int, [_vp, _vp, _vp, C.c_int, C.POINTER(C.c_float), _vp]),
    ("tbsla_hip_get_col_sums", C.c_int, [_vp, _vp, _vp]),
    ("tbsla_hip_normalize_cols", C.c_int, [_vp, _vp, _vp]),
    ("tbsla_hip_make_stochastic", C.c_int, [_vp, _vp, _vp]),
    ("tbsla_hip_read_mm", C.c_int, [_vp, C.c_char_p, _vp]),
    ("tbsla_hip_mm_parse", C.c_int, [C.c_char_p, _ip, _ip, C.POINTER(C.c_longlong), _vp, _vp,
                                     _vp, C.c_longlong]),
    ("tbsla_hip_get_row_sums", C.c_int, [_vp, _vp, _vp]),
    ("tbsla_hip_normalize_rows", C.c_int, [_vp, _vp, _vp]),
    ("tbsla_hip_get_row_max_abs", C.c_int, [_vp, _vp, _vp]),
    ("tbsla_hip_apply_exponential", C.c_int, [_vp, _vp, C.c_int, _vp]),
    ("tbsla_hip_softmax_rows", C.c_int, [_vp, C.c_int, _vp, _vp, _vp]),
    ("tbsla_hip_page_rank", C.c_int, [_vp, C.c_double, C.c_double, C.c_int, _vp, _ip,
                                      C.POINTER(C.c_double), _vp]),
    ("tbsla_hip_page_rank_max_norm", C.c_int, [_vp, C.c_double, C.c_double, C.c_int, _vp, _ip,
                                      C.POINTER(C.c_double), _vp]),
    ("tbsla_hip_cg", C.c_int, [_vp, _vp, _vp, C.c_void_p, _ip, C.POINTER(C.c_double), _vp]),
    ("tbsla_hip_page_rank_batched", C.c_int, [_vp, C.c_int, _ip, C.c_double, C.c_double, C.c_int,
                                              _vp, _ip, C.POINTER(C.c_double),
                                              C.POINTER(C.c_double), _vp]),
    ("tbsla_hip_precond_net_param_count", C.c_int, []),
    ("tbsla_hip_precond_net_create", C.c_int, [C.POINTER(C.c_void_p), C.c_ulonglong]),
    ("tbsla_hip_precond_net_destroy", C.c_int, [_vp]),
    ("tbsla_hip_precond_net_get_params", C.c_int, [_vp, C.POINTER(C.c_float)]),
    ("tbsla_hip_precond_net_set_params", C.c_int, [_vp, C.POINTER(C.c_float)]),
    ("tbsla_hip_precond_net_apply", C.c_int, [_vp, _vp, _vp, _vp, _vp]),
    ("tbsla_hip_precond_net_loss_grad", C.c_int, [_vp, _vp, C.c_int, C.c_ulonglong,
                                                  C.POINTER(C.c_double), C.POINTER(C.c_float),
                                                  _vp]),
    ("tbsla_hip_precond_net_train_step", C.c_int, [_vp, _vp, C.c_int, C.c_ulonglong, C.c_float,
                                                   C.c_float, C.POINTER(C.c_double), _vp]),
    ("tbsla_hip_set_csr_layout", C.c_int, [_vp, C.c_int, C.c_int, C.c_int, _vp]),
    ("tbsla_hip_get_csr_layout", C.c_int, [_vp, _ip, _ip, _ip, C.POINTER(C.c_longlong)]),
    ("tbsla_hip_get_csr_remote_form", C.c_int, [_vp, _ip]),
    ("tbsla_hip_cg_batched", C.c_int, [_vp, C.c_int, _ip, _ip, _vp, _vp, C.c_void_p, C.c_void_p,
                                       _vp]),
    ("tbsla_hip_spmm", C.c_int, [_vp, _vp, _vp, C.c_int, _vp]),
    ("tbsla_hip_personalized_page_rank", C.c_int, [_vp, C.c_double, C.c_double, C.c_int, _ip, _vp,
                                                   _ip, C.POINTER(C.c_double), _vp]),
    ("tbsla_hip_set_csr_schedule", C.c_int, [_vp, C.c_int, C.c_int]),
    ("tbsla_hip_get_csr_schedule", C.c_int, [_vp, _ip, _ip]),
    ("tbsla_hip_set_csr_kernel", C.c_int, [_vp, C.c_int, C.c_int, C.c_int, C.c_int]),
    ("tbsla_hip_get_csr_kernel", C.c_int, [_vp, _ip, _ip, _ip, _ip]),
    ("tbsla_hip_autotune", C.c_int, [_vp, _vp, _vp, C.c_int, C.POINTER(C.c_float), _vp]),
]

_lib = None


class TbslaHipError(RuntimeError):
    """A failed library call; .status is the TBSLA_HIP_ERR_* code (0 when the
    library itself could not be loaded)."""

    def __init__(self, msg: str, status: int = 0):
        super().__init__(msg)
        self.status = status


def lib() -> C.CDLL:
    """The loaded libtbsla_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise TbslaHipError(f"{LIB_PATH} is not built: run `python -m tbsla_amd.build` "
                                "(or __graft_entry__.build())")
        lib_ = C.CDLL(LIB_PATH)
        for name, res, args in SIGNATURES:
            f = getattr(lib_, name)
            f.restype = res
            f.argtypes = args
        _lib = lib_
    return _lib


def check(status: int) -> None:
    if status != 0:
        L = lib()
        raise TbslaHipError(f"{L.tbsla_hip_status_string(status).decode()}: "
                            f"{L.tbsla_hip_last_error().decode()}", status)


def device_count() -> int:
    n = C.c_int(0)
    st = lib().tbsla_hip_device_count(C.byref(n))
    return n.value if st == 0 else 0


class DeviceBuffer:
    """Device memory owned through tbsla_hip_malloc / tbsla_hip_free."""

    def __init__(self, nbytes: int):
        self.nbytes = int(nbytes)
        p = C.c_void_p()
        check(lib().tbsla_hip_malloc(C.byref(p), max(self.nbytes, 1)))
        self.ptr = p.value

    def __del__(self):
        try:
            if self.ptr:
                lib().tbsla_hip_free(self.ptr)
        except Exception:
            pass
        self.ptr = None

    @classmethod
    def from_numpy(cls, a: np.ndarray) -> "DeviceBuffer":
        a = np.ascontiguousarray(a)
        b = cls(a.nbytes)
        if a.nbytes:
            check(lib().tbsla_hip_memcpy_htod(b.ptr, a.ctypes.data, a.nbytes, None))
        return b

    def to_numpy(self, dtype=np.float64, count: int | None = None) -> np.ndarray:
        dtype = np.dtype(dtype)
        count = self.nbytes // dtype.itemsize if count is None else count
        out = np.empty(count, dtype=dtype)
        if out.nbytes:
            check(lib().tbsla_hip_memcpy_dtoh(out.ctypes.data, self.ptr, out.nbytes, None))
        return out

    def upload(self, a: np.ndarray) -> None:
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.nbytes
        if a.nbytes:
            check(lib().tbsla_hip_memcpy_htod(self.ptr, a.ctypes.data, a.nbytes, None))


def lognormal_field(nx, ny, nz, mu=0.0, sigma=1.0, corr_len=4.0, seed=0) -> DeviceBuffer:
    """Smoothed log-normal coefficient field on the device (nx*ny*nz doubles)."""
    d = DeviceBuffer(nx * ny * nz * 8)
    check(lib().tbsla_hip_fill_lognormal_field(d.ptr, nx, ny, nz, mu, sigma, corr_len, seed, None))
    return d


def mm_parse(path):
    """The Matrix Market parse of read_mm into host arrays (no device work):
    (n_row, n_col, rows, cols, values), 0-based, in file order."""
    nr, nc, nnz = C.c_int(), C.c_int(), C.c_longlong()
    p = os.fsencode(path)
    check(lib().tbsla_hip_mm_parse(p, C.byref(nr), C.byref(nc), C.byref(nnz), None, None, None, 0))
    n = nnz.value
    r, c, v = np.zeros(max(n, 1), np.int32), np.zeros(max(n, 1), np.int32), np.zeros(max(n, 1))
    check(lib().tbsla_hip_mm_parse(p, C.byref(nr), C.byref(nc), C.byref(nnz), r.ctypes.data,
                                   c.ctypes.data, v.ctypes.data, max(n, 1)))
    return nr.value, nc.value, r[:n], c[:n], v[:n]


def uniform_vector(n, seed, lo=0.0, hi=1.0) -> DeviceBuffer:
    d = DeviceBuffer(max(n, 1) * 8)
    check(lib().tbsla_hip_fill_uniform_vector(d.ptr, n, seed, lo, hi, None))
    return d


def _dev(v, dtype=np.float64) -> DeviceBuffer:
    return v if isinstance(v, DeviceBuffer) else DeviceBuffer.from_numpy(np.asarray(v, dtype))


class Matrix:
    """A matrix block on the GPU: the reference's tbsla::cpp::Matrix API.

    Vectors may be numpy arrays (copied to the device per call) or
    DeviceBuffers (used in place).  Results come back as numpy arrays.
    """

    def __init__(self, fmt: int, handle=None):
        self.fmt = fmt
        if handle is None:
            h = C.c_void_p()
            check(lib().tbsla_hip_create(C.byref(h), fmt))
            handle = h.value
        self.h = handle

    def __del__(self):
        try:
            if self.h:
                lib().tbsla_hip_destroy(self.h)
        except Exception:
            pass
        self.h = None

    # -- generators ---------------------------------------------------------
    def fill_cqmat(self, n_row, n_col, c, q, seed_mult=1, pr=0, pc=0, NR=1, NC=1,
                   value_seed=-1, stream=None):
        check(lib().tbsla_hip_fill_cqmat(self.h, n_row, n_col, c, q, seed_mult, pr, pc, NR, NC,
                                         value_seed, stream))
        return self

    def fill_cdiag(self, n_row, n_col, cdiag, pr=0, pc=0, NR=1, NC=1, stream=None):
        check(lib().tbsla_hip_fill_cdiag(self.h, n_row, n_col, cdiag, pr, pc, NR, NC, stream))
        return self

    def fill_poisson3d(self, nx, ny, nz, kappa=None, stream=None):
        """7-point -div(k grad u) (CSR); kappa: None (k = 1), numpy or DeviceBuffer."""
        dk = None if kappa is None else _dev(kappa)
        check(lib().tbsla_hip_fill_poisson3d(self.h, nx, ny, nz, dk.ptr if dk else None, stream))
        return self

    def fill_sym_cqmat(self, n, c, q, seed_mult=1, value_seed=0, stream=None):
        check(lib().tbsla_hip_fill_sym_cqmat(self.h, n, c, q, seed_mult, value_seed, stream))
        return self

    def transpose(self) -> "Matrix":
        t = Matrix(CSR)
        check(lib().tbsla_hip_transpose(t.h, self.h, None))
        return t

    @classmethod
    def block_diag(cls, parts) -> "Matrix":
        m = cls(CSR)
        arr = (C.c_void_p * len(parts))(*[p.h for p in parts])
        check(lib().tbsla_hip_block_diag(m.h, arr, len(parts), None))
        return m

    # -- geometry -----------------------------------------------------------
    def geometry(self) -> Geometry:
        g = Geometry()
        check(lib().tbsla_hip_get_geometry(self.h, C.byref(g)))
        return g

    @property
    def geom(self) -> dict:
        g = self.geometry()
        d = {f: getattr(g, f) for f in GEOM_FIELDS}
        d["nnz"] = g.nnz
        if self.fmt == ELL:
            d["max_col"] = g.max_col
        return d

    def get_nnz(self) -> int:
        return self.geometry().nnz

    @property
    def result_len(self) -> int:
        g = self.geometry()
        return g.n_row if self.fmt == COO else g.ln_row

    def arrays(self):
        """Numpy copies in the reference's layout (as oracle.Matrix.arrays)."""
        g = self.geometry()
        L = lib()
        if self.fmt == CSR:
            rp = np.zeros(g.ln_row + 1, np.int32)
            ci = np.zeros(max(g.stored, 0), np.int32)
            va = np.zeros(max(g.stored, 0), np.float64)
            check(L.tbsla_hip_export(self.h, rp.ctypes.data, ci.ctypes.data, va.ctypes.data, None))
            return rp, ci, va
        if self.fmt in (COO, SCOO):
            n = max(g.stored, 0)
            r, c, v = np.zeros(n, np.int32), np.zeros(n, np.int32), np.zeros(n)
            check(L.tbsla_hip_export(self.h, r.ctypes.data, c.ctypes.data, v.ctypes.data, None))
            return r, c, v
        if self.fmt == ELL:
            rows = g.ln_row if g.stored > 0 else 0
            cols = np.zeros((rows, g.max_col), np.int32)
            vals = np.zeros((rows, g.max_col))
            if rows * g.max_col:
                check(L.tbsla_hip_export(self.h, cols.ctypes.data, None, vals.ctypes.data, None))
            return cols, vals
        vals = np.zeros((g.ln_row, g.ln_col))
        if g.stored > 0:
            check(L.tbsla_hip_export(self.h, None, None, vals.ctypes.data, None))
        return (vals,)

    # -- SpMV ---------------------------------------------------------------
    def spmv(self, v, vect_incr: int = 0) -> np.ndarray:
        dv = _dev(v)
        n = self.result_len
        dr = DeviceBuffer(max(n, 1) * 8)
        check(lib().tbsla_hip_spmv(self.h, dr.ptr, dv.ptr, vect_incr, None))
        return dr.to_numpy(np.float64, n)

    def Ax(self, r, v, vect_incr: int = 0) -> np.ndarray:
        """Matrix::Ax into a copy of r (overwrite or accumulate per format)."""
        dv = _dev(v)
        r = np.asarray(r, np.float64)
        dr = DeviceBuffer.from_numpy(r)
        check(lib().tbsla_hip_Ax(self.h, dr.ptr, dv.ptr, vect_incr, None))
        return dr.to_numpy(np.float64, len(r))

    def a_axpx_(self, v, vect_incr: int = 0) -> np.ndarray:
        dv = _dev(v)
        n = self.result_len
        dr, dt = DeviceBuffer(max(n, 1) * 8), DeviceBuffer(max(n, 1) * 8)
        check(lib().tbsla_hip_a_axpx(self.h, dr.ptr, dv.ptr, dt.ptr, vect_incr, None))
        return dr.to_numpy(np.float64, n)

    def AAxpAx(self, r, v, buf, vect_incr: int = 0):
        dr, dv, db = _dev(r), _dev(v), _dev(buf)
        check(lib().tbsla_hip_AAxpAx(self.h, dr.ptr, dv.ptr, db.ptr, vect_incr, None))
        return dr.to_numpy(np.float64, len(r)), db.to_numpy(np.float64, len(buf))

    def AAxpAxpx(self, r, v, buf, vect_incr: int = 0):
        dr, dv, db = _dev(r), _dev(v), _dev(buf)
        check(lib().tbsla_hip_AAxpAxpx(self.h, dr.ptr, dv.ptr, db.ptr, vect_incr, None))
        return dr.to_numpy(np.float64, len(r)), db.to_numpy(np.float64, len(buf))

    def spmv_device(self, dr: DeviceBuffer, dv: DeviceBuffer, vect_incr: int = 0, stream=None):
        check(lib().tbsla_hip_spmv(self.h, dr.ptr, dv.ptr, vect_incr, stream))

    def time_spmv(self, dr: DeviceBuffer, dv: DeviceBuffer, iters: int, stream=None) -> float:
        ms = C.c_float()
        check(lib().tbsla_hip_time_spmv(self.h, dr.ptr, dv.ptr, iters, C.byref(ms), stream))
        return ms.value

    def device_arrays(self):
        a, b, v = C.c_void_p(), C.c_void_p(), C.c_void_p()
        check(lib().tbsla_hip_device_arrays(self.h, C.byref(a), C.byref(b), C.byref(v)))
        return a.value, b.value, v.value

    def write(self) -> bytes:
        """Matrix::write: the block's binary image (the reference's bytes)."""
        n = C.c_size_t()
        check(lib().tbsla_hip_write(self.h, None, 0, C.byref(n), None))
        buf = C.create_string_buffer(max(n.value, 1))
        check(lib().tbsla_hip_write(self.h, buf, n.value, C.byref(n), None))
        return buf.raw[: n.value]

    @classmethod
    def read(cls, fmt: int, image: bytes) -> "Matrix":
        """Matrix::read: a block of format fmt from a binary image."""
        m = cls(fmt)
        check(lib().tbsla_hip_read(m.h, image, len(image), None))
        return m

    def values_changed(self, stream=None) -> None:
        """tbsla_hip_values_changed: after writes through device_arrays."""
        check(lib().tbsla_hip_values_changed(self.h, stream))

    def set_csr_schedule(self, lanes: int, rows: int = 0) -> None:
        check(lib().tbsla_hip_set_csr_schedule(self.h, lanes, rows))

    def get_csr_schedule(self) -> tuple[int, int]:
        a, b = C.c_int(), C.c_int()
        check(lib().tbsla_hip_get_csr_schedule(self.h, C.byref(a), C.byref(b)))
        return a.value, b.value

    def set_csr_kernel(self, lanes: int, rows: int, vec: int = 1, nontemporal: int = 1) -> None:
        check(lib().tbsla_hip_set_csr_kernel(self.h, lanes, rows, vec, nontemporal))

    def get_csr_kernel(self) -> dict:
        a, b, c, d = C.c_int(), C.c_int(), C.c_int(), C.c_int()
        check(lib().tbsla_hip_get_csr_kernel(self.h, C.byref(a), C.byref(b), C.byref(c), C.byref(d)))
        return {"lanes": a.value, "rows": b.value, "vec": c.value, "nontemporal": d.value}

    def tune_layout(self, iters: int = 5) -> float:
        """The autotuner on a unit-value vector of this block's width: picks
        the SpMV kernel and execution layout (the CG engine then reads the
        diagonal layout when it holds every entry); returns us per SpMV."""
        g = self.geometry()
        # COO reads x by the global column and writes y by the global row
        nx = g.n_col if self.fmt == COO else g.ln_col
        dv = DeviceBuffer(max(nx, 1) * 8)
        dr = DeviceBuffer(max(g.n_row if self.fmt == COO else g.ln_row, 1) * 8)
        check(lib().tbsla_hip_fill_unit_vector(dv.ptr, nx, 1, None))
        return self.autotune(dr, dv, iters)

    def autotune(self, dr: DeviceBuffer, dv: DeviceBuffer, iters: int = 10,
                 stream=None) -> float:
        us = C.c_float()
        check(lib().tbsla_hip_autotune(self.h, dr.ptr, dv.ptr, iters, C.byref(us), stream))
        return us.value

    # -- column normalisation, PageRank, CG ----------------------------------
    def get_col_sums(self, s=None) -> np.ndarray:
        n = self.geometry().ln_col
        ds = _dev(np.zeros(n) if s is None else s)
        check(lib().tbsla_hip_get_col_sums(self.h, ds.ptr, None))
        return ds.to_numpy(np.float64, n)

    def normalize_cols(self, s) -> None:
        ds = _dev(s)
        check(lib().tbsla_hip_normalize_cols(self.h, ds.ptr, None))

    def make_stochastic(self, s=None) -> np.ndarray:
        n = self.geometry().ln_col
        ds = _dev(np.zeros(n) if s is None else s)
        check(lib().tbsla_hip_make_stochastic(self.h, ds.ptr, None))
        return ds.to_numpy(np.float64, n)

    # -- row operations ----------------------------------------------------
    def row_sums_len(self) -> int:
        """Length of the row-sum vector: ELL and SCOO index it by global row
        (MatrixELL.cpp:790, MatrixSCOO.cpp:703), CSR and DENSE by local row."""
        g = self.geometry()
        return g.n_row if self.fmt in (ELL, SCOO) else g.ln_row

    def get_row_sums(self, s=None) -> np.ndarray:
        """Matrix::get_row_sums; SCOO accumulates into ``s`` (zeros if None)."""
        n = self.row_sums_len()
        s = np.zeros(n) if s is None else np.asarray(s, np.float64)
        if s.size < n:
            raise ValueError(f"row sums need {n} values, got {s.size}")
        ds = _dev(s)
        check(lib().tbsla_hip_get_row_sums(self.h, ds.ptr, None))
        return ds.to_numpy(np.float64, s.size)

    def normalize_rows(self, s) -> None:
        s = np.asarray(s, np.float64)
        if self.fmt != COO and s.size < self.row_sums_len():
            raise ValueError(f"row sums need {self.row_sums_len()} values, got {s.size}")
        ds = _dev(s)
        check(lib().tbsla_hip_normalize_rows(self.h, ds.ptr, None))

    def get_row_max_abs(self) -> np.ndarray:
        """Matrix::get_row_max_abs (CSR; zeros from the no-op formats)."""
        n = self.geometry().ln_row
        dm = _dev(np.zeros(n))
        check(lib().tbsla_hip_get_row_max_abs(self.h, dm.ptr, None))
        return dm.to_numpy(np.float64, n)

    def apply_exponential(self, max_abs=None, base: int = -1) -> None:
        """Matrix::apply_exponential: base <= 0 uses exp; max_abs None: no shift."""
        if max_abs is None:
            check(lib().tbsla_hip_apply_exponential(self.h, None, base, None))
            return
        max_abs = np.asarray(max_abs, np.float64)
        if self.fmt == CSR and max_abs.size < self.geometry().ln_row:
            raise ValueError("max_abs needs ln_row values")
        dm = _dev(max_abs)
        check(lib().tbsla_hip_apply_exponential(self.h, dm.ptr, base, None))

    def softmax_rows(self, base: int = -1):
        """The attention step's fused row softmax (CSR); returns (max_abs, sums)."""
        n = self.geometry().ln_row
        dm, ds = DeviceBuffer(max(n, 1) * 8), DeviceBuffer(max(n, 1) * 8)
        check(lib().tbsla_hip_softmax_rows(self.h, base, dm.ptr, ds.ptr, None))
        return dm.to_numpy(np.float64, n), ds.to_numpy(np.float64, n)

    def page_rank(self, beta=0.85, epsilon=1e-10, max_iterations=10000):
        n = self.geometry().ln_col
        db = DeviceBuffer(max(n, 1) * 8)
        it, err = C.c_int(), C.c_double()
        check(lib().tbsla_hip_page_rank(self.h, beta, epsilon, max_iterations, db.ptr,
                                        C.byref(it), C.byref(err), None))
        return db.to_numpy(np.float64, n), it.value, err.value

    def page_rank_max_norm(self, beta=0.85, epsilon=1e-10, max_iterations=10000):
        """tbsla::mpi::MatrixCOO::page_rank's max-normalised iteration
        (src/mpi/matrix/MatrixCOO.cpp:108-175); returns (b, iterations, error)."""
        n = self.geometry().n_col
        db = DeviceBuffer(max(n, 1) * 8)
        it, err = C.c_int(), C.c_double()
        check(lib().tbsla_hip_page_rank_max_norm(self.h, beta, epsilon, max_iterations, db.ptr,
                                                 C.byref(it), C.byref(err), None))
        return db.to_numpy(np.float64, n), it.value, err.value

    def page_rank_batched(self, row_offsets, beta=0.85, epsilon=1e-10, max_iterations=10000,
                          times=False):
        """PageRank of each diagonal block [row_offsets[k], row_offsets[k+1]);
        returns (b for all rows, iterations per instance, final errors), and
        the per-instance solve times in microseconds when `times`."""
        offs = np.ascontiguousarray(row_offsets, np.int32)
        k = len(offs) - 1
        n = self.geometry().ln_row
        db = DeviceBuffer(max(n, 1) * 8)
        its, errs, tus = np.zeros(k, np.int32), np.zeros(k), np.zeros(k)
        dp = C.POINTER(C.c_double)
        check(lib().tbsla_hip_page_rank_batched(
            self.h, k, offs.ctypes.data_as(_ip), beta, epsilon, max_iterations, db.ptr,
            its.ctypes.data_as(_ip), errs.ctypes.data_as(dp), tus.ctypes.data_as(dp), None))
        out = (db.to_numpy(np.float64, n), its, errs)
        return out + (tus,) if times else out

    def spmm(self, B) -> np.ndarray:
        """MatrixCSR::dense_multiply: C = A B (B row-major ln_col x k)."""
        B = np.ascontiguousarray(B, np.float64)
        g = self.geometry()
        k = B.shape[1]
        dB, dC = DeviceBuffer.from_numpy(B), DeviceBuffer(max(g.ln_row * k, 1) * 8)
        check(lib().tbsla_hip_spmm(self.h, dB.ptr, dC.ptr, k, None))
        return dC.to_numpy(np.float64, g.ln_row * k).reshape(g.ln_row, k)

    def personalized_page_rank(self, sources, beta=0.85, epsilon=1e-10, max_iterations=10000,
                               out: "DeviceBuffer | None" = None):
        """64 one-hot teleports at once; returns (X n x 64, iterations, max error)."""
        src = np.ascontiguousarray(sources, np.int32)
        assert len(src) == 64
        n = self.geometry().ln_col
        dX = out if out is not None else DeviceBuffer(n * 64 * 8)
        it, err = C.c_int(), C.c_double()
        check(lib().tbsla_hip_personalized_page_rank(
            self.h, beta, epsilon, max_iterations, src.ctypes.data_as(_ip), dX.ptr, C.byref(it),
            C.byref(err), None))
        X = None if out is not None else dX.to_numpy(np.float64, n * 64).reshape(n, 64)
        return X, it.value, err.value

    def set_csr_layout(self, layout=CSR_SPLIT, window=64, slices=8, stream=None):
        check(lib().tbsla_hip_set_csr_layout(self.h, layout, window, slices, stream))
        return self

    def get_csr_layout(self):
        lay, w, s, r = C.c_int(), C.c_int(), C.c_int(), C.c_longlong()
        check(lib().tbsla_hip_get_csr_layout(self.h, C.byref(lay), C.byref(w), C.byref(s),
                                             C.byref(r)))
        h = C.c_int()
        check(lib().tbsla_hip_get_csr_remote_form(self.h, C.byref(h)))
        return {"layout": lay.value, "window": w.value, "slices": s.value, "remote": r.value,
                "lds_block_rows": h.value}

    def cg_options(self, max_iterations=10000, tolerance=1e-10, relative=False,
                   precision=CG_FP64, preconditioner=PRECOND_NONE, replace_every=0,
                   use_initial_guess=False, near_zero=0.0, factor=None, factor_t=None,
                   graph_batch=0, lanes_per_row=0, stall_window=0,
                   factor_form=FACTOR_AUTO, max_blocks_per_system=0,
                   spmv_passes=0) -> CGOptions:
        return CGOptions(max_iterations, tolerance, int(relative), precision, preconditioner,
                         replace_every, int(use_initial_guess), near_zero,
                         factor.h if factor is not None else None,
                         factor_t.h if factor_t is not None else None, graph_batch,
                         lanes_per_row, stall_window, factor_form, max_blocks_per_system,
                         spmv_passes)

    def cg(self, rhs, max_iterations=10000, tolerance=1e-10, relative=False,
           precision=CG_FP64, preconditioner=PRECOND_NONE, replace_every=0, x0=None,
           near_zero=0.0, factor=None, factor_t=None, graph_batch=0, lanes_per_row=0,
           stall_window=0, factor_form=0, max_blocks_per_system=0, spmv_passes=0,
           stream=None):
        """Corrected tbsla::mpi::Matrix::CG; returns (x, iterations, |r|).
        rhs / x0: numpy or DeviceBuffer; M = G G^T with factor=G (a CSR Matrix)."""
        drhs = _dev(rhs)
        n = self.geometry().ln_col
        dx = _dev(np.zeros(n) if x0 is None else x0)
        o = self.cg_options(max_iterations, tolerance, relative, precision, preconditioner,
                            replace_every, x0 is not None, near_zero, factor, factor_t,
                            graph_batch, lanes_per_row, stall_window, factor_form,
                            max_blocks_per_system, spmv_passes)
        it, rn = C.c_int(), C.c_double()
        check(lib().tbsla_hip_cg(self.h, drhs.ptr, dx.ptr, C.byref(o), C.byref(it),
                                 C.byref(rn), stream))
        return dx.to_numpy(np.float64, n), it.value, rn.value

    def cg_device(self, drhs: "DeviceBuffer", dx: "DeviceBuffer", options: CGOptions,
                  stream=None):
        """CG on device vectors; returns (iterations, |r|)."""
        it, rn = C.c_int(), C.c_double()
        check(lib().tbsla_hip_cg(self.h, drhs.ptr, dx.ptr, C.byref(options), C.byref(it),
                                 C.byref(rn), stream))
        return it.value, rn.value

    def cg_batched(self, row_offsets, rhs, options: CGOptions, active=None, x0=None,
                   stream=None):
        """Independent CG solves on the diagonal blocks [row_offsets[k], row_offsets[k+1]);
        returns (x, list of CGResult)."""
        offs = np.ascontiguousarray(row_offsets, np.int32)
        k = len(offs) - 1
        n = int(offs[-1])
        drhs = _dev(rhs)
        dx = _dev(np.zeros(n) if x0 is None else x0)
        act = None if active is None else np.ascontiguousarray(active, np.int32)
        res = (CGResult * k)()
        check(lib().tbsla_hip_cg_batched(
            self.h, k, offs.ctypes.data_as(_ip), None if act is None else act.ctypes.data_as(_ip),
            drhs.ptr, dx.ptr, C.byref(options), res, stream))
        return dx.to_numpy(np.float64, n), list(res)

    # -- conversions --------------------------------------------------------
    def to(self, fmt: int) -> "Matrix":
        """Conversion constructor <fmt>(const MatrixCOO&); self must be COO."""
        m = Matrix(fmt)
        check(lib().tbsla_hip_convert(m.h, fmt, self.h, None))
        return m

    @classmethod
    def read_mm(cls, path) -> "Matrix":
        """MatrixCOO::readMM: a Matrix Market coordinate file into a COO block."""
        m = cls(COO)
        check(lib().tbsla_hip_read_mm(m.h, os.fsencode(path), None))
        return m

    @classmethod
    def from_arrays(cls, fmt: int, geom: dict, a=None, b=None, v=None, stored=None) -> "Matrix":
        """tbsla_hip_import: arrays in the reference layout."""
        g = Geometry()
        for f in GEOM_FIELDS:
            setattr(g, f, int(geom[f]))
        g.nnz = int(geom["nnz"])
        g.max_col = int(geom.get("max_col", 0))
        g.format = fmt
        keep = []

        def ptr(x, dt):
            if x is None:
                return None
            x = np.ascontiguousarray(x, dt)
            keep.append(x)
            return x.ctypes.data
        if stored is None:
            stored = len(v) if v is not None and fmt in (COO, SCOO) else 0
        g.stored = int(stored)
        m = cls(fmt)
        check(lib().tbsla_hip_import(m.h, fmt, C.byref(g), ptr(a, np.int32), ptr(b, np.int32),
                                     ptr(v, np.float64), None))
        return m


class PrecondNet:
    """The learned factored preconditioner (tbsla_hip_precond_net_*): maps an
    SPD CSR Matrix A to G (lower, tril(A)'s pattern) for CG with M = G G^T."""

    def __init__(self, seed: int = 0):
        h = C.c_void_p()
        check(lib().tbsla_hip_precond_net_create(C.byref(h), seed))
        self.h = h.value

    def __del__(self):
        try:
            if self.h:
                lib().tbsla_hip_precond_net_destroy(self.h)
        except Exception:
            pass
        self.h = None

    @staticmethod
    def param_count() -> int:
        return lib().tbsla_hip_precond_net_param_count()

    def get_params(self) -> np.ndarray:
        out = np.zeros(self.param_count(), np.float32)
        check(lib().tbsla_hip_precond_net_get_params(self.h, out.ctypes.data_as(C.POINTER(C.c_float))))
        return out

    def set_params(self, theta) -> None:
        t = np.ascontiguousarray(theta, np.float32)
        assert t.size == self.param_count()
        check(lib().tbsla_hip_precond_net_set_params(self.h, t.ctypes.data_as(C.POINTER(C.c_float))))

    def apply(self, A: Matrix, with_transpose: bool = True, stream=None):
        """(G, G^T or None) as CSR Matrices."""
        G = Matrix(CSR)
        GT = Matrix(CSR) if with_transpose else None
        check(lib().tbsla_hip_precond_net_apply(self.h, A.h, G.h, GT.h if GT else None, stream))
        return G, GT

    def loss_grad(self, A: Matrix, probes: int, seed: int, stream=None):
        loss = C.c_double()
        g = np.zeros(self.param_count(), np.float32)
        check(lib().tbsla_hip_precond_net_loss_grad(self.h, A.h, probes, seed, C.byref(loss),
                                                    g.ctypes.data_as(C.POINTER(C.c_float)), stream))
        return loss.value, g

    def train_step(self, A: Matrix, probes: int, seed: int, lr: float, clip: float = 0.0,
                   stream=None) -> float:
        loss = C.c_double()
        check(lib().tbsla_hip_precond_net_train_step(self.h, A.h, probes, seed, lr, clip,
                                                     C.byref(loss), stream))
        return loss.value

