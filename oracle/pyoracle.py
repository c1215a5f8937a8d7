"""ctypes wrapper of the CPU oracle (``oracle/liboracle.so``).

TEST INFRASTRUCTURE ONLY: imported by tests/, the smoke check and the bench's
cpu_baseline leg, never by the product package.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

ORC_PREC_IDENTITY, ORC_PREC_ILU0 = 0, 1
ORC_CGS, ORC_MGS = 0, 1
ERRORS = {0: "ok", 1: "bad argument", 2: "singular block", 3: "no transposed block",
          4: "allocation failed", 5: "non-finite result"}


class OracleError(RuntimeError):
    def __init__(self, code, what):
        super().__init__(f"oracle {what}: {ERRORS.get(code, code)}")
        self.code = code


class _Pattern(C.Structure):
    _fields_ = [("nrows", C.c_int), ("n", C.c_int), ("nnzb", C.c_int),
                ("row_ptr", C.POINTER(C.c_int32)), ("col", C.POINTER(C.c_int32)),
                ("kind", C.POINTER(C.c_int32)), ("part", C.POINTER(C.c_int32))]


class _Params(C.Structure):
    _fields_ = [("nkrylov", C.c_int), ("tol", C.c_double), ("rtol", C.c_double),
                ("nmax", C.c_int), ("orthogonalization", C.c_int), ("inner_fgmres", C.c_int),
                ("inner_nkrylov", C.c_int), ("inner_tol", C.c_double), ("inner_rtol", C.c_double),
                ("inner_nmax", C.c_int), ("inner_orthogonalization", C.c_int)]


class _Stats(C.Structure):
    _fields_ = [("niter", C.c_int), ("residual_norm", C.c_double), ("nrestart", C.c_int),
                ("nreorth", C.c_int)]


def build() -> None:
    subprocess.run(["make", "-s", "-C", HERE], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = C.CDLL(LIB_PATH)
        d = C.POINTER(C.c_double)
        _lib.orc_mv.argtypes = [C.POINTER(_Pattern), d, d, d]
        _lib.orc_inv.argtypes = [C.c_int, d, d]
        _lib.orc_ilu0_factor.argtypes = [C.POINTER(_Pattern), d, d, d, C.POINTER(C.c_int)]
        _lib.orc_ilu0_apply.argtypes = [C.POINTER(_Pattern), d, d, d, d, d]
        _lib.orc_fgmres_solve.argtypes = [C.POINTER(_Pattern), d, C.c_int, d, d, d, d,
                                          C.POINTER(_Params), C.POINTER(_Stats)]
        _lib.orc_fill_synthetic.argtypes = [C.POINTER(_Pattern), C.c_uint64, C.c_double, C.c_double, d]
        _lib.orc_fill_vector.argtypes = [C.c_int64, C.c_uint64, C.c_double, d]
        _lib.orc_synth_value.argtypes = [C.c_uint64, C.c_int64, C.c_int64]
        _lib.orc_synth_value.restype = C.c_double
        _lib.orc_set_threads.argtypes = [C.c_int]
        _lib.orc_norm.argtypes = [C.c_int64, d]
        _lib.orc_norm.restype = C.c_double
    return _lib


def set_threads(n: int) -> None:
    lib().orc_set_threads(int(n))


def _dp(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(C.POINTER(C.c_double))


class OraclePattern:
    """Keeps the int32 arrays alive for the C struct.  ``part`` (rank of each
    row, or None) restricts ILU(0) to blocks whose parent row has the row's
    rank."""

    def __init__(self, pattern, part=None):
        self.row_ptr = np.ascontiguousarray(pattern.row_ptr, dtype=np.int32)
        self.col = np.ascontiguousarray(pattern.col, dtype=np.int32)
        self.kind = np.ascontiguousarray(pattern.kind, dtype=np.int32)
        self.nrows, self.n, self.nnzb = pattern.nrows, pattern.n, int(self.row_ptr[-1])
        ip = C.POINTER(C.c_int32)
        self.part = None if part is None else np.ascontiguousarray(part, dtype=np.int32)
        pp = self.part.ctypes.data_as(ip) if self.part is not None else ip()
        self.s = _Pattern(self.nrows, self.n, self.nnzb, self.row_ptr.ctypes.data_as(ip),
                          self.col.ctypes.data_as(ip), self.kind.ctypes.data_as(ip), pp)

    @property
    def ptr(self):
        return C.byref(self.s)


def fill_synthetic(pattern, seed: int, s_diag: float, s_off: float) -> np.ndarray:
    op = OraclePattern(pattern)
    blocks = np.empty(op.nnzb * op.n * op.n, dtype=np.float64)
    lib().orc_fill_synthetic(op.ptr, C.c_uint64(seed), s_diag, s_off, _dp(blocks))
    return blocks


def fill_vector(length: int, seed: int, scale: float = 1.0) -> np.ndarray:
    v = np.empty(length, dtype=np.float64)
    lib().orc_fill_vector(length, C.c_uint64(seed), scale, _dp(v))
    return v


def mv(pattern, blocks: np.ndarray, x: np.ndarray) -> np.ndarray:
    op = OraclePattern(pattern)
    y = np.empty(op.nrows * op.n, dtype=np.float64)
    e = lib().orc_mv(op.ptr, _dp(blocks), _dp(np.ascontiguousarray(x, dtype=np.float64)), _dp(y))
    if e:
        raise OracleError(e, "mv")
    return y


def inv(a: np.ndarray) -> np.ndarray:
    """Inverse of a square matrix given in numpy (row-major) form."""
    n = a.shape[0]
    colmajor = np.ascontiguousarray(a.T, dtype=np.float64)
    out = np.empty_like(colmajor)
    e = lib().orc_inv(n, _dp(colmajor), _dp(out))
    if e:
        raise OracleError(e, "inv")
    return out.T.copy()


def ilu0_factor(pattern, blocks: np.ndarray, part=None):
    op = OraclePattern(pattern, part)
    nn = op.n * op.n
    L = np.zeros(op.nnzb * nn, dtype=np.float64)
    Dinv = np.zeros(op.nrows * nn, dtype=np.float64)
    bad = C.c_int(-1)
    e = lib().orc_ilu0_factor(op.ptr, _dp(blocks), _dp(L), _dp(Dinv), C.byref(bad))
    if e:
        err = OracleError(e, "ilu0_factor")
        err.row = bad.value
        raise err
    return L, Dinv


def ilu0_apply(pattern, blocks, L, Dinv, v, part=None) -> np.ndarray:
    op = OraclePattern(pattern, part)
    z = np.empty(op.nrows * op.n, dtype=np.float64)
    e = lib().orc_ilu0_apply(op.ptr, _dp(blocks), _dp(L), _dp(Dinv),
                             _dp(np.ascontiguousarray(v, dtype=np.float64)), _dp(z))
    if e:
        raise OracleError(e, "ilu0_apply")
    return z


def fgmres(pattern, blocks, b, precon: str = "ilu0", factors=None, nkrylov=30, tol=1e-8,
           rtol=1e-8, nmax=-1, orthogonalization="CGS", inner_fgmres=False, inner_nkrylov=100,
           inner_tol=1e-1, inner_rtol=1e-1, inner_nmax=100, inner_orthogonalization="CGS", part=None):
    """Returns (x, niter, residual_norm, nreorth).  ``part``: ranks of the
    rows for the ILU(0) (the matvec uses every block)."""
    op = OraclePattern(pattern, part)
    if precon == "ilu0":
        L, Dinv = factors if factors is not None else ilu0_factor(pattern, blocks, part)
        pk = ORC_PREC_ILU0
    else:
        L = Dinv = np.zeros(1)
        pk = ORC_PREC_IDENTITY
    prm = _Params(nkrylov, tol, rtol, nmax, ORC_MGS if orthogonalization == "MGS" else ORC_CGS,
                  int(inner_fgmres), inner_nkrylov, inner_tol, inner_rtol, inner_nmax,
                  ORC_MGS if inner_orthogonalization == "MGS" else ORC_CGS)
    st = _Stats()
    x = np.empty(op.nrows * op.n, dtype=np.float64)
    e = lib().orc_fgmres_solve(op.ptr, _dp(blocks), pk, _dp(L), _dp(Dinv),
                               _dp(np.ascontiguousarray(b, dtype=np.float64)), _dp(x),
                               C.byref(prm), C.byref(st))
    if e:
        raise OracleError(e, "fgmres")
    return x, st.niter, st.residual_norm, st.nreorth


def synth_value(seed: int, blk: int, elem: int) -> float:
    return lib().orc_synth_value(C.c_uint64(seed), blk, elem)
