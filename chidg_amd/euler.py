"""First-Newton-step DG Euler systems on the periodic uniform hex box [0,1]^3
(``cdg_euler_*`` in include/chidg_hip.h; host code in csrc/euler_jac.cpp).

The system is what ChiDG assembles for the first Newton step of a steady
Euler case (``mod_spatial.f90:40-197``, pseudo-transient term
``type_newton.f90:320-351``): blocks dR/dq + mass/dtau in the hex-box pattern
(``pattern.hex_box``, all directions periodic), the residual R(q), and the
state q (freestream plus a seeded perturbation of every modal coefficient).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from . import pattern as pt


def _check(code: int, what: str):
    if code != _lib.CDG_OK:
        raise _lib.CdgError(code, what, "")


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


def state(N: int, order: int, mach: float = 0.3, seed: int = 1, perturb: float = 0.01) -> np.ndarray:
    """Freestream rho = 1, p = 1/gamma, u = (mach, 0, 0), plus perturb * N(0,1)
    on every modal coefficient."""
    n = pt.block_size(order)
    q = np.empty(N ** 3 * n)
    _check(_lib.load().cdg_euler_state(N, N, N, order, mach, seed, perturb, _ptr(q)), "cdg_euler_state")
    return q


def residual(N: int, order: int, q: np.ndarray, nthreads: int = 0) -> np.ndarray:
    q = np.ascontiguousarray(q, dtype=np.float64)
    R = np.empty_like(q)
    _check(_lib.load().cdg_euler_residual(N, N, N, order, _ptr(q), _ptr(R), nthreads), "cdg_euler_residual")
    return R


def jacobian(N: int, order: int, q: np.ndarray, cfl: float = 1.0, nthreads: int = 0):
    """(pattern, blocks, R): blocks = dR/dq + mass/dtau (cfl > 0), column-major
    per block in the pattern's order."""
    n = pt.block_size(order)
    p = pt.hex_box(N, N, N, n)
    q = np.ascontiguousarray(q, dtype=np.float64)
    blocks = np.empty(p.nnzb * n * n)
    R = np.empty_like(q)
    _check(_lib.load().cdg_euler_jacobian(N, N, N, order, _ptr(q), cfl, _ptr(blocks), _ptr(R), nthreads),
           "cdg_euler_jacobian")
    return p, blocks, R


def system(N: int, order: int, mach: float = 0.3, seed: int = 1, perturb: float = 0.01, cfl: float = 1.0,
           nthreads: int = 0):
    """(pattern, blocks, R, q) of the first Newton step at the perturbed freestream."""
    q = state(N, order, mach, seed, perturb)
    p, blocks, R = jacobian(N, order, q, cfl, nthreads)
    return p, blocks, R, q


def residual_gpu(ctx, N: int, order: int, q, out=None):
    """R(q) on the context's GPU (``cdg_euler_residual_gpu``): q a host array or
    a device tensor; R of the same kind (``out`` if given)."""
    from .linsolve import _after_torch, _before_torch, _empty_like, _ptr as _vptr
    R = out if out is not None else _empty_like(q, len(q))
    _after_torch(ctx, q, R)
    ctx._check(ctx.lib.cdg_euler_residual_gpu(ctx.h, N, N, N, order, _vptr(q), _vptr(R)), "cdg_euler_residual_gpu")
    _before_torch(ctx)
    return R
