"""Host-side mirror of ChiDG's linear-solve interfaces over the gfx950 C ABI.

Names and semantics follow the reference so that calling code reads like
ChiDG's own:

* ``chidg_matrix`` -- ``chidg_matrix_t`` (``src/containers/type_chidg_matrix.f90:38-83``)
  with backend ``'hip'``: ``init`` from a block pattern, ``store``/``assemble``
  uploads the column-major blocks and stamps the matrix.
* ``chidg_mv(A, x)`` -- ``operator_chidg_mv.f90:46-61``.
* ``precon_ILU0`` / ``precon_identity`` -- ``init``/``update(A, b)``/``apply(A, v)``/
  ``tear_down`` (``src/preconditioners/precon_ILU0.f90:64-470``).
* ``fgmres`` -- ``fgmres_t`` with ``solve(A, x, b, M)``; ``niter`` and
  ``residual_norm`` as in ``type_fgmres.f90:81-310,512-542``.

Vectors are numpy float64 arrays (host) or torch float64 tensors on the GPU
(device-resident); results come back in the kind the inputs were given in.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from .pattern import Pattern

_PRECISION = {"fp64": _lib.FP64, "fp32": _lib.FP32, "bf16": _lib.BF16}


def _is_torch(x) -> bool:
    return type(x).__module__.startswith("torch")


def _ptr(x) -> int:
    if _is_torch(x):
        if x.dtype.itemsize != 8 or not x.is_contiguous():
            raise TypeError("device vectors must be contiguous float64 tensors")
        return x.data_ptr()
    if x.dtype != np.float64 or not x.flags.c_contiguous:
        raise TypeError("host vectors must be contiguous float64 numpy arrays")
    return x.ctypes.data


def _empty_like(x, n):
    if _is_torch(x):
        import torch
        return torch.empty(n, dtype=torch.float64, device=x.device)
    return np.empty(n, dtype=np.float64)


def _on_gpu(*xs) -> bool:
    return any(x is not None and _is_torch(x) and x.is_cuda for x in xs)


def _torch_stream():
    import torch
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def _after_torch(ctx, *xs) -> None:
    """The library works on its own stream: order its next work after what
    torch's current stream has queued (the producers of device inputs, and the
    last users of memory the caching allocator hands out as outputs)."""
    if _on_gpu(*xs):
        ctx._check(ctx.lib.cdg_ctx_wait_stream(ctx.h, _torch_stream()), "cdg_ctx_wait_stream")


def _before_torch(ctx) -> None:
    """Order torch's later work after the library's queued work (outputs of
    the enqueue-only calls)."""
    ctx._check(ctx.lib.cdg_ctx_signal_stream(ctx.h, _torch_stream()), "cdg_ctx_signal_stream")


class Context:
    """One HIP device + stream (``cdg_ctx``)."""

    def __init__(self, device: int = 0):
        self.lib = _lib.load()
        h = C.c_void_p()
        self._check(self.lib.cdg_ctx_create(device, C.byref(h)), "cdg_ctx_create", None)
        self.h = h
        self.device = device

    def _check(self, code, what, h=True):
        if code != _lib.CDG_OK:
            msg = ""
            if h is not None and getattr(self, "h", None) is not None:
                msg = (self.lib.cdg_last_error(self.h) or b"").decode()
            raise _lib.CdgError(code, what, msg)

    def mem_used(self) -> int:
        v = C.c_int64()
        self._check(self.lib.cdg_ctx_mem_used(self.h, C.byref(v)), "cdg_ctx_mem_used")
        return v.value

    def stream_ptr(self) -> int:
        """The library's hipStream_t (for torch.cuda.ExternalStream)."""
        v = C.c_void_p()
        self._check(self.lib.cdg_ctx_stream(self.h, C.byref(v)), "cdg_ctx_stream")
        return v.value or 0

    def close(self):
        if getattr(self, "h", None):
            self.lib.cdg_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_default_ctx = None


def default_context() -> Context:
    global _default_ctx
    if _default_ctx is None:
        _default_ctx = Context(0)
    return _default_ctx


class chidg_matrix:
    """Block-sparse Jacobian resident in HBM (``chidg_matrix_t`` with backend 'hip')."""

    def __init__(self, pattern: Pattern, ctx: Context | None = None):
        self.ctx = ctx or default_context()
        self.pattern = pattern
        self.n, self.nrows, self.nnzb = pattern.n, pattern.nrows, pattern.nnzb
        rp = np.ascontiguousarray(pattern.row_ptr, dtype=np.int32)
        co = np.ascontiguousarray(pattern.col, dtype=np.int32)
        ki = np.ascontiguousarray(pattern.kind, dtype=np.int32)
        ip = C.POINTER(C.c_int32)
        h = C.c_void_p()
        self.ctx._check(self.ctx.lib.cdg_mat_create(self.ctx.h, self.nrows, self.n, rp.ctypes.data_as(ip),
                                                    co.ctypes.data_as(ip), ki.ctypes.data_as(ip), C.byref(h)),
                        "cdg_mat_create")
        self.h = h

    @property
    def vec_len(self) -> int:
        return self.nrows * self.n

    def store(self, blocks) -> None:
        """Upload all blocks (pattern order, each n*n column-major) and stamp."""
        if _is_torch(blocks):
            if blocks.numel() != self.nnzb * self.n * self.n:
                raise ValueError("block array has the wrong size")
        else:
            blocks = np.ascontiguousarray(blocks, dtype=np.float64)
            if blocks.size != self.nnzb * self.n * self.n:
                raise ValueError("block array has the wrong size")
        _after_torch(self.ctx, blocks)
        self.ctx._check(self.ctx.lib.cdg_mat_upload(self.ctx.h, self.h, _ptr(blocks)), "cdg_mat_upload")

    assemble = store

    def set_partition(self, part) -> None:
        """Rank of each block row (None: one rank).  ILU(0) then uses only the
        blocks whose parent row is on the row's rank (the reference's
        parent_proc == IRANK test, precon_ILU0.f90:180,222,344,368):
        block-Jacobi ILU(0) over the ranks.  Preconditioners created
        afterwards follow it; the matvec uses every block."""
        if part is None:
            code = self.ctx.lib.cdg_mat_set_partition(self.ctx.h, self.h, None)
        else:
            arr = np.ascontiguousarray(part, dtype=np.int32)
            if arr.shape != (self.nrows,):
                raise ValueError("part needs one rank per block row")
            code = self.ctx.lib.cdg_mat_set_partition(self.ctx.h, self.h, arr.ctypes.data_as(C.POINTER(C.c_int32)))
        self.ctx._check(code, "cdg_mat_set_partition")
        self.part = None if part is None else np.asarray(part, dtype=np.int32).copy()

    def fill_synthetic(self, seed: int, s_diag: float, s_off: float) -> None:
        self.ctx._check(self.ctx.lib.cdg_mat_fill_synthetic(self.ctx.h, self.h, seed, s_diag, s_off),
                        "cdg_mat_fill_synthetic")

    def synthetic_vector(self, seed: int, scale: float = 1.0, like=None):
        v = _empty_like(like if like is not None else np.empty(0), self.vec_len)
        _after_torch(self.ctx, v)
        self.ctx._check(self.ctx.lib.cdg_vec_fill_synthetic(self.ctx.h, self.h, seed, scale, _ptr(v)),
                        "cdg_vec_fill_synthetic")
        return v

    def to_host(self) -> np.ndarray:
        out = np.empty(self.nnzb * self.n * self.n)
        self.ctx._check(self.ctx.lib.cdg_mat_download(self.ctx.h, self.h, _ptr(out)), "cdg_mat_download")
        return out

    @property
    def stamp(self) -> int:
        v = C.c_uint64()
        self.ctx._check(self.ctx.lib.cdg_mat_stamp(self.h, C.byref(v)), "cdg_mat_stamp")
        return v.value

    def tune(self, reps: int = 5) -> float:
        """Launch-shape autotuner of the matvec: keeps the fastest shape for
        this matrix and later ones of the same block size.  Returns its time
        per product in seconds."""
        ms = C.c_double(0.0)
        self.ctx._check(self.ctx.lib.cdg_mat_tune(self.ctx.h, self.h, int(reps), C.byref(ms)), "cdg_mat_tune")
        return ms.value * 1e-3

    def launch_shape(self) -> dict:
        out = (C.c_int32 * 2)()
        self.ctx._check(self.ctx.lib.cdg_mat_launch_shape(self.h, out), "cdg_mat_launch_shape")
        return dict(threads=out[0], loads_in_flight=out[1])

    def release(self):
        if getattr(self, "h", None):
            self.ctx.lib.cdg_mat_destroy(self.ctx.h, self.h)
            self.h = None

    def __del__(self):
        try:
            self.release()
        except Exception:
            pass


def chidg_mv(A: chidg_matrix, x, out=None):
    """y = A x over every block class (operator_chidg_mv.f90:46-61)."""
    y = out if out is not None else _empty_like(x, A.vec_len)
    _after_torch(A.ctx, x, y)
    A.ctx._check(A.ctx.lib.cdg_mv(A.ctx.h, A.h, _ptr(x), _ptr(y)), "cdg_mv")
    return y


def chidg_mv_async(A: chidg_matrix, x, y, order: bool = True) -> None:
    """Enqueue y = A x on the library stream (device tensors only).  With
    ``order`` (the default) the product waits for torch's queued work and
    torch's later work waits for the product; a caller that has synchronized
    the device already (a timing loop) may pass False."""
    if order:
        _after_torch(A.ctx, x, y)
    A.ctx._check(A.ctx.lib.cdg_mv_async(A.ctx.h, A.h, _ptr(x), _ptr(y)), "cdg_mv_async")
    if order:
        _before_torch(A.ctx)


class _Preconditioner:
    kind = _lib.PREC_IDENTITY

    def __init__(self, precision: str = "fp64", schedule: str | None = None):
        self.precision = precision
        self.schedule = schedule
        self.h = None
        self.ctx = None
        self.initialized = False
        self.stamp = 0

    def init(self, A: chidg_matrix):
        self.tear_down()
        self.ctx = A.ctx
        h = C.c_void_p()
        self.ctx._check(self.ctx.lib.cdg_prec_create(self.ctx.h, A.h, self.kind, _PRECISION[self.precision],
                                                     C.byref(h)), "cdg_prec_create")
        self.h = h
        self.initialized = True
        if self.schedule is not None:
            self.set_schedule(self.schedule)
        return self

    def set_schedule(self, schedule: str):
        """ILU(0) apply schedule: 'levels' (one launch per wavefront level
        and stage), 'dataflow' (one persistent launch over column slices with
        partial sums), 'folded' (levels, with Dinv folded into the upper blocks
        at the first apply after each factor; FP64/FP32), 'stream' (one
        persistent launch over row slices of the factor, W = Dinv U formed at
        the first apply after each factor), 'rowslice' (the same row-slice
        tasks, one launch per wavefront level) or 'auto' (the default:
        'levels' for the first 12 applies after each factor, 'rowslice' from
        then on, or from the first apply when the previous factor served more
        than 12)."""
        code = {"levels": 0, "dataflow": 1, "folded": 2, "auto": 3, "stream": 4, "rowslice": 5}[schedule]
        self.ctx._check(self.ctx.lib.cdg_prec_set_schedule(self.ctx.h, self.h, code), "cdg_prec_set_schedule")
        self.schedule = schedule

    def tune(self, A: chidg_matrix, reps: int = 5) -> float:
        """Launch-shape autotuner of the apply on the current factor (after
        update): keeps the fastest shape for this preconditioner and for later
        ones of the same block size and precision.  Returns its time per
        apply in seconds."""
        ms = C.c_double(0.0)
        self.ctx._check(self.ctx.lib.cdg_prec_tune(self.ctx.h, self.h, A.h, int(reps), C.byref(ms)), "cdg_prec_tune")
        return ms.value * 1e-3

    def launch_shape(self) -> dict:
        out = (C.c_int32 * 4)()
        self.ctx._check(self.ctx.lib.cdg_prec_launch_shape(self.h, out), "cdg_prec_launch_shape")
        return dict(slices=out[0], slice_width=out[1], threads=out[2], loads_in_flight=out[3])

    def update(self, A: chidg_matrix, b=None):
        if not self.initialized:
            self.init(A)
        bad = C.c_int32(-1)
        code = self.ctx.lib.cdg_prec_update(self.ctx.h, self.h, A.h, C.byref(bad))
        if code != _lib.CDG_OK:
            err = _lib.CdgError(code, "cdg_prec_update", (self.ctx.lib.cdg_last_error(self.ctx.h) or b"").decode())
            err.row = bad.value
            raise err
        self.stamp = A.stamp

    def apply(self, A: chidg_matrix, v, out=None):
        z = out if out is not None else _empty_like(v, A.vec_len)
        _after_torch(self.ctx, v, z)
        self.ctx._check(self.ctx.lib.cdg_prec_apply(self.ctx.h, self.h, A.h, _ptr(v), _ptr(z)), "cdg_prec_apply")
        return z

    def apply_async(self, A: chidg_matrix, v, z, order: bool = True) -> None:
        """Enqueue z = M^-1 v on the library stream (device tensors only);
        ``order`` as in ``chidg_mv_async``."""
        if order:
            _after_torch(self.ctx, v, z)
        self.ctx._check(self.ctx.lib.cdg_prec_apply_async(self.ctx.h, self.h, A.h, _ptr(v), _ptr(z)),
                        "cdg_prec_apply_async")
        if order:
            _before_torch(self.ctx)

    def factors(self, A: chidg_matrix):
        """(L, Dinv) as FP64 host arrays, L in the matrix's block slots."""
        L = np.zeros(A.nnzb * A.n * A.n)
        D = np.zeros(A.nrows * A.n * A.n)
        self.ctx._check(self.ctx.lib.cdg_prec_download(self.ctx.h, self.h, _ptr(L), _ptr(D)), "cdg_prec_download")
        return L, D

    def tear_down(self):
        if self.h is not None and self.ctx is not None:
            self.ctx.lib.cdg_prec_destroy(self.ctx.h, self.h)
        self.h = None
        self.initialized = False
        self.stamp = 0

    def __del__(self):
        try:
            self.tear_down()
        except Exception:
            pass


class precon_identity(_Preconditioner):
    kind = _lib.PREC_IDENTITY


class precon_ILU0(_Preconditioner):
    """Block ILU(0): ``precision`` 'fp64' (reference), 'fp32', or 'bf16'."""
    kind = _lib.PREC_ILU0


class precon_jacobi(_Preconditioner):
    """Block Jacobi (``precon_jacobi_t``, precon_jacobi.f90:95-215):
    ``update`` inverts every diagonal block, ``apply`` is z_i = Dinv_i v_i."""
    kind = _lib.PREC_JACOBI


def create_preconditioner(name: str, precision: str = "fp64") -> _Preconditioner:
    """Factory with the reference's names (mod_preconditioner.f90:49-108)."""
    key = name.strip().lower()
    if key == "ilu0":
        return precon_ILU0(precision)
    if key in ("jacobi", "blockjacobi"):
        return precon_jacobi(precision)
    if key == "identity":
        return precon_identity(precision)
    raise ValueError(f"unknown preconditioner {name!r}")


class fgmres:
    """``fgmres_t``: restarted flexible GMRES.  Defaults are the reference's
    (type_linear_solver.f90:21-40)."""

    def __init__(self, nkrylov=2000, tol=1e-8, rtol=1e-10, nmax=-1, orthogonalization="CGS",
                 inner_fgmres=True, inner_nkrylov=100, inner_tol=1e-1, inner_rtol=1e-1, inner_nmax=100,
                 inner_orthogonalization="CGS"):
        self.nkrylov, self.tol, self.rtol, self.nmax = nkrylov, tol, rtol, nmax
        self.orthogonalization = orthogonalization
        self.inner_fgmres, self.inner_nkrylov = inner_fgmres, inner_nkrylov
        self.inner_tol, self.inner_rtol, self.inner_nmax = inner_tol, inner_rtol, inner_nmax
        self.inner_orthogonalization = inner_orthogonalization
        self.niter = 0
        self.residual_norm = float("nan")
        self.stats = None

    def set(self, options: dict):
        """linear_solver_t%set(options) (type_linear_solver.f90:113-131)."""
        for k, v in options.items():
            if not hasattr(self, k):
                raise KeyError(k)
            setattr(self, k, v)

    def params(self, update_preconditioner: bool = True) -> _lib.FgmresParams:
        o = {"CGS": _lib.CGS, "MGS": _lib.MGS}
        return _lib.FgmresParams(self.nkrylov, self.tol, self.rtol, self.nmax, o[self.orthogonalization],
                                 int(self.inner_fgmres), self.inner_nkrylov, self.inner_tol, self.inner_rtol,
                                 self.inner_nmax, o[self.inner_orthogonalization], int(update_preconditioner))

    def solve(self, A: chidg_matrix, x, b, M: _Preconditioner, solver_controller=None, comm=None):
        """Solve A x = b (x overwritten, zero initial guess).  Without a
        controller the preconditioner is updated every call, as in the
        reference; with one, when ``solver_controller.update_preconditioner``
        says so (type_fgmres.f90:118-122)."""
        if not M.initialized:
            M.init(A)
        update = True if solver_controller is None else solver_controller.update_preconditioner(A, M, comm)
        prm = self.params(update_preconditioner=update)
        st = _lib.FgmresStats()
        _after_torch(A.ctx, x, b)
        A.ctx._check(A.ctx.lib.cdg_fgmres_solve(A.ctx.h, A.h, M.h, _ptr(b), _ptr(x), C.byref(prm), C.byref(st)),
                     "cdg_fgmres_solve")
        self.niter = st.niter
        self.residual_norm = st.residual_norm
        self.stats = st
        if st.factored:
            M.stamp = A.stamp
        return x

    def solve_multi(self, A: chidg_matrix, X, B, M: _Preconditioner, solver_controller=None, comm=None):
        """k right-hand sides with one preconditioner (``cdg_fgmres_solve_multi``):
        B and X are (k, nrows*n); every row is the solve of that right-hand
        side alone, the Arnoldi steps share k-vector applies and matvecs.
        ``niter`` / ``residual_norm`` become lists, ``stats`` a list of k."""
        if not M.initialized:
            M.init(A)
        k = int(B.shape[0])
        assert tuple(X.shape) == tuple(B.shape) and B.shape[1] == A.vec_len
        update = True if solver_controller is None else solver_controller.update_preconditioner(A, M, comm)
        prm = self.params(update_preconditioner=update)
        st = (_lib.FgmresStats * k)()
        _after_torch(A.ctx, X, B)
        A.ctx._check(A.ctx.lib.cdg_fgmres_solve_multi(A.ctx.h, A.h, M.h, k, _ptr(B), _ptr(X), C.byref(prm), st),
                     "cdg_fgmres_solve_multi")
        self.stats = list(st)
        self.niter = [s.niter for s in st]
        self.residual_norm = [s.residual_norm for s in st]
        if st[0].factored:
            M.stamp = A.stamp
        return X


class solver_controller:
    """``solver_controller_t`` (type_solver_controller.f90:25-124): controls
    the update of the lhs and of the preconditioner inside a nonlinear solver.

    ``update_preconditioner(lhs, preconditioner, comm)`` asks for a refactor
    when the lhs changed since the last call (its stamp differs from the
    remembered ``lhs_stamp``), when ``force_update_preconditioner`` is set, or
    when the preconditioner was last updated with another version of the
    matrix; it then remembers the lhs stamp and clears the force flag.  The
    reference reduces the decision with a logical OR over the ranks of
    ``comm``; on one GPU (one rank) that reduction is the local decision."""

    def __init__(self):
        self.force_update_lhs = False
        self.lhs_updated = False
        self.force_update_preconditioner = False
        self.lhs_stamp = 0
        self.preconditioner_updated = False

    def update_lhs(self, A, niter: int = 0, residual_ratio: float = 0.0) -> bool:
        """Default: always update the lhs (type_solver_controller.f90:58-74)."""
        self.lhs_updated = True
        return True

    def update_preconditioner(self, lhs: chidg_matrix, preconditioner: _Preconditioner, comm=None) -> bool:
        stamp = lhs.stamp
        update = (self.lhs_stamp != stamp or self.force_update_preconditioner
                  or stamp != preconditioner.stamp)
        self.lhs_stamp = stamp
        self.force_update_preconditioner = False
        self.preconditioner_updated = bool(update)
        return self.preconditioner_updated


class batch:
    """Many independent systems solved together (``cdg_batch_*``): the
    systems of each block size joined in one block-diagonal matrix with one
    ILU(0), every right-hand side an FGMRES(nkrylov) run in lockstep with the
    others, each cycle of Arnoldi steps one captured HIP graph.  ``mats`` are
    stored ``chidg_matrix`` objects of one context; ``nrhs[s]`` right-hand
    sides (1..8) for system s."""

    def __init__(self, mats, nrhs, precision: str = "fp32", nkrylov: int = 30):
        self.ctx = mats[0].ctx
        self.mats = list(mats)
        self.nrhs = [int(k) for k in nrhs]
        handles = (C.c_void_p * len(mats))(*[A.h for A in mats])
        ks = (C.c_int32 * len(mats))(*self.nrhs)
        h = C.c_void_p()
        prec = {"fp64": _lib.FP64, "fp32": _lib.FP32, "bf16": _lib.BF16}[precision]
        self.ctx._check(self.ctx.lib.cdg_batch_create(self.ctx.h, len(mats), handles, ks, prec, nkrylov, C.byref(h)),
                        "cdg_batch_create")
        self.h = h
        self.nkrylov = nkrylov

    def solve(self, B, X, tol=0.0, rtol=1e-8, nmax=-1):
        """B[s], X[s]: (nrhs[s], nrows*n) arrays or device tensors; X overwritten.
        Returns one cdg_fgmres_stats per right-hand side, systems in order."""
        ls = fgmres(nkrylov=self.nkrylov, tol=tol, rtol=rtol, nmax=nmax, inner_fgmres=False)
        prm = ls.params(update_preconditioner=False)
        nl = sum(self.nrhs)
        st = (_lib.FgmresStats * nl)()
        _after_torch(self.ctx, *B, *X)
        bp = (C.c_void_p * len(B))(*[_ptr(b) for b in B])
        xp = (C.c_void_p * len(X))(*[_ptr(x) for x in X])
        self.ctx._check(self.ctx.lib.cdg_batch_solve(self.ctx.h, self.h, bp, xp, C.byref(prm), st), "cdg_batch_solve")
        _before_torch(self.ctx)
        return list(st)

    def update(self, mats=None):
        """New blocks for the same systems (``mats``: stored matrices with the
        patterns the batch was created from): copied into the joined matrices
        and refactored, nothing allocated.  ``mats`` None: the joined matrices
        refactored as they are."""
        handles = None
        if mats is not None:
            handles = (C.c_void_p * len(mats))(*[A.h for A in mats])
        self.ctx._check(self.ctx.lib.cdg_batch_update(self.ctx.h, self.h, handles), "cdg_batch_update")

    def info(self) -> dict:
        nl, ng, nv = C.c_int32(), C.c_int32(), C.c_int64()
        self.ctx._check(self.ctx.lib.cdg_batch_info(self.h, C.byref(nl), C.byref(ng), C.byref(nv)), "cdg_batch_info")
        return dict(lanes=nl.value, groups=ng.value, values=nv.value)

    def close(self):
        if getattr(self, "h", None):
            self.ctx.lib.cdg_batch_destroy(self.ctx.h, self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
