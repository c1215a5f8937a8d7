"""Request loop for parameter-sweep solves (``cdg_server_*`` in
include/chidg_hip.h): the Python side submits synthetic systems and collects
results; the server's worker threads, admission against the device-memory
budget and per-slot HIP streams are native.

``poisson_load`` drives a server with Poisson arrivals of requests drawn from
a size distribution and reports throughput and latency percentiles.
"""
from __future__ import annotations

import ctypes as C
import time
from dataclasses import dataclass

import numpy as np

from . import _lib
from .linsolve import _PRECISION, fgmres


@dataclass
class Request:
    """A synthetic system: hex box nx*ny*nz, degree ``order`` (P0..P3), periodic
    faces by bit (1: xi, 2: eta, 4: zeta); blocks 2I + G*s_diag on the diagonal
    and G*s_off elsewhere (seed_A), right-hand side G (seed_b)."""
    nx: int
    ny: int
    nz: int
    order: int
    seed_A: int
    seed_b: int = 1
    periodic: int = 7
    precision: str = "fp32"
    s_diag: float | None = None
    s_off: float | None = None
    nrhs: int = 1

    @property
    def n(self) -> int:
        return (self.order + 1) ** 3 * 5

    @property
    def nrows(self) -> int:
        return self.nx * self.ny * self.nz

    def c(self) -> _lib.Request:
        sd = self.s_diag if self.s_diag is not None else 1.0 / np.sqrt(self.n)
        so = self.s_off if self.s_off is not None else 0.1 / np.sqrt(self.n)
        return _lib.Request(self.nx, self.ny, self.nz, self.order, self.periodic, _PRECISION[self.precision],
                            self.seed_A, self.seed_b, sd, so, self.nrhs)


@dataclass
class Result:
    status: int
    niter: int
    residual_norm: float
    t_queue: float
    t_run: float
    t_latency: float
    t_factor: float
    slot: int
    factored: int = 0
    x: np.ndarray | None = None


class Server:
    """``nslots`` workers on GPU ``device``; ``hbm_budget`` bytes of device memory
    for running requests (None: 80% of the free memory); FGMRES parameters as
    in ``linsolve.fgmres``."""

    def __init__(self, nslots: int = 4, device: int = 0, hbm_budget: int | None = None, **fgmres_kw):
        self.lib = _lib.load()
        ls = fgmres(**fgmres_kw)
        self.params = ls.params()
        h = C.c_void_p()
        code = self.lib.cdg_server_create(device, nslots, int(hbm_budget or 0), C.byref(self.params), C.byref(h))
        if code != _lib.CDG_OK:
            raise _lib.CdgError(code, "cdg_server_create", "")
        self.h = h
        self._out = {}
        self._args = {}      # inputs of queued requests (kept alive until their wait)
        self._keep = {}      # matrices of registered systems

    def footprint(self, req: Request) -> int:
        b = C.c_int64()
        r = req.c()
        code = self.lib.cdg_server_footprint(C.byref(r), C.byref(self.params), C.byref(b))
        if code != _lib.CDG_OK:
            raise _lib.CdgError(code, "cdg_server_footprint", "")
        return b.value

    def submit(self, req: Request, want_x: bool = False) -> int:
        x = np.empty(req.nrows * req.n) if want_x else None
        t = C.c_uint64()
        r = req.c()
        xp = x.ctypes.data_as(C.c_void_p) if x is not None else None
        code = self.lib.cdg_server_submit(self.h, C.byref(r), xp, C.byref(t))
        if code != _lib.CDG_OK:
            raise _lib.CdgError(code, "cdg_server_submit", "")
        self._out[t.value] = x
        return t.value

    def wait(self, ticket: int) -> Result:
        r = _lib.Result()
        code = self.lib.cdg_server_wait(self.h, ticket, C.byref(r))
        if code != _lib.CDG_OK:
            raise _lib.CdgError(code, "cdg_server_wait", "")
        x = self._out.pop(ticket, None)
        self._args.pop(ticket, None)
        res = Result(r.status, r.niter, r.residual_norm, r.t_queue, r.t_run, r.t_latency, r.t_factor, r.slot,
                     r.factored, x)
        if res.status != _lib.CDG_OK:
            msg = (self.lib.cdg_server_last_error(self.h) or b"").decode()
            raise _lib.CdgError(res.status, "request", msg)
        return res

    def register(self, A, precision: str = "fp32") -> int:
        """A caller's matrix (``linsolve.chidg_matrix``, stored) as a resident
        system: its ILU(0) factor stays on the device between requests and is
        recomputed only when the matrix is stored again (a new stamp)."""
        sid = C.c_uint64()
        code = self.lib.cdg_server_register(self.h, A.h, _PRECISION[precision], C.byref(sid))
        if code != _lib.CDG_OK:
            raise _lib.CdgError(code, "cdg_server_register", "")
        self._keep.setdefault(sid.value, []).append(A)
        return sid.value

    def submit_system(self, system: int, b, want_x: bool = False) -> int:
        """Right-hand sides b (one vector, or nrhs of them as rows of a 2-D
        array; host float64) on a registered system."""
        b = np.ascontiguousarray(b, dtype=np.float64)
        nrhs = 1 if b.ndim == 1 else b.shape[0]
        x = np.empty_like(b) if want_x else None
        t = C.c_uint64()
        xp = x.ctypes.data_as(C.c_void_p) if x is not None else None
        code = self.lib.cdg_server_submit_system(self.h, system, b.ctypes.data_as(C.c_void_p), nrhs, xp, C.byref(t))
        if code != _lib.CDG_OK:
            raise _lib.CdgError(code, "cdg_server_submit_system", "")
        self._out[t.value] = x
        self._args[t.value] = b
        return t.value

    def unregister(self, system: int) -> None:
        code = self.lib.cdg_server_unregister(self.h, system)
        if code != _lib.CDG_OK:
            raise _lib.CdgError(code, "cdg_server_unregister", "")
        self._keep.pop(system, None)

    def info(self) -> dict:
        q, a, pk = C.c_int32(), C.c_int32(), C.c_int32()
        use, bud = C.c_int64(), C.c_int64()
        self.lib.cdg_server_info(self.h, C.byref(q), C.byref(a), C.byref(pk), C.byref(use), C.byref(bud))
        return dict(queued=q.value, active=a.value, peak_active=pk.value, in_use=use.value, budget=bud.value)

    def close(self):
        if getattr(self, "h", None) is not None:
            self.lib.cdg_server_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def sweep_requests(count: int, rng: np.random.Generator, meshes=(4, 8), orders=(1, 2, 3),
                   precision: str = "fp32", mach=(0.2, 0.6), nrhs=(1, 1)) -> list[Request]:
    """Parameter-sweep requests: cube meshes N^3 with N uniform in [meshes[0],
    meshes[1]], degree uniform over ``orders``, right-hand sides uniform in
    [nrhs[0], nrhs[1]]; the Mach number M ~ U(mach) scales the off-diagonal
    coupling, s_off = (M / 0.4) * 0.1 / sqrt(n)."""
    out = []
    for k in range(count):
        N = int(rng.integers(meshes[0], meshes[1] + 1))
        P = int(rng.choice(orders))
        M = float(rng.uniform(*mach))
        n = (P + 1) ** 3 * 5
        out.append(Request(N, N, N, P, seed_A=int(rng.integers(1, 2**62)), seed_b=int(rng.integers(1, 2**62)),
                           precision=precision, s_off=(M / 0.4) * 0.1 / np.sqrt(n),
                           nrhs=int(rng.integers(nrhs[0], nrhs[1] + 1))))
    return out


def batch(server: Server, reqs: list[Request]) -> dict:
    """Submit every request at once and wait for all (a batch of independent
    systems); throughput in right-hand sides solved per second."""
    t0 = time.perf_counter()
    tickets = [server.submit(r) for r in reqs]
    res = [server.wait(t) for t in tickets]
    wall = time.perf_counter() - t0
    nrhs = sum(max(1, r.nrhs) for r in reqs)
    return dict(systems=len(reqs), rhs=nrhs, wall_s=wall, rhs_per_s=nrhs / wall, systems_per_s=len(reqs) / wall,
                iterations=int(sum(r.niter for r in res)), max_residual=float(max(r.residual_norm for r in res)),
                peak_active=server.info()["peak_active"])


def joined(reqs: list[Request], device: int = 0, nkrylov: int = 30, tol: float = 0.0, rtol: float = 1e-8,
           nmax: int = -1, warm: bool = True) -> dict:
    """The same batch of independent systems as one joined solve
    (``linsolve.batch``, ``cdg_batch_*``): each request's matrix filled on
    the device from seed_A (as the request loop fills it), its right-hand
    sides from seeds seed_b + k, the systems of each block size joined and
    factored once, every right-hand side an FGMRES lane.  The whole job is
    timed: fill, join + factor (allocating the joined matrices, factors and
    Krylov space), solve; the requests' matrices are released once joined (the
    batch holds its copy: both at once would fill the device at this size).
    With ``warm``, the joined matrices refactored in place
    (``batch.update()``, nothing allocated) and solved again: reported under
    "warm", with no fill."""
    import torch
    from .linsolve import Context, batch as joined_batch, chidg_matrix
    from .pattern import hex_box
    precision = reqs[0].precision
    if any(r.precision != precision for r in reqs):
        raise ValueError("joined: one ILU precision for the whole batch")
    ctx = Context(device)
    dev = torch.empty(0, dtype=torch.float64, device=f"cuda:{device}")
    nrhs = sum(max(1, r.nrhs) for r in reqs)

    def fill(mats):
        for A, r in zip(mats, reqs):
            c = r.c()
            A.fill_synthetic(r.seed_A, c.s_diag, c.s_off)

    def timed_solve(bt, B):
        X = [torch.zeros_like(b) for b in B]
        st = bt.solve(B, X, tol=tol, rtol=rtol, nmax=nmax)
        torch.cuda.synchronize(device)
        return st

    def summary(st, B, t0, t1, t2, t3):
        bnorm = [float(torch.linalg.norm(b[i])) for b in B for i in range(b.shape[0])]
        return dict(systems=len(reqs), rhs=nrhs, wall_s=t3 - t0, rhs_per_s=nrhs / (t3 - t0),
                    systems_per_s=len(reqs) / (t3 - t0), t_fill_s=t1 - t0, t_join_factor_s=t2 - t1,
                    t_solve_s=t3 - t2, rhs_per_s_solve=nrhs / (t3 - t2),
                    iterations=int(sum(s.niter for s in st)), max_iterations=int(max(s.niter for s in st)),
                    max_residual=float(max(s.residual_norm for s in st)),
                    max_rel_residual=float(max(s.residual_norm / nb for s, nb in zip(st, bnorm))))

    try:
        torch.cuda.synchronize(device)
        t0 = time.perf_counter()
        mats = [chidg_matrix(hex_box(r.nx, r.ny, r.nz, r.n, periodic=tuple(bool(r.periodic >> d & 1) for d in range(3))),
                             ctx) for r in reqs]
        fill(mats)
        B = [torch.stack([A.synthetic_vector(r.seed_b + i, like=dev) for i in range(max(1, r.nrhs))])
             for A, r in zip(mats, reqs)]
        torch.cuda.synchronize(device)
        t1 = time.perf_counter()
        bt = joined_batch(mats, [b.shape[0] for b in B], precision=precision, nkrylov=nkrylov)
        for A in mats:
            A.release()
        t2 = time.perf_counter()
        st = timed_solve(bt, B)
        t3 = time.perf_counter()
        out = summary(st, B, t0, t1, t2, t3)
        out["groups"] = bt.info()["groups"]
        if warm:
            torch.cuda.synchronize(device)
            t0 = t1 = time.perf_counter()
            bt.update()
            t2 = time.perf_counter()
            st2 = timed_solve(bt, B)
            t3 = time.perf_counter()
            out["warm"] = summary(st2, B, t0, t1, t2, t3)
            out["warm"]["same_iterations"] = [s.niter for s in st2] == [s.niter for s in st]
        bt.close()
    finally:
        ctx.close()
    return out


def poisson_load(server: Server, reqs: list[Request], rate: float, rng: np.random.Generator) -> dict:
    """Submit ``reqs`` with exponential inter-arrival times of mean 1/rate
    (seconds), wait for all; latency = submit -> done (queueing included)."""
    gaps = rng.exponential(1.0 / rate, size=len(reqs))
    tickets = []
    t0 = time.perf_counter()
    t_next = t0
    for r, g in zip(reqs, gaps):
        t_next += g
        d = t_next - time.perf_counter()
        if d > 0:
            time.sleep(d)
        tickets.append(server.submit(r))
    res = [server.wait(t) for t in tickets]
    wall = time.perf_counter() - t0
    lat = np.array([r.t_latency for r in res])
    return dict(rate=rate, requests=len(reqs), wall_s=wall, solves_per_s=len(reqs) / wall,
                p50_ms=1e3 * float(np.percentile(lat, 50)), p99_ms=1e3 * float(np.percentile(lat, 99)),
                mean_queue_ms=1e3 * float(np.mean([r.t_queue for r in res])),
                mean_run_ms=1e3 * float(np.mean([r.t_run for r in res])),
                max_residual=float(max(r.residual_norm for r in res)),
                peak_active=server.info()["peak_active"])
