#!/usr/bin/env python3
"""Benchmark: P2 16^3 Newton-Krylov linear solves/s on one MI355X.

Workload (one "step" = one complete linear solve): the first-Newton-step
Jacobian of steady DG Euler on the uniform periodic hex box [0,1]^3, 16x16x16
elements, P2 (27 terms x 5 equations -> 135x135 FP64 blocks, 28,672 blocks):
dR/dq + mass/dtau (CFL 1) at freestream Mach 0.3 plus a seeded N(0, 0.01)
perturbation of every modal coefficient (chidg_amd/euler.py), RHS seeded
N(0,1); block ILU(0) update + FGMRES(30) to rtol 1e-8 (inner FGMRES off, CGS).
Inputs are resident in HBM when timing starts.

``value`` is the FP32-ILU form's solves/s (ILU factor + apply in FP32, Arnoldi,
residual and solution in FP64); the plain FP64 form is measured beside it.
The BF16-stored-ILU form is measured the same way.  The roofline object
describes the ILU(0) apply (the metric's "ILU apply GB/s") in the schedule a
long solve settles in (the row-slice schedule: one launch per wavefront level
over row slices of the factor, W = Dinv U), timed with HIP events on the
library's own stream.  The CPU baseline
runs the C oracle (the reference algorithm restated) on the host cores, on the
same system.
"""
from __future__ import annotations

import argparse
import json
import os
import statistics
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

N_ELEM = 16
ORDER = 2
NFIELDS = 5
MACH = 0.3
PERTURB = 0.01
CFL = 1.0
SEED_Q = 1          # state perturbation
SEED_B = 1          # right-hand side
SEED_A = 20261202   # synthetic systems (P3 section, synthetic P2 section)
HBM_PEAK_GBS = 8000.0
# The FP32 ILU(0) apply in the row-slice schedule: one call = 92 launches of
# k_rs_level<float,...>, one per wavefront level (46 forward, 46 backward),
# 40,960 tasks of (block row, 27-row slice) in all.
APPLY_KERNEL = "ILU(0) apply FP32, row-slice schedule: k_rs_level<float,4,R,E> x 92 launches per call"
# HBM bytes per FP32 row-slice apply call: rocprofv3 FETCH_SIZE (x 2.0) + WRITE_SIZE
# (x 1.0) on the Euler 16^3 P2 system, the factors being what the same counters read for
# a 512 MiB device copy (1,015,833 + 7,520 KiB raw per call).
APPLY_TRAFFIC_FP32 = 2.088e9


def sig(v, digits=5):
    """Floats to a few significant digits for the JSON line."""
    return float(f"{v:.{digits}g}") if isinstance(v, float) else v


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def ilu_apply_bytes(p, itemsize):
    """s*(n_L + n_U + N_e)*n^2 + 3*8*N_e*n (read v, read z parents, write z)."""
    c = p.counts()
    n = p.n
    return itemsize * (c["lower"] + c["upper"] + c["diag"]) * n * n + 3 * 8 * p.nrows * n


def mv_bytes(p):
    return 8 * (p.nnzb * p.n * p.n + 2 * p.nrows * p.n)


def time_async(ctx, fn, reps):
    """Average duration of fn() (enqueue-only) from HIP events on the library stream."""
    import torch
    s = torch.cuda.ExternalStream(ctx.stream_ptr())
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    fn()
    torch.cuda.synchronize()
    e0.record(s)
    for _ in range(reps):
        fn()
    e1.record(s)
    e1.synchronize()
    return e0.elapsed_time(e1) / reps / 1e3


def euler_system():
    """(pattern, blocks, b): the first-Newton-step Euler Jacobian at 16^3 P2
    and a seeded N(0,1) right-hand side, as host arrays."""
    from chidg_amd import euler
    t0 = time.perf_counter()
    p, blocks, _, _ = euler.system(N_ELEM, ORDER, mach=MACH, seed=SEED_Q, perturb=PERTURB, cfl=CFL)
    b = np.random.default_rng(SEED_B).standard_normal(p.vec_len)
    log(f"Euler 16^3 P2 Jacobian generated in {time.perf_counter() - t0:.2f} s")
    return p, blocks, b


def time_first(ctx, fn):
    """Duration of one fn() (enqueue-only) from HIP events on the library stream."""
    import torch
    s = torch.cuda.ExternalStream(ctx.stream_ptr())
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record(s)
    fn()
    e1.record(s)
    e1.synchronize()
    return e0.elapsed_time(e1) / 1e3


ITEMSIZE = {"fp64": 8, "fp32": 4, "bf16": 2}


def measure(args, p, A, b, steps, warmup, kernels=True, precisions=("fp32", "fp64", "bf16")):
    """The FP32-ILU, FP64 and BF16-stored-ILU forms of one system: solves/s
    over ``steps`` timed solves (default schedule), and with ``kernels`` the
    apply under the level, row-slice, stream and folded schedules, the
    row-slice storage's preparation (W = Dinv U and the row-slice copies) and
    the matvec, timed on the library stream."""
    import torch
    from chidg_amd.linsolve import chidg_mv_async, fgmres, precon_ILU0
    ctx = A.ctx
    bnorm = float(torch.linalg.norm(b))
    out = {}
    if args.tune:
        A.tune(reps=5)
    for prec in precisions:
        M = precon_ILU0(prec, schedule=args.schedule).init(A)
        ls = fgmres(nkrylov=30, tol=0.0, rtol=1e-8, inner_fgmres=False, orthogonalization="CGS")
        x = torch.zeros_like(b)
        for _ in range(warmup):
            ls.solve(A, x, b, M)
        tune = {}
        if args.tune:
            # launch-shape autotuner of the level schedule on the live factor
            # (setup, outside the timed region)
            if not warmup:
                M.update(A)
            t0 = time.perf_counter()
            M.tune(A, reps=3)
            tune = dict(apply_shape=M.launch_shape(), tune_s=time.perf_counter() - t0)
        torch.cuda.synchronize()
        lat = []
        t0 = time.perf_counter()
        for _ in range(steps):
            ts = time.perf_counter()
            ls.solve(A, x, b, M)
            lat.append(time.perf_counter() - ts)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        st = ls.stats
        true_res = float(torch.linalg.norm(b - _mv(A, x)) / bnorm)
        r = dict(solves_per_s=steps / wall, ms_per_solve=1e3 * wall / steps,
                 p50_ms=1e3 * statistics.median(lat), p99_ms=1e3 * float(np.percentile(lat, 99)),
                 niter=st.niter, residual_rel=st.residual_norm / bnorm, true_residual_rel=true_res,
                 t_factor_ms=1e3 * st.t_factor, t_precon_ms=1e3 * st.t_precon, t_mv_ms=1e3 * st.t_mv,
                 t_dot_ms=1e3 * st.t_dot, t_norm_ms=1e3 * st.t_norm, wall_s=wall, x=x, **tune)
        if kernels:
            nbytes = ilu_apply_bytes(p, ITEMSIZE[prec])
            v = b / bnorm
            z = torch.empty_like(b)
            y = torch.empty_like(b)
            torch.cuda.synchronize()
            # inputs are complete (synchronized above): enqueue without cross-stream waits
            app = lambda: M.apply_async(A, v, z, order=False)   # noqa: E731
            M.update(A)
            M.set_schedule("levels")
            t_lv = time_async(ctx, app, args.kreps)
            M.set_schedule("rowslice")          # W = Dinv U and the row-slice copies at the first apply
            t_first = time_first(ctx, app)
            t_rs = time_async(ctx, app, args.kreps)
            M.set_schedule("stream")            # the same storage, one persistent launch
            t_st = time_async(ctx, app, args.kreps)
            r.update(apply_levels_ms=1e3 * t_lv, apply_levels_gbs=nbytes / t_lv / 1e9,
                     apply_rowslice_ms=1e3 * t_rs, apply_rowslice_gbs=nbytes / t_rs / 1e9,
                     apply_stream_ms=1e3 * t_st, apply_stream_gbs=nbytes / t_st / 1e9,
                     rowslice_prep_ms=1e3 * (t_first - t_rs))
            if prec != "bf16" and args.folded:
                M.update(A)
                M.set_schedule("folded")
                t_ff = time_first(ctx, app)
                t_fd = time_async(ctx, app, args.kreps)
                r.update(apply_folded_ms=1e3 * t_fd, apply_folded_gbs=nbytes / t_fd / 1e9,
                         fold_ms=1e3 * (t_ff - t_fd))
            M.set_schedule(args.schedule or "auto")
            t_mv = time_async(ctx, lambda: chidg_mv_async(A, v, y, order=False), args.kreps)
            r.update(mv_ms=1e3 * t_mv, mv_gbs=mv_bytes(p) / t_mv / 1e9)
        out[prec] = r
        log(f"{prec}: {r['solves_per_s']:.2f} solves/s, niter {st.niter}, res {r['residual_rel']:.2e}, "
            f"factor {r['t_factor_ms']:.2f} ms, precon {r['t_precon_ms']:.2f} ms, mv {r['t_mv_ms']:.2f} ms"
            + (f"; apply levels {r['apply_levels_ms']:.3f} ms ({r['apply_levels_gbs']:.0f} GB/s), row-slice "
               f"{r['apply_rowslice_ms']:.3f} ms ({r['apply_rowslice_gbs']:.0f} GB/s), stream "
               f"{r['apply_stream_ms']:.3f} ms ({r['apply_stream_gbs']:.0f} GB/s), prep {r['rowslice_prep_ms']:.2f} "
               f"ms, mv {r['mv_ms']:.3f} ms ({r['mv_gbs']:.0f} GB/s)" if kernels else ""))
        M.tear_down()
    x64 = out["fp64"].pop("x")
    for prec in precisions:
        if prec != "fp64":
            xp = out[prec].pop("x")
            out[prec]["x_rel_vs_fp64"] = float(torch.linalg.norm(xp - x64) / torch.linalg.norm(x64))
    out["mv_shape"] = A.launch_shape()
    return out, x64


def gpu_run(args, p, blocks, b_host):
    import torch
    from chidg_amd.linsolve import Context, chidg_matrix
    ctx = Context(0)
    A = chidg_matrix(p, ctx)
    A.store(blocks)
    b = torch.from_numpy(b_host).to("cuda:0")
    out, x64 = measure(args, p, A, b, args.steps, args.warmup)
    return out, x64.cpu().numpy(), A


def synthetic_run(args):
    """The seeded synthetic 16^3 P2 system (diag blocks 2I + G/sqrt(n),
    off-diagonal 0.1 G/sqrt(n), RHS N(0,1) from the counter hash): solves/s
    of both forms, for comparison with the Euler system's."""
    import torch
    from chidg_amd import pattern as pt
    from chidg_amd.linsolve import Context, chidg_matrix
    n = pt.block_size(ORDER, NFIELDS)
    p = pt.hex_box(N_ELEM, N_ELEM, N_ELEM, n)
    ctx = Context(0)
    A = chidg_matrix(p, ctx)
    A.fill_synthetic(SEED_A, 1.0 / np.sqrt(n), 0.1 / np.sqrt(n))
    b = A.synthetic_vector(SEED_B, 1.0, like=torch.empty(0, dtype=torch.float64, device="cuda:0"))
    out, _ = measure(args, p, A, b, 5, 1, kernels=False, precisions=("fp32", "fp64"))
    A.release()
    ctx.close()            # its cached device memory back to the device
    return {k: {kk: sig(vv) for kk, vv in v.items()} for k, v in out.items()}


def p3_run(args):
    """16^3 periodic P3 (320 x 320 blocks, A = 23.5 GB FP64): one FGMRES(30)
    solve to rtol 1e-8 per ILU(0) precision; the FP64-ILU solution is the
    reference the FP32 and BF16-stored forms are measured against."""
    import torch
    from chidg_amd import pattern as pt
    from chidg_amd.linsolve import Context, chidg_matrix, fgmres, precon_ILU0
    n = pt.block_size(3, NFIELDS)
    p = pt.hex_box(N_ELEM, N_ELEM, N_ELEM, n)
    ctx = Context(0)
    A = chidg_matrix(p, ctx)
    A.fill_synthetic(SEED_A, 1.0 / np.sqrt(n), 0.1 / np.sqrt(n))
    b = A.synthetic_vector(SEED_B, 1.0, like=torch.empty(0, dtype=torch.float64, device="cuda:0"))
    if args.tune:
        A.tune(reps=3)
    out, xs = {}, {}
    for prec in ("fp64", "fp32", "bf16"):
        M = precon_ILU0(prec).init(A)
        ls = fgmres(nkrylov=30, tol=0.0, rtol=1e-8, inner_fgmres=False)
        x = torch.zeros_like(b)
        ls.solve(A, x, b, M)             # warm-up (allocations)
        if args.tune:
            M.tune(A, reps=2)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        ls.solve(A, x, b, M)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        xs[prec] = x
        out[prec] = dict(solves_per_s=1.0 / dt, ms_per_solve=1e3 * dt, niter=ls.niter,
                         residual_rel=ls.residual_norm / float(torch.linalg.norm(b)),
                         t_factor_ms=1e3 * ls.stats.t_factor)
        log(f"P3 {prec}: {1e3 * dt:.1f} ms/solve, niter {ls.niter}, factor {1e3 * ls.stats.t_factor:.1f} ms")
        M.tear_down()
    for prec in ("fp32", "bf16"):
        out[prec]["x_rel_vs_fp64_ilu"] = float(torch.linalg.norm(xs[prec] - xs["fp64"]) / torch.linalg.norm(xs["fp64"]))
    A.release()
    ctx.close()
    return {k: {kk: sig(vv) for kk, vv in v.items()} for k, v in out.items()}


def pretune_small(args):
    """Launch shapes for the request loop's block sizes (P1-P3, FP32 ILU),
    tuned on 8^3 synthetic systems; the loop's own contexts start from them."""
    from chidg_amd import pattern as pt
    from chidg_amd.linsolve import Context, chidg_matrix, precon_ILU0
    ctx = Context(0)
    shapes = {}
    for order in (1, 2, 3):
        n = pt.block_size(order, NFIELDS)
        p = pt.hex_box(8, 8, 8, n)
        A = chidg_matrix(p, ctx)
        A.fill_synthetic(SEED_A, 1.0 / np.sqrt(n), 0.1 / np.sqrt(n))
        A.tune(reps=3)
        M = precon_ILU0("fp32").init(A)
        M.update(A)
        M.tune(A, reps=3)
        shapes[f"P{order}"] = dict(apply=M.launch_shape(), mv=A.launch_shape())
        M.tear_down()
        A.release()
    ctx.close()
    log(f"pre-tuned launch shapes: {shapes}")
    return shapes


def batch_run(args):
    """64 independent systems, meshes N^3 with N in 4..12, P1..P3 and 1..8
    right-hand sides drawn uniformly (seeded), FP32 ILU(0), solved two ways,
    each timed over the whole job (matrix fill, factor, solve), cold (the
    first batch: every allocation fresh) and warm (the same batch again on the
    same objects):
    * request loop: all systems submitted at once (6 slots by default); warm =
      the same server again (its slots' device memory cached);
    * joined: the systems of each block size in one block-diagonal matrix, one
      factor, every right-hand side a lockstep FGMRES lane, per-step graphs;
      warm = the joined matrices refactored in place (no allocation, no
      fill) and solved."""
    import torch
    from chidg_amd.serve import Server, batch, joined, sweep_requests
    torch.cuda.empty_cache()   # the joined batch takes about 200 GB of device memory
    rng = np.random.default_rng(20261202)
    reqs = sweep_requests(64, rng, meshes=(4, 12), orders=(1, 2, 3), nrhs=(1, 8))
    fmt = lambda d: {k: sig(float(v)) if isinstance(v, float) else v for k, v in d.items()}
    with Server(nslots=args.batch_slots, nkrylov=30, tol=0.0, rtol=1e-8, inner_fgmres=False) as srv:
        loop = [batch(srv, reqs) for _ in range(2)]
    for tag, d in zip(("cold", "warm"), loop):
        log(f"batch (request loop, {tag}): {d['rhs']} rhs in {d['wall_s']:.2f} s ({d['rhs_per_s']:.1f} rhs/s), "
            f"peak {d['peak_active']} concurrent")
    joined(reqs[:4], rtol=1e-8, warm=False)          # kernels loaded, graph path exercised
    jn = joined(reqs, rtol=1e-8)
    jw = jn.pop("warm")
    for tag, d in (("cold", jn), ("warm", jw)):
        log(f"batch (joined, {tag}): {d['rhs']} rhs in {d['wall_s']:.2f} s ({d['rhs_per_s']:.1f} rhs/s; fill "
            f"{d['t_fill_s']:.2f} s, join/update + factor {d['t_join_factor_s']:.2f} s, solve {d['t_solve_s']:.2f} s "
            f"= {d['rhs_per_s_solve']:.0f} rhs/s), max relative residual {d['max_rel_residual']:.2e}")
    out = {"request_loop": {"cold": fmt(loop[0]), "warm": fmt(loop[1])}, "joined": {"cold": fmt(jn), "warm": fmt(jw)}}
    best = max((("request_loop", loop[1]["rhs_per_s"]), ("joined", jw["rhs_per_s"])), key=lambda t: t[1])
    out["rhs_per_s_warm"] = sig(float(best[1]))
    out["rhs_per_s_warm_path"] = best[0]
    return out


SERVE_RATES = (10.0, 50.0, 100.0, 200.0, 500.0)


def serve_run(args):
    """Poisson arrivals of sweep solves at the config-3/4 size mix (meshes
    4^3..12^3, P1..P3, Mach-scaled coupling), FP32 ILU(0), 6 slots by
    default, at each rate of SERVE_RATES (requests per second); p50/p99
    latency from submit to done over ``--serve-requests`` requests per rate."""
    from chidg_amd.serve import Server, poisson_load, sweep_requests
    out = []
    with Server(nslots=args.serve_slots, nkrylov=30, tol=0.0, rtol=1e-8, inner_fgmres=False) as srv:
        warm = sweep_requests(4, np.random.default_rng(1), meshes=(4, 6))
        poisson_load(srv, warm, rate=1000.0, rng=np.random.default_rng(2))
        for rate in SERVE_RATES:
            rng = np.random.default_rng(int(rate))
            reqs = sweep_requests(args.serve_requests, rng, meshes=(4, 12))
            r = poisson_load(srv, reqs, rate=rate, rng=rng)
            log(f"serve rate {rate:.0f}/s: {r['solves_per_s']:.1f} solves/s, p50 {r['p50_ms']:.1f} ms, "
                f"p99 {r['p99_ms']:.1f} ms, peak {r['peak_active']} concurrent")
            out.append({k: sig(float(v)) if isinstance(v, float) else v for k, v in r.items()})
    return out


def _guarded(fn, args, what):
    try:
        return fn(args)
    except Exception as e:   # an extra section never costs the headline line
        log(f"{what} failed: {e!r}")
        return {"error": repr(e)}


def _mv(A, x):
    from chidg_amd.linsolve import chidg_mv
    return chidg_mv(A, x)


def cpu_baseline(p, blocks, b, x_gpu):
    """The oracle (the reference algorithm restated in C) on the same system,
    ILU(0) update + FGMRES(30) to rtol 1e-8, on the host's cores:

    * all cores of this process's CPU share (OpenMP over the rows of a
      wavefront level and over matvec rows; the order of every sum is the
      sequential one, so iteration counts do not change): full solves, repeated
      until about 10 s of work (at most 4);
    * one core (the reference at one MPI rank, the only rank count whose ILU(0)
      is the full ILU(0)): one ILU(0) update and the first 8 FGMRES
      iterations, timed; the solve time is the update plus 8 iterations'
      time scaled to the all-core run's iteration count (a bounded sample: the
      full one-core solve takes about a minute).

    Also the GPU FP64 solution's distance from the oracle's."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as po
    share = len(os.sched_getaffinity(0))
    cores = min(share, 16)
    po.set_threads(cores)
    tf = ts = 0.0
    runs = 0
    while runs < 4 and (runs == 0 or tf + ts < 10.0):
        t0 = time.perf_counter()
        L, D = po.ilu0_factor(p, blocks)
        t1 = time.perf_counter()
        x, niter, res, _ = po.fgmres(p, blocks, b, factors=(L, D), nkrylov=30, tol=0.0, rtol=1e-8)
        t2 = time.perf_counter()
        tf += t1 - t0
        ts += t2 - t1
        runs += 1
        del L, D
    total = (tf + ts) / runs
    log(f"cpu oracle ({cores} threads, {runs} solves): factor {tf / runs:.2f} s, fgmres {ts / runs:.2f} s "
        f"({niter} it)")
    po.set_threads(1)
    t0 = time.perf_counter()
    L, D = po.ilu0_factor(p, blocks)
    t1 = time.perf_counter()
    k1 = 8
    po.fgmres(p, blocks, b, factors=(L, D), nkrylov=30, tol=0.0, rtol=1e-8, nmax=k1)
    t2 = time.perf_counter()
    del L, D
    one_factor, one_iter = t1 - t0, (t2 - t1) / k1
    one_total = one_factor + niter * one_iter
    log(f"cpu oracle (1 thread): factor {one_factor:.2f} s, {one_iter * 1e3:.0f} ms per FGMRES iteration "
        f"-> {one_total:.1f} s per {niter}-iteration solve")
    po.set_threads(cores)
    parity = dict(oracle_niter=niter, oracle_residual_rel=res / float(np.linalg.norm(b)),
                  gpu_fp64_x_rel_vs_oracle=float(np.linalg.norm(x_gpu - x) / np.linalg.norm(x)))
    return dict(value=round(1.0 / total, 4), unit="solves/s", cores=cores, kind="port",
                sample=f"{runs} full solve(s) of the same Euler 16^3 P2 system (ILU0 update {tf / runs:.2f} s + "
                       f"FGMRES(30) {niter} iterations {ts / runs:.2f} s each), C oracle with {cores} OpenMP "
                       f"threads (all cores of the CPU share: sched_getaffinity {share}, os.cpu_count "
                       f"{os.cpu_count()})",
                nproc=os.cpu_count(), cpu_share=share,
                one_core={"value": round(1.0 / one_total, 5), "unit": "solves/s", "cores": 1, "kind": "port",
                          "sample": f"one ILU0 update ({one_factor:.2f} s) + the first {k1} FGMRES(30) iterations "
                                    f"({one_iter * 1e3:.0f} ms each) with 1 thread; solve = update + {niter} "
                                    f"iterations"}), parity


def profile_workload(name):
    """Short fixed workloads for counter and trace passes (no JSON line):
    * apply: the FP32 row-slice ILU(0) apply on the Euler 16^3 P2 system, 5
      times after the one that forms the row-slice storage, then a 512 MiB
      device copy 3 times (read 512 MiB + write 512 MiB each: the reference
      for scaling HBM counters);
    * factor: the ILU(0) factor of the Euler 16^3 P2 system, twice in FP64 and
      twice in FP32, then the seeded synthetic 16^3 P3 system once in FP32 and
      once in FP64;
    * batch: the 64-system uneven batch as one joined solve, cold then warm."""
    import torch
    from chidg_amd import pattern as pt
    from chidg_amd.linsolve import Context, chidg_matrix, precon_ILU0
    if name == "batch":
        from chidg_amd.serve import joined, sweep_requests
        reqs = sweep_requests(64, np.random.default_rng(20261202), meshes=(4, 12), orders=(1, 2, 3), nrhs=(1, 8))
        out = joined(reqs, rtol=1e-8)
        log(f"batch workload: cold {out['wall_s']:.2f} s, warm {out['warm']['wall_s']:.2f} s")
        return
    ctx = Context(0)
    p, blocks, b_host = euler_system()
    A = chidg_matrix(p, ctx)
    A.store(blocks)
    if name == "apply":
        b = torch.from_numpy(b_host).to("cuda:0")
        v = b / torch.linalg.norm(b)
        z = torch.empty_like(b)
        M = precon_ILU0("fp32", schedule="rowslice").init(A)
        M.update(A)
        M.apply_async(A, v, z)               # forms W and the row-slice copies
        torch.cuda.synchronize()
        for _ in range(5):
            M.apply_async(A, v, z)
        torch.cuda.synchronize()
        log("apply workload: 5 row-slice applies done")
        src = torch.ones(64 * 1024 * 1024, dtype=torch.float64, device="cuda:0")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        log("apply workload: 3 x 512 MiB copies done")
    elif name == "factor":
        for prec in ("fp64", "fp32"):
            M = precon_ILU0(prec).init(A)
            for _ in range(2):
                M.update(A)
            torch.cuda.synchronize()
            log(f"factor workload: Euler 16^3 P2 {prec} x2 done")
            M.tear_down()
        A.release()
        n = pt.block_size(3, NFIELDS)
        A = chidg_matrix(pt.hex_box(N_ELEM, N_ELEM, N_ELEM, n), ctx)
        A.fill_synthetic(SEED_A, 1.0 / np.sqrt(n), 0.1 / np.sqrt(n))
        for prec in ("fp32", "fp64"):
            M = precon_ILU0(prec).init(A)
            M.update(A)
            torch.cuda.synchronize()
            log(f"factor workload: synthetic 16^3 P3 {prec} done")
            M.tear_down()
    A.release()
    ctx.close()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--kreps", type=int, default=20, help="kernel timing repetitions")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--schedule", choices=["levels", "dataflow", "folded", "stream", "rowslice", "auto"], default=None,
                    help="ILU(0) apply schedule (default: the library's)")
    ap.add_argument("--extras", default="synthetic,p3,batch,serve",
                    help="comma list of extra sections: synthetic (seeded synthetic 16^3 P2 system), p3 "
                         "(16^3 P3 mixed precision), batch (64 uneven systems), serve (Poisson request "
                         "loop); empty for none")
    ap.add_argument("--serve-requests", type=int, default=200, help="requests per arrival rate")
    # 6 slots: the process gets 4 hardware queues; at 100 req/s 6 slots gave p99 63-76 ms
    # against 82-104 ms with 8 (4 slots: 123-126 ms), the same batch throughput
    ap.add_argument("--serve-slots", type=int, default=6, help="request-loop slots (contexts) for the serve section")
    ap.add_argument("--batch-slots", type=int, default=6, help="request-loop slots (contexts) for the batch section")
    ap.add_argument("--no-tune", dest="tune", action="store_false",
                    help="skip the launch-shape autotuner (run before each timed section, outside the timed region)")
    ap.add_argument("--no-folded", dest="folded", action="store_false",
                    help="skip timing the folded level schedule (W = U Dinv) of the FP64 and FP32 forms")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="also run a plain 512 MiB device copy (reference for HBM counter scaling)")
    ap.add_argument("--workload", choices=["apply", "factor", "batch"], default=None,
                    help="run one short fixed workload for a profiler pass and exit (no JSON line)")
    args = ap.parse_args()
    if args.workload:
        profile_workload(args.workload)
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: there is no multi-GPU path (each system is solved on "
              "one GPU)", file=sys.stderr)
        sys.exit(2)
    if args.calibrate_copy:
        import torch
        src = torch.ones(64 * 1024 * 1024, dtype=torch.float64, device="cuda:0")   # 512 MiB
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        log("calibration copy: 3 x 512 MiB device-to-device (read 512 MiB + write 512 MiB each)")
        del src, dst
    p, blocks, b_host = euler_system()
    g, x64, A = gpu_run(args, p, blocks, b_host)
    A.release()
    A.ctx.close()          # the headline section's cached device memory back to the device
    f32, f64 = g["fp32"], g["fp64"]
    extras = [e for e in args.extras.split(",") if e]
    sections = {}
    if "synthetic" in extras:
        sections["synthetic_16c_p2"] = _guarded(synthetic_run, args, "synthetic")
    if "p3" in extras:
        sections["p3_16c_mixed_precision"] = _guarded(p3_run, args, "p3")
    if args.tune and ("batch" in extras or "serve" in extras):
        sections["pretuned_shapes"] = _guarded(pretune_small, args, "pretune")
    if "batch" in extras:
        sections["batch_64_uneven"] = _guarded(batch_run, args, "batch")
    if "serve" in extras:
        sections["serving"] = _guarded(serve_run, args, "serve")
    cpu, parity = None, None
    if not args.no_cpu:
        cpu, parity = cpu_baseline(p, blocks, b_host, x64)
        parity["gpu_fp64_niter"] = f64["niter"]
        log(f"parity: GPU FP64 {f64['niter']} it vs oracle {parity['oracle_niter']} it, "
            f"x rel {parity['gpu_fp64_x_rel_vs_oracle']:.2e}")
    # the apply a long solve settles in: the row-slice schedule
    t_ap = f32["apply_rowslice_ms"] / 1e3
    gbs = ilu_apply_bytes(p, 4) / t_ap / 1e9
    bf = g.get("bf16")
    roof = {"bound": "hbm", "kernel": APPLY_KERNEL,
            "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(gbs / HBM_PEAK_GBS, 4),
            "traffic": APPLY_TRAFFIC_FP32, "algorithmic_bytes": ilu_apply_bytes(p, 4),
            "avg_launch_ms": round(1e3 * t_ap, 4)}
    line = {
        "metric": "P2 16^3 Newton-Krylov linear solves/s (FP32-ILU vs FP64), ILU apply GB/s, p99 latency",
        "value": round(f32["solves_per_s"], 3), "unit": "solves/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(f32["ms_per_solve"], 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64 Krylov + f32 ILU(0)",
        "data": "synthetic: first-Newton-step DG Euler Jacobian generated in-process (seeded state "
                "perturbation), seeded N(0,1) right-hand side",
        "config": {"workload": "steady DG Euler, first Newton step, 16x16x16 periodic hex [0,1]^3, P2 "
                               "(135x135 blocks), Mach 0.3 + N(0,0.01) perturbation, CFL 1; "
                               "FGMRES(30)+ILU0 to rtol 1e-8",
                   "nrows": p.nrows, "block_n": p.n, "nnzb": p.nnzb, "seed_state": SEED_Q, "seed_b": SEED_B},
        "fp64": {k: sig(v) for k, v in f64.items()},
        "fp32_ilu": {k: sig(v) for k, v in f32.items()},
        "bf16_ilu": {k: sig(v) for k, v in bf.items()} if bf else None,
        "mv_shape": g.get("mv_shape"),
        "speedup_fp32_vs_fp64": round(f32["solves_per_s"] / f64["solves_per_s"], 3),
        "speedup_bf16_vs_fp64": round(bf["solves_per_s"] / f64["solves_per_s"], 3) if bf else None,
        "ilu_apply_gbs_fp64": round(ilu_apply_bytes(p, 8) / (f64["apply_rowslice_ms"] / 1e3) / 1e9, 1),
        "ilu_apply_gbs_fp32": round(gbs, 1),
        "ilu_apply_gbs_bf16": round(ilu_apply_bytes(p, 2) / (bf["apply_rowslice_ms"] / 1e3) / 1e9, 1) if bf else None,
        "latency_ms": {"p50": round(f32["p50_ms"], 3), "p99": round(f32["p99_ms"], 3)},
        "parity_vs_oracle": parity,
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    line.update(sections)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
