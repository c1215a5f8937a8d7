#!/usr/bin/env python3
"""Benchmark: P2 16^3 Newton-Krylov linear solves/s on one MI355X.

Workload (one "step" = one complete linear solve): uniform periodic hex box
16x16x16, P2 (27 terms x 5 equations -> 135x135 FP64 blocks, 28,672 blocks),
seeded synthetic Jacobian (diag blocks 2I + G/sqrt(n), off-diagonal
0.1*G/sqrt(n)), RHS ~N(0,1); block ILU(0) update + FGMRES(30) to rtol 1e-8
(inner FGMRES off, CGS).  Inputs are resident in HBM when timing starts.

``value`` is the FP32-ILU form's solves/s (ILU factor + apply in FP32, Arnoldi,
residual and solution in FP64); the plain FP64 form is measured beside it.
The roofline object describes the ILU(0) apply sweep (the metric's "ILU apply
GB/s"), timed with HIP events on the library's own stream.  The CPU baseline
runs the C oracle (the reference algorithm restated) on the host cores.
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
SEED_A = 20261202
SEED_B = 1
HBM_PEAK_GBS = 8000.0
# The FP32 ILU(0) apply, one call = 138 k_kslice launches + 1 k_sum_parts.
APPLY_KERNEL = "ILU(0) apply FP32: 138 x k_kslice<float,float,4> + k_sum_parts (one call)"
# HBM bytes per apply call from rocprofv3 FETCH_SIZE (x2.0) + WRITE_SIZE (x1.0), the
# factors being what the same counters read for a 512 MiB device copy
# (profiles/hbm_traffic_16x16x16_p2.md).
APPLY_TRAFFIC_FP32 = 2.236e9


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


def gpu_run(args):
    import torch
    from chidg_amd import pattern as pt
    from chidg_amd.linsolve import (Context, chidg_matrix, chidg_mv_async, fgmres, precon_ILU0)

    n = pt.block_size(ORDER, NFIELDS)
    p = pt.hex_box(N_ELEM, N_ELEM, N_ELEM, n)
    ctx = Context(0)
    A = chidg_matrix(p, ctx)
    A.fill_synthetic(SEED_A, 1.0 / np.sqrt(n), 0.1 / np.sqrt(n))
    dev = torch.empty(0, dtype=torch.float64, device="cuda:0")
    b = A.synthetic_vector(SEED_B, 1.0, like=dev)
    bnorm = float(torch.linalg.norm(b))
    out = {}
    for prec in ("fp32", "fp64"):
        M = precon_ILU0(prec, schedule=args.schedule).init(A)
        ls = fgmres(nkrylov=30, tol=0.0, rtol=1e-8, inner_fgmres=False, orthogonalization="CGS")
        x = torch.zeros_like(b)
        for _ in range(args.warmup):
            ls.solve(A, x, b, M)
        torch.cuda.synchronize()
        lat = []
        t0 = time.perf_counter()
        for _ in range(args.steps):
            ts = time.perf_counter()
            ls.solve(A, x, b, M)
            lat.append(time.perf_counter() - ts)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        st = ls.stats
        # kernel timings on the library stream
        v = b / bnorm
        z = torch.empty_like(b)
        y = torch.empty_like(b)
        t_apply = time_async(ctx, lambda: M.apply_async(A, v, z), args.kreps)
        t_mv = time_async(ctx, lambda: chidg_mv_async(A, v, y), args.kreps)
        folded = {}
        if (args.schedule or "levels") == "levels":
            # the same factor under the folded schedule (W = U Dinv formed at the
            # first apply): fold cost and apply time, for factors serving many applies
            M.set_schedule("folded")
            s = torch.cuda.ExternalStream(ctx.stream_ptr())
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            torch.cuda.synchronize()
            e0.record(s)
            M.apply_async(A, v, z)
            e1.record(s)
            e1.synchronize()
            t_first = e0.elapsed_time(e1) / 1e3
            t_af = time_async(ctx, lambda: M.apply_async(A, v, z), args.kreps)
            folded = dict(apply_folded_ms=1e3 * t_af,
                          apply_folded_gbs=ilu_apply_bytes(p, 4 if prec == "fp32" else 8) / t_af / 1e9,
                          fold_ms=1e3 * (t_first - t_af))
            M.set_schedule("levels")
        true_res = float(torch.linalg.norm(b - _mv(A, x)) / bnorm)
        out[prec] = dict(solves_per_s=args.steps / wall, ms_per_solve=1e3 * wall / args.steps,
                         p50_ms=1e3 * statistics.median(lat), p99_ms=1e3 * float(np.percentile(lat, 99)),
                         niter=st.niter, residual_rel=st.residual_norm / bnorm, true_residual_rel=true_res,
                         t_factor_ms=1e3 * st.t_factor, t_precon_ms=1e3 * st.t_precon, t_mv_ms=1e3 * st.t_mv,
                         t_dot_ms=1e3 * st.t_dot, t_norm_ms=1e3 * st.t_norm,
                         apply_ms=1e3 * t_apply, apply_gbs=ilu_apply_bytes(p, 4 if prec == "fp32" else 8) / t_apply / 1e9,
                         mv_ms=1e3 * t_mv, mv_gbs=mv_bytes(p) / t_mv / 1e9, wall_s=wall, x=x, **folded)
        log(f"{prec}: {out[prec]['solves_per_s']:.2f} solves/s, niter {st.niter}, "
            f"res {out[prec]['residual_rel']:.2e}, factor {1e3 * st.t_factor:.2f} ms, "
            f"apply {1e3 * t_apply:.3f} ms ({out[prec]['apply_gbs']:.0f} GB/s), "
            f"mv {1e3 * t_mv:.3f} ms ({out[prec]['mv_gbs']:.0f} GB/s)"
            + (f", folded apply {folded['apply_folded_ms']:.3f} ms ({folded['apply_folded_gbs']:.0f} GB/s, "
               f"fold {folded['fold_ms']:.2f} ms)" if folded else ""))
        M.tear_down()
    x64, x32 = out["fp64"].pop("x"), out["fp32"].pop("x")
    out["fp32"]["x_rel_vs_fp64"] = float(torch.linalg.norm(x32 - x64) / torch.linalg.norm(x64))
    return p, out


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
    out, xs = {}, {}
    for prec in ("fp64", "fp32", "bf16"):
        M = precon_ILU0(prec).init(A)
        ls = fgmres(nkrylov=30, tol=0.0, rtol=1e-8, inner_fgmres=False)
        x = torch.zeros_like(b)
        ls.solve(A, x, b, M)             # warm-up (allocations)
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
    return {k: {kk: sig(vv) for kk, vv in v.items()} for k, v in out.items()}


def batch_run(args):
    """64 independent systems, meshes N^3 with N in 4..12, P1..P3 and 1..8
    right-hand sides drawn uniformly (seeded), FP32 ILU(0), all submitted at
    once to the request loop (8 slots)."""
    from chidg_amd.serve import Server, batch, sweep_requests
    rng = np.random.default_rng(20261202)
    reqs = sweep_requests(64, rng, meshes=(4, 12), orders=(1, 2, 3), nrhs=(1, 8))
    with Server(nslots=8, nkrylov=30, tol=0.0, rtol=1e-8, inner_fgmres=False) as srv:
        out = batch(srv, reqs)
    log(f"batch: {out['systems']} systems / {out['rhs']} rhs in {out['wall_s']:.2f} s "
        f"({out['rhs_per_s']:.1f} rhs/s), peak {out['peak_active']} concurrent")
    return {k: sig(float(v)) if isinstance(v, float) else v for k, v in out.items()}


def serve_run(args):
    """Poisson arrivals of sweep solves (meshes 4^3..8^3, P1..P3, Mach-scaled
    coupling), FP32 ILU(0), 8 slots; p50/p99 latency from submit to done."""
    from chidg_amd.serve import Server, poisson_load, sweep_requests
    out = []
    with Server(nslots=8, nkrylov=30, tol=0.0, rtol=1e-8, inner_fgmres=False) as srv:
        warm = sweep_requests(4, np.random.default_rng(1), meshes=(4, 6))
        poisson_load(srv, warm, rate=1000.0, rng=np.random.default_rng(2))
        for rate in (10.0, 100.0, 500.0):
            rng = np.random.default_rng(int(rate))
            reqs = sweep_requests(args.serve_requests, rng, meshes=(4, 8))
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


def cpu_baseline(p, niter_target):
    """Oracle (reference algorithm restated in C, OpenMP over rows of a
    wavefront level and over matvec rows) on the same 16^3 P2 system: one ILU
    update + FGMRES(30) to rtol 1e-8."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as po
    cores = min(len(os.sched_getaffinity(0)), 16)
    po.set_threads(cores)
    n = p.n
    t0 = time.perf_counter()
    blocks = po.fill_synthetic(p, SEED_A, 1.0 / np.sqrt(n), 0.1 / np.sqrt(n))
    b = po.fill_vector(p.vec_len, SEED_B, 1.0)
    tgen = time.perf_counter() - t0
    t0 = time.perf_counter()
    L, D = po.ilu0_factor(p, blocks)
    tf = time.perf_counter() - t0
    t1 = time.perf_counter()
    x, niter, res, _ = po.fgmres(p, blocks, b, factors=(L, D), nkrylov=30, tol=0.0, rtol=1e-8)
    ts = time.perf_counter() - t1
    total = tf + ts
    log(f"cpu oracle ({cores} threads): factor {tf:.2f} s, fgmres {ts:.2f} s ({niter} it), gen {tgen:.1f} s")
    return dict(value=1.0 / total, unit="solves/s", cores=cores, kind="port",
                sample=f"one full solve of the same 16^3 P2 system (ILU0 update {tf:.2f} s + FGMRES(30) "
                       f"{niter} iterations {ts:.2f} s), C oracle with {cores} OpenMP threads"), x, niter


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--kreps", type=int, default=20, help="kernel timing repetitions")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--schedule", choices=["levels", "dataflow", "folded"], default=None,
                    help="ILU(0) apply schedule (default: the library's)")
    ap.add_argument("--extras", default="p3,batch,serve",
                    help="comma list of extra sections: p3 (16^3 P3 mixed precision), batch (64 uneven "
                         "systems), serve (Poisson request loop); empty for none")
    ap.add_argument("--serve-requests", type=int, default=40, help="requests per arrival rate")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="also run a plain 512 MiB device copy (reference for HBM counter scaling)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (one system per GPU; "
              "sharding is described in DESIGN.md)", file=sys.stderr)
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
    p, g = gpu_run(args)
    f32, f64 = g["fp32"], g["fp64"]
    extras = [e for e in args.extras.split(",") if e]
    sections = {}
    if "p3" in extras:
        sections["p3_16c_mixed_precision"] = _guarded(p3_run, args, "p3")
    if "batch" in extras:
        sections["batch_64_uneven"] = _guarded(batch_run, args, "batch")
    if "serve" in extras:
        sections["serving"] = _guarded(serve_run, args, "serve")
    cpu = None
    if not args.no_cpu:
        cpu, x_cpu, n_cpu = cpu_baseline(p, f64["niter"])
    roof_ms = f32["apply_ms"] / 1e3
    roof = {"bound": "hbm", "kernel": APPLY_KERNEL,
            "achieved": round(f32["apply_gbs"], 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(f32["apply_gbs"] / HBM_PEAK_GBS, 4),
            "traffic": APPLY_TRAFFIC_FP32, "algorithmic_bytes": ilu_apply_bytes(p, 4),
            "avg_launch_ms": round(1e3 * roof_ms, 4)}
    line = {
        "metric": "P2 16^3 Newton-Krylov linear solves/s (FP32-ILU vs FP64), ILU apply GB/s, p99 latency",
        "value": round(f32["solves_per_s"], 3), "unit": "solves/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(f32["ms_per_solve"], 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64 Krylov + f32 ILU(0)", "data": "synthetic",
        "config": {"workload": "16x16x16 periodic hex, P2 (135x135 blocks), FGMRES(30)+ILU0 to rtol 1e-8",
                   "nrows": p.nrows, "block_n": p.n, "nnzb": p.nnzb, "seed_A": SEED_A, "seed_b": SEED_B},
        "fp64": {k: sig(v) for k, v in f64.items()},
        "fp32_ilu": {k: sig(v) for k, v in f32.items()},
        "speedup_fp32_vs_fp64": round(f32["solves_per_s"] / f64["solves_per_s"], 3),
        "ilu_apply_gbs_fp64": round(f64["apply_gbs"], 1), "ilu_apply_gbs_fp32": round(f32["apply_gbs"], 1),
        "latency_ms": {"p50": round(f32["p50_ms"], 3), "p99": round(f32["p99_ms"], 3)},
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    line.update(sections)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
