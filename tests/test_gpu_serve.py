"""The request loop (cdg_server_*): results alone and in a mix are the same,
admission keeps the running footprint inside the budget, and every request
of a Poisson burst completes with its residual below rtol."""
import numpy as np
import pytest

from chidg_amd.serve import Request, Server, poisson_load, sweep_requests

pytestmark = pytest.mark.gpu

KW = dict(nkrylov=30, tol=0.0, rtol=1e-8, inner_fgmres=False)


def test_alone_equals_in_mix():
    rng = np.random.default_rng(5)
    probe = [Request(6, 6, 6, 2, seed_A=101, seed_b=7), Request(5, 5, 5, 3, seed_A=102, seed_b=8, precision="bf16"),
             Request(8, 8, 8, 1, seed_A=103, seed_b=9, precision="fp64")]
    alone = []
    with Server(nslots=1, **KW) as s:
        for r in probe:
            alone.append(s.wait(s.submit(r, want_x=True)))
    mix = sweep_requests(12, rng, meshes=(4, 8))
    with Server(nslots=6, **KW) as s:
        tickets = [s.submit(r) for r in mix[:6]]
        mine = [s.submit(r, want_x=True) for r in probe]
        tickets += [s.submit(r) for r in mix[6:]]
        got = [s.wait(t) for t in mine]
        others = [s.wait(t) for t in tickets]
        assert s.info()["peak_active"] > 1          # the work did overlap
    for a, g in zip(alone, got):
        # same kernels, same fixed-order sums: bitwise identical
        assert g.niter == a.niter
        assert np.array_equal(g.x, a.x)
        assert a.residual_norm < 1e-8
    assert all(o.residual_norm < 1e-8 for o in others)


def test_admission_respects_budget():
    reqs = [Request(6, 6, 6, 2, seed_A=200 + k) for k in range(6)]
    with Server(nslots=4, **KW) as probe:
        fp = probe.footprint(reqs[0])
    # room for two running requests at a time
    with Server(nslots=4, hbm_budget=int(2.5 * fp), **KW) as s:
        res = [s.wait(t) for t in [s.submit(r) for r in reqs]]
        info = s.info()
    assert info["peak_active"] <= 2
    assert all(r.residual_norm < 1e-8 for r in res)


def test_poisson_burst_completes():
    rng = np.random.default_rng(9)
    reqs = sweep_requests(16, rng, meshes=(4, 6))
    with Server(nslots=4, **KW) as s:
        out = poisson_load(s, reqs, rate=200.0, rng=rng)
    assert out["requests"] == 16
    assert out["max_residual"] < 1e-8
    assert out["p99_ms"] >= out["p50_ms"] > 0


def test_multi_rhs_shares_factor():
    """nrhs right-hand sides on one factor (folded schedule: Dinv folded into
    the upper blocks): the first agrees with a one-RHS request up to the
    re-association of the backward products, every one converges, and the
    same request gives the same bits again."""
    one = Request(6, 6, 6, 2, seed_A=301, seed_b=11)
    three = Request(6, 6, 6, 2, seed_A=301, seed_b=11, nrhs=3)
    with Server(nslots=2, **KW) as s:
        a = s.wait(s.submit(one, want_x=True))
        b = s.wait(s.submit(three, want_x=True))
        c = s.wait(s.submit(three, want_x=True))
    assert np.linalg.norm(a.x - b.x) / np.linalg.norm(a.x) < 1e-10
    assert np.array_equal(b.x, c.x)
    assert b.niter >= 3 * 1 and b.niter >= a.niter
    assert b.residual_norm < 1e-8
