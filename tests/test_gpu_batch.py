"""A batch of independent systems of uneven sizes and block sizes
(cdg_batch_*): every right-hand side gets the iterations and (to rounding)
the solution of its own lockstep solve on its own system -- one
fgmres_t%solve per right-hand side in the reference, type_fgmres.f90:81-310
-- and the oracle's solution; the graph-captured cycles give the eager
launches' bits on every replay."""
import numpy as np
import pytest

import pyoracle as po
from chidg_amd import pattern as pt
from chidg_amd.linsolve import Context, batch, chidg_matrix, chidg_mv, fgmres, precon_ILU0

pytestmark = pytest.mark.gpu

NMAX = 600    # every solve here converges well inside this; the bound keeps a stalled run finite
KW = dict(nkrylov=30, tol=0.0, rtol=1e-10, nmax=NMAX, inner_fgmres=False)


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(np.asarray(b)), 1e-300))


@pytest.fixture(scope="module")
def ctx():
    c = Context(0)
    yield c
    c.close()


# (nx, ny, nz, degree, nrhs, s_off scale)
SMALL = [(3, 3, 3, 1, 2, 0.1), (4, 4, 4, 2, 1, 0.1), (2, 2, 2, 3, 3, 0.1), (3, 3, 2, 1, 1, 0.3),
         (4, 4, 4, 1, 4, 0.1), (3, 3, 3, 2, 2, 0.2), (2, 2, 2, 0, 5, 0.1)]


def make(ctx, spec, seed0=100):
    mats, rhs, host = [], [], []
    for i, (nx, ny, nz, deg, k, so) in enumerate(spec):
        n = pt.block_size(deg)
        p = pt.hex_box(nx, ny, nz, n)
        blocks = po.fill_synthetic(p, seed0 + i, 1 / np.sqrt(n), so / np.sqrt(n))
        A = chidg_matrix(p, ctx)
        A.store(blocks)
        mats.append(A)
        rhs.append(np.stack([po.fill_vector(p.vec_len, seed0 + 10 * i + r) for r in range(k)]))
        host.append((p, blocks))
    return mats, rhs, host


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_batch_matches_own_solves_and_oracle(ctx, precision):
    mats, rhs, host = make(ctx, SMALL)
    bt = batch(mats, [b.shape[0] for b in rhs], precision=precision, nkrylov=30)
    assert bt.info()["groups"] == 4            # P0, P1, P2, P3
    X = [np.zeros_like(b) for b in rhs]
    st = bt.solve(rhs, X, rtol=KW["rtol"], nmax=NMAX)
    lane = 0
    for s, (A, B) in enumerate(zip(mats, rhs)):
        Xs = np.zeros_like(B)
        ls = fgmres(**KW)
        ls.solve_multi(A, Xs, B, precon_ILU0(precision, schedule="rowslice").init(A))
        for r in range(B.shape[0]):
            assert abs(st[lane].niter - ls.niter[r]) <= 1, (s, r, st[lane].niter, ls.niter[r])
            assert rel(X[s][r], Xs[r]) < 1e-8, (s, r)
            assert st[lane].residual_norm / np.linalg.norm(B[r]) < 1e-9
            if precision == "fp64":
                p, blocks = host[s]
                x_ref, n_ref, _, _ = po.fgmres(p, blocks, B[r], nkrylov=30, tol=0.0, rtol=1e-10, nmax=NMAX)
                assert abs(st[lane].niter - n_ref) <= 1
                assert rel(X[s][r], x_ref) < 1e-8
            lane += 1
    bt.close()


def test_batch_graph_replays_equal_eager(ctx, monkeypatch):
    """The cycles as captured HIP-graph replays and as eager launches give the
    same bits; three solves in a row (each replaying the graph once per
    cycle, eager work and synchronizations in between) all equal the eager one."""
    import torch
    mats, rhs, _ = make(ctx, SMALL[:5], seed0=300)
    bt = batch(mats, [b.shape[0] for b in rhs], precision="fp32", nkrylov=8)   # restarts: several cycles
    monkeypatch.setenv("CDG_BATCH_GRAPH", "0")
    Xe = [np.zeros_like(b) for b in rhs]
    se = bt.solve(rhs, Xe, rtol=1e-9, nmax=NMAX)
    monkeypatch.delenv("CDG_BATCH_GRAPH")
    for rep in range(3):
        Xg = [np.zeros_like(b) for b in rhs]
        sg = bt.solve(rhs, Xg, rtol=1e-9, nmax=NMAX)
        torch.cuda.synchronize()
        assert [s.niter for s in sg] == [s.niter for s in se]
        for a, b in zip(Xg, Xe):
            assert np.array_equal(a, b), rep
        y = chidg_mv(mats[0], Xg[0][0])          # eager work between the replays
        assert np.isfinite(y).all()
    assert max(s.nrestart for s in se) >= 1
    bt.close()


@pytest.mark.parametrize("nmax", [NMAX, 10])
def test_batch_restarts_and_device_vectors(ctx, nmax):
    """nkrylov 6 with harder coupling (the lanes restart at different cycles:
    26, 40 and 81 iterations in the oracle), right-hand sides and solutions
    as device tensors: each lane as its own solve (iterations within 1,
    solution to 1e-8).  With nmax 10 every lane stops inside its second cycle
    (the iteration-count stop), and x is that of its own solve stopped there."""
    import torch
    spec = [(3, 3, 3, 1, 3, 0.4), (4, 3, 2, 2, 2, 0.5), (2, 2, 2, 3, 1, 0.6)]
    mats, rhs, _ = make(ctx, spec, seed0=500)
    bt = batch(mats, [b.shape[0] for b in rhs], precision="fp64", nkrylov=6)
    Bd = [torch.from_numpy(b).to("cuda:0") for b in rhs]
    Xd = [torch.zeros_like(b) for b in Bd]
    st = bt.solve(Bd, Xd, rtol=1e-10, nmax=nmax)
    lane = 0
    for s, (A, B) in enumerate(zip(mats, rhs)):
        Xs = np.zeros_like(B)
        ls = fgmres(nkrylov=6, tol=0.0, rtol=1e-10, nmax=nmax, inner_fgmres=False)
        ls.solve_multi(A, Xs, B, precon_ILU0("fp64", schedule="rowslice").init(A))
        for r in range(B.shape[0]):
            if nmax == 10:
                assert st[lane].niter == ls.niter[r] == 10 and st[lane].nrestart == 1
            else:
                assert abs(st[lane].niter - ls.niter[r]) <= 1 and st[lane].niter < nmax
            assert rel(Xd[s][r].cpu().numpy(), Xs[r]) < 1e-8
            lane += 1
    assert max(x.nrestart for x in st) >= 1
    bt.close()


def test_batch_64_uneven_config4(ctx):
    """The config-4 batch at its real size (64 seeded systems, meshes 4^3..12^3,
    P1..P3, 1..8 right-hand sides, FP32 ILU): every true residual below 1e-8,
    and sampled systems take the iterations of their own lockstep solve."""
    import torch
    from chidg_amd.serve import sweep_requests
    rng = np.random.default_rng(20261202)
    reqs = sweep_requests(64, rng, meshes=(4, 12), orders=(1, 2, 3), nrhs=(1, 8))
    dev = torch.empty(0, dtype=torch.float64, device="cuda:0")
    mats, B = [], []
    for r in reqs:
        A = chidg_matrix(pt.hex_box(r.nx, r.ny, r.nz, r.n), ctx)
        A.fill_synthetic(r.seed_A, 1 / np.sqrt(r.n), r.s_off)
        mats.append(A)
        B.append(torch.stack([A.synthetic_vector(r.seed_b + k, like=dev) for k in range(r.nrhs)]))
    bt = batch(mats, [r.nrhs for r in reqs], precision="fp32", nkrylov=30)
    X = [torch.zeros_like(b) for b in B]
    st = bt.solve(B, X, rtol=1e-8, nmax=NMAX)
    lane = 0
    checked = 0
    for s, (A, Bs) in enumerate(zip(mats, B)):
        for r in range(Bs.shape[0]):
            assert st[lane].residual_norm / float(torch.linalg.norm(Bs[r])) < 1e-8, (s, r)
            lane += 1
        if s % 9 == 0:
            Xs = torch.zeros_like(Bs)
            ls = fgmres(nkrylov=30, tol=0.0, rtol=1e-8, nmax=NMAX, inner_fgmres=False)
            ls.solve_multi(A, Xs, Bs, precon_ILU0("fp32", schedule="rowslice").init(A))
            first = lane - Bs.shape[0]
            for r in range(Bs.shape[0]):
                assert abs(st[first + r].niter - ls.niter[r]) <= 1
            checked += 1
    assert checked >= 7
    bt.close()
    for A in mats:
        A.release()


def test_batch_update_equals_fresh_batch(ctx):
    """cdg_batch_update without matrices refactors in place (the same bits);
    with new blocks (the same patterns, other seeds: the next Newton step's
    Jacobians) it gives the bits of a batch created from the new matrices,
    and each lane the iterations of its own solve on them."""
    mats, rhs, _ = make(ctx, SMALL[:5], seed0=700)
    nrhs = [b.shape[0] for b in rhs]
    bt = batch(mats, nrhs, precision="fp32", nkrylov=30)
    X0 = [np.zeros_like(b) for b in rhs]
    s0 = bt.solve(rhs, X0, rtol=1e-9, nmax=NMAX)
    bt.update()                                           # refactored in place: the same bits
    Xr = [np.zeros_like(b) for b in rhs]
    sr = bt.solve(rhs, Xr, rtol=1e-9, nmax=NMAX)
    assert [s.niter for s in sr] == [s.niter for s in s0]
    assert all(np.array_equal(a, b) for a, b in zip(Xr, X0))
    new, _, _ = make(ctx, SMALL[:5], seed0=900)          # same patterns, other blocks
    for A, An in zip(mats, new):
        A.store(An.to_host())
    bt.update(mats)
    X1 = [np.zeros_like(b) for b in rhs]
    s1 = bt.solve(rhs, X1, rtol=1e-9, nmax=NMAX)
    fresh = batch(new, nrhs, precision="fp32", nkrylov=30)
    X2 = [np.zeros_like(b) for b in rhs]
    s2 = fresh.solve(rhs, X2, rtol=1e-9, nmax=NMAX)
    assert [s.niter for s in s1] == [s.niter for s in s2]
    for a, b in zip(X1, X2):
        assert np.array_equal(a, b)
    assert any(not np.allclose(a, b) for a, b in zip(X0, X1))     # the update took effect
    lane = 0
    for A, B in zip(new, rhs):
        Xs = np.zeros_like(B)
        ls = fgmres(nkrylov=30, tol=0.0, rtol=1e-9, nmax=NMAX, inner_fgmres=False)
        ls.solve_multi(A, Xs, B, precon_ILU0("fp32", schedule="rowslice").init(A))
        for r in range(B.shape[0]):
            assert abs(s1[lane].niter - ls.niter[r]) <= 1
            assert rel(X1[lane_sys(nrhs, lane)][r], Xs[r]) < 1e-8
            lane += 1
    with pytest.raises(Exception):
        bt.update(list(reversed(mats)))                       # another pattern: refused
    bt.close()
    fresh.close()


def lane_sys(nrhs, lane):
    s = 0
    while lane >= nrhs[s]:
        lane -= nrhs[s]
        s += 1
    return s
