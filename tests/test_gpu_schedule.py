"""The ILU(0) apply under its two schedules (one launch per wavefront level,
or one persistent dataflow launch) against the oracle, against each other,
and across repeated calls (the dataflow counters carry over between calls)."""
import numpy as np
import pytest

import pyoracle as po
from chidg_amd import pattern as pt
from chidg_amd.linsolve import Context, chidg_matrix, fgmres, precon_ILU0
from helpers import ilu_test_system, synthetic_system

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(np.asarray(b)), 1e-300))


@pytest.fixture(scope="module")
def ctx():
    c = Context(0)
    yield c
    c.close()


def upload(ctx, p, blocks):
    A = chidg_matrix(p, ctx)
    A.store(blocks)
    return A


def test_dataflow_known_answer(ctx):
    """test_ILU0_mod.pf:294-332 through the dataflow sweep (1x1 blocks)."""
    p, blocks, b = ilu_test_system()
    A = upload(ctx, p, blocks)
    M = precon_ILU0(schedule="dataflow").init(A)
    M.update(A, b)
    for _ in range(3):
        np.testing.assert_allclose(M.apply(A, b), [43 / 7, 18 / 7, 3 / 14], rtol=1e-14)


# (dims, degree, precision, tolerance against the FP64 oracle)
CASES = [((3, 3, 3), 1, "fp64", 1e-12), ((4, 3, 2), 2, "fp64", 1e-12), ((5, 2, 2), 0, "fp64", 1e-12),
         ((4, 4, 4), 1, "fp32", 1e-5), ((3, 3, 3), 2, "fp32", 1e-5), ((3, 3, 3), 2, "bf16", 5e-2)]


@pytest.mark.parametrize("dims,degree,precision,tol", CASES)
def test_dataflow_matches_oracle_and_levels(ctx, dims, degree, precision, tol):
    p, blocks = synthetic_system(*dims, degree, seed=17)
    A = upload(ctx, p, blocks)
    Ml = precon_ILU0(precision, schedule="levels").init(A)
    Md = precon_ILU0(precision, schedule="dataflow").init(A)
    Ml.update(A)
    Md.update(A)
    L_ref, D_ref = po.ilu0_factor(p, blocks)
    for seed in (3, 4):
        v = po.fill_vector(p.vec_len, seed)
        z_ref = po.ilu0_apply(p, blocks, L_ref, D_ref, v)
        zd = Md.apply(A, v)
        assert rel(zd, z_ref) < tol
        # the two schedules differ only in how partial sums are grouped
        assert rel(zd, Ml.apply(A, v)) < (1e-13 if precision == "fp64" else tol)
        # repeated calls: bitwise identical (fixed-order sums, counters carried over)
        assert np.array_equal(Md.apply(A, v), zd)


def test_dataflow_fgmres_config1(ctx):
    """4^3 periodic P1, FGMRES(30)+ILU0 to rtol 1e-8 with the dataflow sweep."""
    p, blocks = synthetic_system(4, 4, 4, 1, seed=20261202)
    b = po.fill_vector(p.vec_len, 1)
    A = upload(ctx, p, blocks)
    ls = fgmres(nkrylov=30, tol=1e-30, rtol=1e-8, inner_fgmres=False)
    x = np.zeros(p.vec_len)
    ls.solve(A, x, b, precon_ILU0(schedule="dataflow"))
    x_ref, n_ref, _, _ = po.fgmres(p, blocks, b, nkrylov=30, tol=1e-30, rtol=1e-8)
    assert abs(ls.niter - n_ref) <= 1
    assert rel(x, x_ref) < 1e-8


def test_dataflow_full_size(ctx):
    """16^3 periodic P2 (4096 block rows): dataflow and level sweeps agree, and
    an FGMRES(30) solve with the FP32 dataflow sweep reaches rtol 1e-8."""
    import torch
    n = pt.block_size(2)
    p = pt.hex_box(16, 16, 16, n)
    A = chidg_matrix(p, ctx)
    A.fill_synthetic(20261202, 1 / np.sqrt(n), 0.1 / np.sqrt(n))
    dev = torch.empty(0, dtype=torch.float64, device="cuda:0")
    v = A.synthetic_vector(1, like=dev)
    for precision, tol in (("fp64", 1e-13), ("fp32", 1e-6)):
        Ml = precon_ILU0(precision, schedule="levels").init(A)
        Md = precon_ILU0(precision, schedule="dataflow").init(A)
        Ml.update(A)
        Md.update(A)
        zl = Ml.apply(A, v)
        zd = Md.apply(A, v)
        assert float(torch.linalg.norm(zd - zl) / torch.linalg.norm(zl)) < tol
        assert torch.equal(Md.apply(A, v), zd)
        Ml.tear_down()
        if precision == "fp32":
            x = torch.zeros_like(v)
            ls = fgmres(nkrylov=30, tol=1e-30, rtol=1e-8, inner_fgmres=False)
            ls.solve(A, x, v, Md)
            assert ls.residual_norm / float(torch.linalg.norm(v)) < 1e-8
        Md.tear_down()
    A.release()


def test_dataflow_gives_up_cleanly(ctx, monkeypatch):
    """A wait that exceeds its spin limit (forced from the host: CDG_DF_SPIN=1)
    ends the sweep normally: the call reports an error, the counters are
    reset, and a fresh preconditioner on the same context works."""
    p, blocks = synthetic_system(4, 4, 4, 1, seed=17)
    A = upload(ctx, p, blocks)
    v = po.fill_vector(p.vec_len, 3)
    monkeypatch.setenv("CDG_DF_SPIN", "1")
    M = precon_ILU0(schedule="dataflow").init(A)
    M.update(A)
    with pytest.raises(Exception):
        for _ in range(20):          # a run may finish before any wait starts; many will not
            M.apply(A, v)
    M.tear_down()
    monkeypatch.delenv("CDG_DF_SPIN")
    M2 = precon_ILU0(schedule="dataflow").init(A)
    M2.update(A)
    L_ref, D_ref = po.ilu0_factor(p, blocks)
    assert rel(M2.apply(A, v), po.ilu0_apply(p, blocks, L_ref, D_ref, v)) < 1e-12


@pytest.mark.parametrize("dims,degree,precision,tol", [c for c in CASES if c[2] != "bf16"])
def test_folded_matches_oracle_and_levels(ctx, dims, degree, precision, tol):
    """Dinv folded into the upper blocks (W = U Dinv formed by the factor):
    same apply as the oracle and the level schedule up to rounding, and
    bitwise repeatable."""
    p, blocks = synthetic_system(*dims, degree, seed=17)
    A = upload(ctx, p, blocks)
    Ml = precon_ILU0(precision, schedule="levels").init(A)
    Mf = precon_ILU0(precision, schedule="folded").init(A)
    Ml.update(A)
    Mf.update(A)
    L_ref, D_ref = po.ilu0_factor(p, blocks)
    for seed in (3, 4):
        v = po.fill_vector(p.vec_len, seed)
        z_ref = po.ilu0_apply(p, blocks, L_ref, D_ref, v)
        zf = Mf.apply(A, v)
        assert rel(zf, z_ref) < tol
        assert rel(zf, Ml.apply(A, v)) < (1e-13 if precision == "fp64" else tol)
        assert np.array_equal(Mf.apply(A, v), zf)


def test_folded_schedule_set_after_factor(ctx):
    """Choosing the folded schedule after an update forms W at the next apply."""
    p, blocks = synthetic_system(4, 3, 2, 1, seed=5)
    A = upload(ctx, p, blocks)
    M = precon_ILU0("fp64").init(A)
    M.update(A)
    v = po.fill_vector(p.vec_len, 9)
    z_levels = M.apply(A, v)
    M.set_schedule("folded")
    assert rel(M.apply(A, v), z_levels) < 1e-13
    M.update(A)                      # a new factor forms W with it
    assert rel(M.apply(A, v), z_levels) < 1e-13


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-13), ("fp32", 1e-6), ("bf16", 1e-2)])
def test_auto_schedule_streams_after_12_applies(ctx, precision, tol):
    """The default schedule: the first 12 applies after a factor are the level
    schedule's bit for bit, the 13th on the row-slice schedule's (the stream
    sweep's bits: the same tasks and sums); after a factor
    that served more than 12 applies the next switches at its first apply,
    after one that served fewer it waits for 12 again; choosing 'levels' gives
    the level sweep."""
    p, blocks = synthetic_system(4, 3, 2, 2, seed=21)
    A = upload(ctx, p, blocks)
    Ma = precon_ILU0(precision).init(A)
    Ml = precon_ILU0(precision, schedule="levels").init(A)
    Ms = precon_ILU0(precision, schedule="stream").init(A)
    for M in (Ma, Ml, Ms):
        M.update(A)
    v = po.fill_vector(p.vec_len, 6)
    zl, zs = Ml.apply(A, v), Ms.apply(A, v)
    assert rel(zs, zl) < tol
    for k in range(12):
        assert np.array_equal(Ma.apply(A, v), zl), k
    for k in range(3):
        assert np.array_equal(Ma.apply(A, v), zs), k
    Ma.update(A)                    # the last factor served 15 applies
    for k in range(3):
        assert np.array_equal(Ma.apply(A, v), zs), k
    Ma.update(A)                    # this one served 3
    for k in range(12):
        assert np.array_equal(Ma.apply(A, v), zl), k
    assert np.array_equal(Ma.apply(A, v), zs)
    Ma.update(A)
    Ma.set_schedule("levels")
    for _ in range(14):
        assert np.array_equal(Ma.apply(A, v), zl)


def test_folded_fgmres_full_size(ctx):
    """16^3 periodic P2, FP32 ILU, folded and auto schedules: FGMRES(30) reaches
    rtol 1e-8 and matches the level schedule's solution."""
    import torch
    n = pt.block_size(2)
    p = pt.hex_box(16, 16, 16, n)
    A = chidg_matrix(p, ctx)
    A.fill_synthetic(20261202, 1 / np.sqrt(n), 0.1 / np.sqrt(n))
    dev = torch.empty(0, dtype=torch.float64, device="cuda:0")
    b = A.synthetic_vector(1, like=dev)
    xs = {}
    for sched in ("levels", "folded", "stream", "auto"):
        x = torch.zeros_like(b)
        ls = fgmres(nkrylov=30, tol=1e-30, rtol=1e-8, inner_fgmres=False)
        M = precon_ILU0("fp32", schedule=sched)
        ls.solve(A, x, b, M)
        assert ls.residual_norm / float(torch.linalg.norm(b)) < 1e-8
        xs[sched] = x
        M.tear_down()
    for sched in ("folded", "stream", "auto"):
        assert float(torch.linalg.norm(xs[sched] - xs["levels"]) / torch.linalg.norm(xs["levels"])) < 1e-10
    A.release()


# ------------------------------------------------------------ stream schedule
def test_stream_known_answer(ctx):
    """test_ILU0_mod.pf:294-332 through the stream sweep (1x1 blocks)."""
    p, blocks, b = ilu_test_system()
    A = upload(ctx, p, blocks)
    M = precon_ILU0(schedule="stream").init(A)
    M.update(A, b)
    for _ in range(3):
        np.testing.assert_allclose(M.apply(A, b), [43 / 7, 18 / 7, 3 / 14], rtol=1e-14)


STREAM_CASES = CASES + [((2, 2, 2), 3, "fp64", 1e-12), ((3, 2, 2), 3, "fp32", 1e-5), ((3, 2, 2), 3, "bf16", 5e-2),
                        ((6, 1, 1), 1, "fp64", 1e-12)]


@pytest.mark.parametrize("dims,degree,precision,tol", STREAM_CASES)
def test_stream_matches_oracle_and_levels(ctx, dims, degree, precision, tol):
    """Row-slice sweep with W = Dinv U: the oracle's apply to the stated
    tolerance, the level schedule's to rounding (FP64), bitwise repeatable
    (the exchange tags change every call, the sums keep their order)."""
    p, blocks = synthetic_system(*dims, degree, seed=17, periodic=(True, dims[1] > 1, dims[2] > 1))
    A = upload(ctx, p, blocks)
    Ml = precon_ILU0(precision, schedule="levels").init(A)
    Ms = precon_ILU0(precision, schedule="stream").init(A)
    Ml.update(A)
    Ms.update(A)
    L_ref, D_ref = po.ilu0_factor(p, blocks)
    for seed in (3, 4):
        v = po.fill_vector(p.vec_len, seed)
        z_ref = po.ilu0_apply(p, blocks, L_ref, D_ref, v)
        zs = Ms.apply(A, v)
        assert rel(zs, z_ref) < tol
        assert rel(zs, Ml.apply(A, v)) < (1e-13 if precision == "fp64" else tol)
        for _ in range(3):
            assert np.array_equal(Ms.apply(A, v), zs)


@pytest.mark.parametrize("dims,degree,precision,tol", STREAM_CASES)
def test_rowslice_matches_oracle_and_stream(ctx, dims, degree, precision, tol):
    """The row-slice tasks one launch per level: the oracle's apply to the
    stated tolerance, and bit for bit the persistent sweep's result (same
    products, same sums, same roundings), on every repeat."""
    p, blocks = synthetic_system(*dims, degree, seed=17, periodic=(True, dims[1] > 1, dims[2] > 1))
    A = upload(ctx, p, blocks)
    Mr = precon_ILU0(precision, schedule="rowslice").init(A)
    Ms = precon_ILU0(precision, schedule="stream").init(A)
    Mr.update(A)
    Ms.update(A)
    L_ref, D_ref = po.ilu0_factor(p, blocks)
    for seed in (3, 4):
        v = po.fill_vector(p.vec_len, seed)
        z_ref = po.ilu0_apply(p, blocks, L_ref, D_ref, v)
        zr = Mr.apply(A, v)
        assert rel(zr, z_ref) < tol
        assert np.array_equal(zr, Ms.apply(A, v))
        for _ in range(2):
            assert np.array_equal(Mr.apply(A, v), zr)


def test_stream_refactor_and_schedule_switch(ctx):
    """A new factor re-forms W and the row-slice copies at its first stream
    apply; switching schedules back and forth keeps the results."""
    p, blocks = synthetic_system(4, 3, 2, 2, seed=23)
    A = upload(ctx, p, blocks)
    M = precon_ILU0("fp64", schedule="stream").init(A)
    M.update(A)
    v = po.fill_vector(p.vec_len, 9)
    L_ref, D_ref = po.ilu0_factor(p, blocks)
    z_ref = po.ilu0_apply(p, blocks, L_ref, D_ref, v)
    z1 = M.apply(A, v)
    assert rel(z1, z_ref) < 1e-12
    blocks2 = blocks * 1.5
    A.store(blocks2)
    M.update(A)
    L2, D2 = po.ilu0_factor(p, blocks2)
    assert rel(M.apply(A, v), po.ilu0_apply(p, blocks2, L2, D2, v)) < 1e-12
    M.set_schedule("levels")
    z_lv = M.apply(A, v)
    M.set_schedule("stream")
    z_st = M.apply(A, v)
    assert rel(z_st, z_lv) < 1e-13
    M.set_schedule("rowslice")
    assert np.array_equal(M.apply(A, v), z_st)
    A.store(blocks)                  # a new factor while the row-slice schedule is set
    M.update(A)
    assert np.array_equal(M.apply(A, v), z1)


def test_stream_fgmres_full_size_all_precisions(ctx):
    """16^3 periodic P2: the stream sweep agrees with the level sweep (FP64 to
    rounding; FP32 and BF16 to their storage precision), repeats bitwise, and
    FGMRES(30) reaches rtol 1e-8 with it in every precision."""
    import torch
    n = pt.block_size(2)
    p = pt.hex_box(16, 16, 16, n)
    A = chidg_matrix(p, ctx)
    A.fill_synthetic(20261202, 1 / np.sqrt(n), 0.1 / np.sqrt(n))
    dev = torch.empty(0, dtype=torch.float64, device="cuda:0")
    v = A.synthetic_vector(1, like=dev)
    for precision, tol in (("fp64", 1e-13), ("fp32", 1e-6), ("bf16", 1e-2)):
        Ml = precon_ILU0(precision, schedule="levels").init(A)
        Ms = precon_ILU0(precision, schedule="stream").init(A)
        Ml.update(A)
        Ms.update(A)
        zl = Ml.apply(A, v)
        zs = Ms.apply(A, v)
        assert float(torch.linalg.norm(zs - zl) / torch.linalg.norm(zl)) < tol, precision
        assert torch.equal(Ms.apply(A, v), zs)
        Ml.tear_down()
        Ms.set_schedule("rowslice")
        assert torch.equal(Ms.apply(A, v), zs)
        x = torch.zeros_like(v)
        ls = fgmres(nkrylov=30, tol=1e-30, rtol=1e-8, inner_fgmres=False)
        ls.solve(A, x, v, Ms)
        assert ls.residual_norm / float(torch.linalg.norm(v)) < 1e-8
        Ms.tear_down()
    A.release()


def test_stream_gives_up_cleanly(ctx, monkeypatch):
    """A wait that exceeds its spin limit (forced from the host: CDG_RS_SPIN=1)
    ends the sweep normally: every workgroup leaves, the call reports an
    error, and later calls on a fresh preconditioner are right."""
    p, blocks = synthetic_system(4, 4, 4, 1, seed=17)
    A = upload(ctx, p, blocks)
    v = po.fill_vector(p.vec_len, 3)
    monkeypatch.setenv("CDG_RS_SPIN", "1")
    M = precon_ILU0(schedule="stream").init(A)
    M.update(A)
    with pytest.raises(Exception):
        for _ in range(20):          # a call may find every input ready; most will not
            M.apply(A, v)
    M.tear_down()
    monkeypatch.delenv("CDG_RS_SPIN")
    M2 = precon_ILU0(schedule="stream").init(A)
    M2.update(A)
    L_ref, D_ref = po.ilu0_factor(p, blocks)
    assert rel(M2.apply(A, v), po.ilu0_apply(p, blocks, L_ref, D_ref, v)) < 1e-12
