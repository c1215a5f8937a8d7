"""The gfx950 path on DG Euler first-Newton-step systems (chidg_amd/euler.py)
against the oracle on the same blocks.  Tolerances as for the synthetic
systems: matvec 1e-13 relative, FP64 solution 1e-8 relative, iteration
counts within +-1; FP32/BF16-stored ILU reach the same residual with FP64
Krylov vectors."""
import numpy as np
import pytest

import pyoracle as po
from chidg_amd import euler
from chidg_amd.linsolve import Context, chidg_matrix, chidg_mv, fgmres, precon_ILU0
from helpers import dense

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(np.asarray(b)), 1e-300))


@pytest.fixture(scope="module")
def ctx():
    c = Context(0)
    yield c
    c.close()


@pytest.fixture(scope="module")
def config1():
    """4^3 periodic P1 Euler, Mach 0.3 + N(0, 0.01) perturbation, CFL 1."""
    p, blocks, R, _ = euler.system(4, 1, mach=0.3, seed=1, perturb=0.01, cfl=1.0)
    return p, blocks, R


def test_euler_mv(ctx, config1):
    p, blocks, _ = config1
    A = chidg_matrix(p, ctx)
    A.store(blocks)
    x = np.random.default_rng(5).standard_normal(p.vec_len)
    assert rel(chidg_mv(A, x), dense(p, blocks) @ x) < 1e-13


def test_euler_ilu0_apply(ctx, config1):
    """Factor (checked-diagonal inverse with its partial-pivoting fallback)
    and apply against the oracle."""
    p, blocks, R = config1
    A = chidg_matrix(p, ctx)
    A.store(blocks)
    M = precon_ILU0().init(A)
    M.update(A)
    L_ref, D_ref = po.ilu0_factor(p, blocks)
    z_ref = po.ilu0_apply(p, blocks, L_ref, D_ref, -R)
    assert rel(M.apply(A, -R), z_ref) < 1e-12


@pytest.mark.parametrize("schedule", ["levels", "folded", "stream", "auto"])
def test_euler_fgmres_config1(ctx, config1, schedule):
    """b = -R: FGMRES(30)+ILU(0) to rtol 1e-8, FP64, against the oracle."""
    p, blocks, R = config1
    A = chidg_matrix(p, ctx)
    A.store(blocks)
    b = -R
    ls = fgmres(nkrylov=30, tol=0.0, rtol=1e-8, inner_fgmres=False)
    x = np.zeros(p.vec_len)
    ls.solve(A, x, b, precon_ILU0(schedule=schedule))
    x_ref, n_ref, _, _ = po.fgmres(p, blocks, b, nkrylov=30, tol=0.0, rtol=1e-8)
    assert abs(ls.niter - n_ref) <= 1
    assert rel(x, x_ref) < 1e-8


@pytest.mark.parametrize("precision", ["fp32", "bf16"])
def test_euler_mixed_precision(ctx, config1, precision):
    p, blocks, R = config1
    A = chidg_matrix(p, ctx)
    A.store(blocks)
    b = -R
    x64 = np.zeros(p.vec_len)
    fgmres(nkrylov=30, tol=0.0, rtol=1e-8, inner_fgmres=False).solve(A, x64, b, precon_ILU0("fp64"))
    ls = fgmres(nkrylov=30, tol=0.0, rtol=1e-8, inner_fgmres=False)
    x = np.zeros(p.vec_len)
    ls.solve(A, x, b, precon_ILU0(precision))
    assert ls.residual_norm / np.linalg.norm(b) < 1e-8
    # both solutions have relative residual <= 1e-8; cond(A) ~ 2e2 bounds their distance
    assert rel(x, x64) < 1e-5


def test_euler_p2_8cube(ctx):
    """8^3 P2 (135 x 135 blocks), RHS N(0,1): FP64 against the oracle, FP32 ILU
    to the same residual."""
    p, blocks, _, _ = euler.system(8, 2, mach=0.3, seed=2, perturb=0.01, cfl=1.0)
    A = chidg_matrix(p, ctx)
    A.store(blocks)
    b = np.random.default_rng(11).standard_normal(p.vec_len)
    x_ref, n_ref, _, _ = po.fgmres(p, blocks, b, nkrylov=30, tol=0.0, rtol=1e-8)
    for precision in ("fp64", "fp32"):
        ls = fgmres(nkrylov=30, tol=0.0, rtol=1e-8, inner_fgmres=False)
        x = np.zeros(p.vec_len)
        ls.solve(A, x, b, precon_ILU0(precision))
        assert ls.residual_norm / np.linalg.norm(b) < 1e-8
        if precision == "fp64":
            assert abs(ls.niter - n_ref) <= 1
            assert rel(x, x_ref) < 1e-8
        else:
            assert rel(x, x_ref) < 1e-5


@pytest.mark.parametrize("orth,inner", [("MGS", False), ("CGS", True), ("MGS", True)])
def test_euler_fgmres_orthogonalization_and_inner(ctx, config1, orth, inner):
    """b = -R with modified Gram-Schmidt, and with the reference's default inner
    FGMRES (nkrylov 100, tol and rtol 1e-1, preconditioned by the same ILU(0))
    on each outer direction: FP64 against the oracle with the same parameters."""
    p, blocks, R = config1
    A = chidg_matrix(p, ctx)
    A.store(blocks)
    b = -R
    ls = fgmres(nkrylov=30, tol=0.0, rtol=1e-8, orthogonalization=orth, inner_fgmres=inner)
    x = np.zeros(p.vec_len)
    ls.solve(A, x, b, precon_ILU0())
    x_ref, n_ref, _, _ = po.fgmres(p, blocks, b, nkrylov=30, tol=0.0, rtol=1e-8, orthogonalization=orth,
                                   inner_fgmres=inner)
    assert abs(ls.niter - n_ref) <= 1
    assert rel(x, x_ref) < 1e-8
    assert ls.residual_norm / np.linalg.norm(b) < 1e-8


def test_euler_p3_3cube(ctx):
    """3^3 P3 (320 x 320 blocks: the global-memory block inverse and its
    partial-pivoting fallback on Euler diagonal blocks), RHS N(0,1), 58 oracle
    iterations (one restart of FGMRES(30)): FP64 against the oracle, FP32 ILU to
    the same residual."""
    p, blocks, _, _ = euler.system(3, 3, mach=0.3, seed=3, perturb=0.01, cfl=1.0)
    A = chidg_matrix(p, ctx)
    A.store(blocks)
    b = np.random.default_rng(13).standard_normal(p.vec_len)
    x_ref, n_ref, _, _ = po.fgmres(p, blocks, b, nkrylov=30, tol=0.0, rtol=1e-8)
    for precision in ("fp64", "fp32"):
        ls = fgmres(nkrylov=30, tol=0.0, rtol=1e-8, inner_fgmres=False)
        x = np.zeros(p.vec_len)
        ls.solve(A, x, b, precon_ILU0(precision))
        assert ls.residual_norm / np.linalg.norm(b) < 1e-8
        if precision == "fp64":
            assert abs(ls.niter - n_ref) <= 1
            assert rel(x, x_ref) < 1e-8
        else:
            assert rel(x, x_ref) < 1e-5
