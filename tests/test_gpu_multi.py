"""k right-hand sides of one system sharing its matrix and ILU(0) factor:
the k-vector matvec and apply give every vector the bits of the
single-vector kernels, and the lockstep FGMRES gives every right-hand side
the single solve's iterations and bits (reference: one fgmres_t%solve per
right-hand side, type_fgmres.f90:81-310), and the oracle's solution."""
import ctypes as C

import numpy as np
import pytest

import pyoracle as po
from chidg_amd import euler
from chidg_amd import pattern as pt
from chidg_amd.linsolve import Context, chidg_matrix, chidg_mv, fgmres, precon_ILU0
from helpers import synthetic_system

pytestmark = pytest.mark.gpu

KW = dict(nkrylov=30, tol=0.0, rtol=1e-8, inner_fgmres=False)


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


def vectors(p, k, seed):
    return np.stack([po.fill_vector(p.vec_len, seed + r) for r in range(k)])


@pytest.mark.parametrize("dims,degree", [((4, 3, 3), 1), ((3, 3, 2), 2), ((2, 2, 2), 3), ((5, 1, 1), 0)])
@pytest.mark.parametrize("k", [2, 3, 8, 11])
def test_mv_multi_bits(ctx, dims, degree, k):
    """cdg_mv_multi: every Y_r is cdg_mv(X_r) bit for bit, and the dense product."""
    p, blocks = synthetic_system(*dims, degree, seed=5, periodic=(True, dims[1] > 1, dims[2] > 1))
    A = upload(ctx, p, blocks)
    X = vectors(p, k, 11)
    Y = np.empty_like(X)
    ctx._check(ctx.lib.cdg_mv_multi(ctx.h, A.h, k, X.ctypes.data_as(C.c_void_p), Y.ctypes.data_as(C.c_void_p)),
               "cdg_mv_multi")
    for r in range(k):
        assert np.array_equal(Y[r], chidg_mv(A, X[r])), r
        assert rel(Y[r], po.mv(p, blocks, X[r])) < 1e-13


@pytest.mark.parametrize("dims,degree,precision,tol", [((4, 3, 3), 1, "fp64", 1e-12), ((3, 3, 2), 2, "fp64", 1e-12),
                                                        ((3, 3, 2), 2, "fp32", 1e-5), ((2, 2, 2), 3, "fp64", 1e-12),
                                                        ((2, 2, 2), 3, "fp32", 1e-5), ((3, 2, 2), 2, "bf16", 5e-2),
                                                        ((6, 1, 1), 1, "fp64", 1e-12)])
@pytest.mark.parametrize("k", [2, 5, 8])
def test_apply_multi_bits(ctx, dims, degree, precision, tol, k):
    """cdg_prec_apply_multi: every Z_r is the row-slice apply of V_r alone bit
    for bit (groups of right-hand sides per launch included), and the
    oracle's apply to the precision's tolerance."""
    p, blocks = synthetic_system(*dims, degree, seed=17, periodic=(True, dims[1] > 1, dims[2] > 1))
    A = upload(ctx, p, blocks)
    M = precon_ILU0(precision, schedule="rowslice").init(A)
    M.update(A)
    V = vectors(p, k, 3)
    Z = np.empty_like(V)
    ctx._check(ctx.lib.cdg_prec_apply_multi(ctx.h, M.h, A.h, k, V.ctypes.data_as(C.c_void_p),
                                            Z.ctypes.data_as(C.c_void_p)), "cdg_prec_apply_multi")
    L_ref, D_ref = po.ilu0_factor(p, blocks)
    for r in range(k):
        assert np.array_equal(Z[r], M.apply(A, V[r])), r
        assert rel(Z[r], po.ilu0_apply(p, blocks, L_ref, D_ref, V[r])) < tol


@pytest.mark.parametrize("orth", ["CGS", "MGS"])
@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_solve_multi_matches_single_and_oracle(ctx, orth, precision):
    """4^3 P1 Euler (config 1 system), 5 right-hand sides: each one's
    iterations and bits equal the single solve's under the row-slice
    schedule; FP64 against the oracle (iterations +-1, x within 1e-8)."""
    p, blocks, _, _ = euler.system(4, 1, mach=0.3, seed=1, perturb=0.01, cfl=1.0)
    A = upload(ctx, p, blocks)
    k = 5
    B = vectors(p, k, 40)
    X = np.zeros_like(B)
    ls = fgmres(orthogonalization=orth, **KW)
    M = precon_ILU0(precision).init(A)
    ls.solve_multi(A, X, B, M)
    for r in range(k):
        Ms = precon_ILU0(precision, schedule="rowslice").init(A)
        x1 = np.zeros(p.vec_len)
        ls1 = fgmres(orthogonalization=orth, **KW)
        ls1.solve(A, x1, B[r], Ms)
        assert ls.niter[r] == ls1.niter, (r, ls.niter[r], ls1.niter)
        assert np.array_equal(X[r], x1), r
        assert ls.stats[r].residual_norm == ls1.residual_norm
        Ms.tear_down()
        if precision == "fp64":
            x_ref, n_ref, _, _ = po.fgmres(p, blocks, B[r], nkrylov=30, tol=0.0, rtol=1e-8,
                                           orthogonalization=orth)
            assert abs(ls.niter[r] - n_ref) <= 1
            assert rel(X[r], x_ref) < 1e-8


def test_solve_multi_restarts_and_nmax(ctx):
    """Restarts (nkrylov 5) and an iteration cap (nmax 12) per right-hand
    side: the same iterations and bits as the single solves, where the
    vectors stop at different steps."""
    p, blocks = synthetic_system(4, 4, 2, 1, seed=29, s_off_scale=0.6)
    A = upload(ctx, p, blocks)
    k = 4
    B = vectors(p, k, 60)
    B[2] *= 1e-3                     # a different scale: the same relative stopping point
    for kw in (dict(nkrylov=5, tol=0.0, rtol=1e-9, inner_fgmres=False),
               dict(nkrylov=30, tol=0.0, rtol=1e-12, nmax=12, inner_fgmres=False)):
        X = np.zeros_like(B)
        ls = fgmres(**kw)
        ls.solve_multi(A, X, B, precon_ILU0("fp64").init(A))
        for r in range(k):
            x1 = np.zeros(p.vec_len)
            ls1 = fgmres(**kw)
            ls1.solve(A, x1, B[r], precon_ILU0("fp64", schedule="rowslice").init(A))
            assert ls.niter[r] == ls1.niter and ls.stats[r].nrestart == ls1.stats.nrestart, r
            assert np.array_equal(X[r], x1), r


def test_solve_multi_zero_rhs_and_inner(ctx):
    """A zero right-hand side among others gives x = 0 in 0 iterations; with
    the nested inner FGMRES the vectors are solved one after another, each
    as alone."""
    p, blocks = synthetic_system(3, 3, 3, 1, seed=31)
    A = upload(ctx, p, blocks)
    B = vectors(p, 3, 70)
    B[1] = 0.0
    X = np.zeros_like(B)
    ls = fgmres(**KW)
    ls.solve_multi(A, X, B, precon_ILU0("fp64").init(A))
    assert ls.niter[1] == 0 and not X[1].any()
    assert ls.niter[0] > 0 and ls.niter[2] > 0
    X2 = np.zeros_like(B)
    li = fgmres(nkrylov=30, tol=0.0, rtol=1e-8, inner_fgmres=True)
    li.solve_multi(A, X2, B, precon_ILU0("fp64").init(A))
    for r in (0, 2):
        x1 = np.zeros(p.vec_len)
        l1 = fgmres(nkrylov=30, tol=0.0, rtol=1e-8, inner_fgmres=True)
        l1.solve(A, x1, B[r], precon_ILU0("fp64").init(A))
        assert li.niter[r] == l1.niter
        assert rel(X2[r], x1) < 1e-12


def test_solve_multi_euler_16c_p2(ctx):
    """The headline system, 4 right-hand sides on the device, FP32 ILU:
    every true residual below 1e-8; the first equals the single solve bit
    for bit."""
    import torch
    p, blocks, _, _ = euler.system(16, 2, mach=0.3, seed=1, perturb=0.01, cfl=1.0)
    A = upload(ctx, p, blocks)
    k = 4
    B = torch.from_numpy(np.stack([np.random.default_rng(1 + r).standard_normal(p.vec_len) for r in range(k)]))
    B = B.to("cuda:0")
    X = torch.zeros_like(B)
    ls = fgmres(**KW)
    ls.solve_multi(A, X, B, precon_ILU0("fp32").init(A))
    for r in range(k):
        tr = float(torch.linalg.norm(chidg_mv(A, X[r]) - B[r]) / torch.linalg.norm(B[r]))
        assert tr < 1e-8, (r, tr)
    x1 = torch.zeros_like(B[0])
    ls1 = fgmres(**KW)
    ls1.solve(A, x1, B[0], precon_ILU0("fp32", schedule="rowslice").init(A))
    assert ls.niter[0] == ls1.niter
    assert torch.equal(X[0], x1)
    A.release()
