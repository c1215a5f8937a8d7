"""Partitioned ILU(0) (the reference's multi-rank semantics: a block enters
the factor and the sweeps only when its parent row is on the row's rank,
precon_ILU0.f90:180,222,344,368).  Oracle checks on the CPU against closed
forms; GPU parity against the oracle with the same partition."""
import numpy as np
import pytest

import pyoracle as po
from chidg_amd import pattern as pt
from helpers import dense, synthetic_system


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(np.asarray(b)), 1e-300))


def test_one_rank_equals_no_partition():
    p, blocks = synthetic_system(3, 3, 2, 1, seed=61)
    L0, D0 = po.ilu0_factor(p, blocks)
    L1, D1 = po.ilu0_factor(p, blocks, part=np.zeros(p.nrows, np.int32))
    assert np.array_equal(L0, L1) and np.array_equal(D0, D1)


def test_every_row_its_own_rank_is_block_jacobi():
    """nrows ranks: no lower or upper block survives, M = blockdiag(A_ii)."""
    p, blocks = synthetic_system(3, 2, 2, 1, seed=62)
    part = np.arange(p.nrows, dtype=np.int32)
    L, D = po.ilu0_factor(p, blocks, part=part)
    assert not L.any()
    v = po.fill_vector(p.vec_len, 3)
    z = po.ilu0_apply(p, blocks, L, D, v, part=part)
    n = p.n
    B = blocks.reshape(p.nnzb, n, n)
    for i in range(p.nrows):
        b = [k for k in range(p.row_ptr[i], p.row_ptr[i + 1]) if p.kind[k] == 0 and p.col[k] == i][0]
        zi = np.linalg.solve(B[b].T, v[i * n:(i + 1) * n])
        assert rel(z[i * n:(i + 1) * n], zi) < 1e-12


def test_chain_slabs_give_exact_slab_inverses():
    """A 1-D chain split into two slabs: ILU(0) of a chain is its exact block
    LU, so the partitioned preconditioner is the inverse of each slab's
    diagonal part of A (the coupling between the slabs dropped)."""
    p, blocks = synthetic_system(8, 1, 1, 1, seed=63, periodic=(False, False, False))
    part = np.array([0] * 3 + [1] * 5, dtype=np.int32)
    L, D = po.ilu0_factor(p, blocks, part=part)
    v = po.fill_vector(p.vec_len, 4)
    z = po.ilu0_apply(p, blocks, L, D, v, part=part)
    Ad = dense(p, blocks, kinds=(0,))
    n = p.n
    cut = 3 * n
    Ad[:cut, cut:] = 0.0
    Ad[cut:, :cut] = 0.0
    assert rel(z, np.linalg.solve(Ad, v)) < 1e-11
    # the matvec still uses the coupling between the slabs
    assert rel(po.mv(p, blocks, v), dense(p, blocks) @ v) < 1e-13


def test_partitioned_fgmres_converges_with_more_iterations():
    p, blocks = synthetic_system(4, 4, 4, 1, seed=64, s_off_scale=0.3)
    b = po.fill_vector(p.vec_len, 5)
    part = (np.arange(p.nrows) // 16).astype(np.int32)     # 4 slabs of 4x4 elements
    x1, n1, r1, _ = po.fgmres(p, blocks, b, nkrylov=30, tol=0.0, rtol=1e-8)
    x4, n4, r4, _ = po.fgmres(p, blocks, b, nkrylov=30, tol=0.0, rtol=1e-8, part=part)
    assert r4 / np.linalg.norm(b) < 1e-8 and n4 >= n1
    assert rel(x4, x1) < 1e-6


# ------------------------------------------------------------------- GPU
@pytest.mark.gpu
@pytest.mark.parametrize("schedule", ["levels", "stream"])
@pytest.mark.parametrize("precision,tol", [("fp64", 1e-12), ("fp32", 1e-5)])
def test_gpu_partitioned_ilu_matches_oracle(schedule, precision, tol):
    from chidg_amd.linsolve import Context, chidg_matrix, chidg_mv, fgmres, precon_ILU0
    ctx = Context(0)
    p, blocks = synthetic_system(4, 4, 4, 2, seed=65)
    part = (np.arange(p.nrows) // 32).astype(np.int32)     # 2 slabs in z
    A = chidg_matrix(p, ctx)
    A.store(blocks)
    A.set_partition(part)
    M = precon_ILU0(precision, schedule=schedule).init(A)
    M.update(A)
    L_ref, D_ref = po.ilu0_factor(p, blocks, part=part)
    v = po.fill_vector(p.vec_len, 6)
    assert rel(M.apply(A, v), po.ilu0_apply(p, blocks, L_ref, D_ref, v, part=part)) < tol
    if precision == "fp64":
        _, D = M.factors(A)
        assert rel(D, D_ref) < 1e-12
    x = np.zeros(p.vec_len)
    ls = fgmres(nkrylov=30, tol=0.0, rtol=1e-8, inner_fgmres=False)
    ls.solve(A, x, v, M)
    x_ref, n_ref, _, _ = po.fgmres(p, blocks, v, nkrylov=30, tol=0.0, rtol=1e-8, part=part)
    assert ls.residual_norm / np.linalg.norm(v) < 1e-8
    if precision == "fp64":
        assert abs(ls.niter - n_ref) <= 1
        assert rel(x, x_ref) < 1e-8
    # the matvec keeps the cross-slab blocks
    assert rel(chidg_mv(A, v), po.mv(p, blocks, v)) < 1e-13
    M.tear_down()
    A.release()
    ctx.close()


# ------------------------------------------------------ block Jacobi (GPU)
@pytest.mark.gpu
@pytest.mark.parametrize("schedule", ["levels", "stream"])
@pytest.mark.parametrize("precision,tol", [("fp64", 1e-12), ("fp32", 1e-5), ("bf16", 5e-2)])
def test_gpu_block_jacobi(schedule, precision, tol):
    """precon_jacobi (precon_jacobi.f90:95-215): z_i = inv(A_ii) v_i per block
    row, against numpy; FGMRES with it reaches rtol 1e-8."""
    from chidg_amd.linsolve import Context, chidg_matrix, create_preconditioner, fgmres
    ctx = Context(0)
    p, blocks = synthetic_system(3, 3, 3, 2, seed=66)
    A = chidg_matrix(p, ctx)
    A.store(blocks)
    M = create_preconditioner("jacobi", precision).init(A)
    M.set_schedule(schedule)
    M.update(A)
    v = po.fill_vector(p.vec_len, 7)
    z = M.apply(A, v)
    n = p.n
    B = blocks.reshape(p.nnzb, n, n)
    ref = np.empty_like(v)
    for i in range(p.nrows):
        b = [k for k in range(p.row_ptr[i], p.row_ptr[i + 1]) if p.kind[k] == 0 and p.col[k] == i][0]
        ref[i * n:(i + 1) * n] = np.linalg.solve(B[b].T, v[i * n:(i + 1) * n])
    assert rel(z, ref) < tol
    part = np.arange(p.nrows, dtype=np.int32)
    L_ref, D_ref = po.ilu0_factor(p, blocks, part=part)
    assert rel(z, po.ilu0_apply(p, blocks, L_ref, D_ref, v, part=part)) < tol
    x = np.zeros(p.vec_len)
    ls = fgmres(nkrylov=30, tol=0.0, rtol=1e-8, inner_fgmres=False)
    ls.solve(A, x, v, M)
    assert ls.residual_norm / np.linalg.norm(v) < 1e-8
    M.tear_down()
    A.release()
    ctx.close()
