"""Parity of the gfx950 path (through the C ABI) with the CPU oracle and the
reference's known answers.  Tolerances: matvec 1e-13
relative, FP64 solution 1e-8 relative, iteration counts within +-1."""
import numpy as np
import pytest

import pyoracle as po
from chidg_amd import pattern as pt
from chidg_amd import _lib
from chidg_amd.linsolve import (Context, chidg_matrix, chidg_mv, fgmres, precon_identity, precon_ILU0)
from helpers import (bc_mv_test_system, chimera_mv_test_system, fgmres_ref_test_system, fgmres_test_system,
                     ilu_test_system, set_block, synthetic_system)

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


# ------------------------------------------------------------- known answers
def test_known_ilu0_factor_and_apply(ctx):
    """test_ILU0_mod.pf:227-332."""
    p, blocks, b = ilu_test_system()
    A = upload(ctx, p, blocks)
    M = precon_ILU0().init(A)
    M.update(A, b)
    L, D = M.factors(A)
    blk = {(i, int(p.col[k])): k for i in range(3) for k in range(p.row_ptr[i], p.row_ptr[i + 1])}
    assert D[0] == 1.0 and D[1] == -1.0
    assert D[2] == pytest.approx(1 / 14, rel=1e-15)
    assert L[blk[(1, 0)]] == 2.0 and L[blk[(2, 1)]] == -2.0
    z = M.apply(A, b)
    np.testing.assert_allclose(z, [6.142857143, 2.57142857, 0.2142857143], atol=1e-5)
    np.testing.assert_allclose(z, [43 / 7, 18 / 7, 3 / 14], rtol=1e-14)


@pytest.mark.parametrize("inner", [True, False])
def test_known_fgmres_3x3(ctx, inner):
    """test_fgmres.pf: x = [3,4,-5] (tol 1e-9), identity preconditioner, tol 1e-13."""
    p, blocks, b, x_ref = fgmres_test_system()
    A = upload(ctx, p, blocks)
    ls = fgmres(tol=1e-13, inner_fgmres=inner)
    x = np.zeros(3)
    ls.solve(A, x, b, precon_identity())
    np.testing.assert_allclose(x, x_ref, atol=1e-9)
    _, n_orc, _, _ = po.fgmres(p, blocks, b, precon="identity", nkrylov=2000, tol=1e-13, rtol=1e-10,
                               inner_fgmres=inner)
    assert abs(ls.niter - n_orc) <= 1


def test_known_fgmres_2x2_skew(ctx):
    """test_fgmres_ref.pf: x = [-1, 1]."""
    p, blocks, b, x_ref = fgmres_ref_test_system()
    A = upload(ctx, p, blocks)
    x = np.zeros(2)
    fgmres(tol=1e-13).solve(A, x, b, precon_identity())
    np.testing.assert_allclose(x, x_ref, atol=1e-9)


def test_known_mv_chimera(ctx):
    """test_chidg_matrix_operator_mv.pf (tol 1e-13)."""
    p, blocks, x, expect = chimera_mv_test_system()
    y = chidg_mv(upload(ctx, p, blocks), x)
    for row, val in expect.items():
        np.testing.assert_allclose(y[row * p.n:(row + 1) * p.n], val, atol=1e-13)


def test_known_mv_bc(ctx):
    """test_comm_bc_D1E8_mod.pf, NRANK = 1 (tol 1e-10)."""
    p, blocks, x, expect = bc_mv_test_system()
    y = chidg_mv(upload(ctx, p, blocks), x)
    for row, val in expect.items():
        assert abs(y[row] - val) < 1e-10


# --------------------------------------------------------- seeded synthetic
# (mesh, polynomial degree): P1 -> 40x40, P2 -> 135x135, P3 -> 320x320 blocks
CASES = [((4, 4, 4), 1), ((3, 3, 3), 2), ((4, 4, 4), 2), ((2, 3, 2), 3)]


@pytest.mark.parametrize("dims,degree", CASES)
def test_mv_parity(ctx, dims, degree):
    p, blocks = synthetic_system(*dims, degree, seed=20261202)
    x = po.fill_vector(p.vec_len, 3)
    y = chidg_mv(upload(ctx, p, blocks), x)
    y_ref = po.mv(p, blocks, x)
    assert rel(y, y_ref) < 1e-13


@pytest.mark.parametrize("dims,degree", CASES[:3])
def test_ilu0_factor_apply_parity(ctx, dims, degree):
    p, blocks = synthetic_system(*dims, degree, seed=7)
    A = upload(ctx, p, blocks)
    M = precon_ILU0().init(A)
    M.update(A)
    L, D = M.factors(A)
    L_ref, D_ref = po.ilu0_factor(p, blocks)
    lower = np.zeros(p.nnzb, dtype=bool)
    rows = np.repeat(np.arange(p.nrows), np.diff(p.row_ptr))
    lower[(p.kind == 0) & (p.col < rows)] = True
    nn = p.n * p.n
    Lm = L.reshape(p.nnzb, nn)[lower]
    Lr = L_ref.reshape(p.nnzb, nn)[lower]
    assert rel(Lm, Lr) < 1e-12
    assert rel(D, D_ref) < 1e-12
    v = po.fill_vector(p.vec_len, 11)
    z = M.apply(A, v)
    z_ref = po.ilu0_apply(p, blocks, L_ref, D_ref, v)
    assert rel(z, z_ref) < 1e-12
    # repeated applies are bitwise identical (deterministic sweep)
    assert np.array_equal(M.apply(A, v), z)


@pytest.mark.parametrize("degree", [1, 2, 3])
@pytest.mark.parametrize("precision,tol", [("fp64", 1e-12), ("fp32", 1e-5)])
def test_factor_blocks_needing_row_pivots(ctx, degree, precision, tol):
    """Diagonal blocks 2P + 0.1 G/sqrt(n) (P a cyclic shift: the diagonal
    entries are small next to each column's largest) on every other block
    row.  Those blocks fail the diagonal-pivot threshold and take the
    partial-pivoting inverse, the others keep diagonal pivots; factor and
    apply match the oracle's (which pivots every block partially)."""
    p, blocks = synthetic_system(3, 3, 2, degree, seed=31)
    n = p.n
    B = blocks.reshape(p.nnzb, n, n)            # per block: [column][row]
    rng = np.random.default_rng(8)
    rows = np.repeat(np.arange(p.nrows), np.diff(p.row_ptr))
    shift = np.zeros((n, n))
    shift[np.arange(n), (np.arange(n) + 1) % n] = 2.0
    for b in np.flatnonzero((p.kind == pt.KIND_LOCAL) & (p.col == rows) & (rows % 2 == 0)):
        B[b] = (shift + 0.1 * rng.standard_normal((n, n)) / np.sqrt(n)).T
    A = upload(ctx, p, blocks)
    M = precon_ILU0(precision).init(A)
    M.update(A)
    L_ref, D_ref = po.ilu0_factor(p, blocks)
    _, D = M.factors(A)
    assert rel(D, D_ref) < tol
    v = po.fill_vector(p.vec_len, 12)
    assert rel(M.apply(A, v), po.ilu0_apply(p, blocks, L_ref, D_ref, v)) < tol


@pytest.mark.parametrize("orth", ["CGS", "MGS"])
def test_fgmres_parity_config1(ctx, orth):
    """Config 1 shape: 4^3 periodic P1, FGMRES(30)+ILU0, rtol 1e-8, inner off."""
    p, blocks = synthetic_system(4, 4, 4, 1, seed=20261202)
    b = po.fill_vector(p.vec_len, 1)
    A = upload(ctx, p, blocks)
    ls = fgmres(nkrylov=30, tol=1e-30, rtol=1e-8, inner_fgmres=False, orthogonalization=orth)
    x = np.zeros(p.vec_len)
    ls.solve(A, x, b, precon_ILU0())
    x_ref, n_ref, r_ref, _ = po.fgmres(p, blocks, b, nkrylov=30, tol=1e-30, rtol=1e-8, orthogonalization=orth)
    assert abs(ls.niter - n_ref) <= 1
    assert rel(x, x_ref) < 1e-8
    assert ls.residual_norm / np.linalg.norm(b) < 1e-8


def test_fgmres_inner_parity(ctx):
    p, blocks = synthetic_system(3, 3, 3, 1, seed=5)
    b = po.fill_vector(p.vec_len, 2)
    A = upload(ctx, p, blocks)
    ls = fgmres(nkrylov=10, tol=1e-30, rtol=1e-9, inner_fgmres=True, inner_nkrylov=5)
    x = np.zeros(p.vec_len)
    ls.solve(A, x, b, precon_ILU0())
    x_ref, n_ref, _, _ = po.fgmres(p, blocks, b, nkrylov=10, tol=1e-30, rtol=1e-9, inner_fgmres=True,
                                   inner_nkrylov=5)
    assert abs(ls.niter - n_ref) <= 1
    assert rel(x, x_ref) < 1e-8


def test_fgmres_restart_nmax(ctx):
    p, blocks = synthetic_system(3, 3, 3, 1, seed=5)
    b = po.fill_vector(p.vec_len, 6)
    A = upload(ctx, p, blocks)
    ls = fgmres(nkrylov=5, tol=1e-30, rtol=1e-10, inner_fgmres=False)
    x = np.zeros(p.vec_len)
    ls.solve(A, x, b, precon_identity())
    _, n_ref, _, _ = po.fgmres(p, blocks, b, precon="identity", nkrylov=5, tol=1e-30, rtol=1e-10)
    assert abs(ls.niter - n_ref) <= 1 and ls.stats.nrestart >= 1
    ls.nmax = 7
    ls.solve(A, x, b, precon_identity())
    assert ls.niter == 7


@pytest.mark.parametrize("precision,tol", [("fp32", 1e-5), ("bf16", 5e-2)])
def test_mixed_precision_ilu(ctx, precision, tol):
    """FP32 / BF16-stored ILU(0): apply error against the FP64 oracle is stated
    by the tolerance; FGMRES still reaches rtol 1e-8 in FP64."""
    p, blocks = synthetic_system(4, 4, 4, 2, seed=3)
    A = upload(ctx, p, blocks)
    M = precon_ILU0(precision).init(A)
    M.update(A)
    v = po.fill_vector(p.vec_len, 4)
    L_ref, D_ref = po.ilu0_factor(p, blocks)
    z_ref = po.ilu0_apply(p, blocks, L_ref, D_ref, v)
    assert rel(M.apply(A, v), z_ref) < tol
    b = po.fill_vector(p.vec_len, 1)
    x = np.zeros(p.vec_len)
    ls = fgmres(nkrylov=30, tol=1e-30, rtol=1e-8, inner_fgmres=False)
    ls.solve(A, x, b, M)
    x_ref, n_ref, _, _ = po.fgmres(p, blocks, b, nkrylov=30, tol=1e-30, rtol=1e-8)
    assert ls.residual_norm / np.linalg.norm(b) < 1e-8
    assert rel(x, x_ref) < 1e-6
    assert ls.niter <= n_ref + 3


def test_device_generator_matches_host(ctx):
    p = pt.hex_box(3, 2, 2, 40)
    A = chidg_matrix(p, ctx)
    A.fill_synthetic(99, 1 / np.sqrt(40), 0.1 / np.sqrt(40))
    ref = po.fill_synthetic(p, 99, 1 / np.sqrt(40), 0.1 / np.sqrt(40))
    assert np.array_equal(A.to_host(), ref)
    v = A.synthetic_vector(5, 1.0)
    assert np.array_equal(v, po.fill_vector(p.vec_len, 5, 1.0))


# ----------------------------------------------------------------- errors
def test_singular_block_reports_row(ctx):
    p, blocks = synthetic_system(2, 2, 2, 1, seed=1)
    n = p.n
    # D_5 = A_55 - sum L_5j A_j5 vanishes when A_55 and row 5's lower blocks are zero
    for parent in (5, 4, 1):
        set_block(p, blocks, 5, parent, np.zeros((n, n)))
    A = upload(ctx, p, blocks)
    M = precon_ILU0().init(A)
    with pytest.raises(_lib.CdgError) as e:
        M.update(A)
    assert e.value.code == 3 and e.value.row == 5


def test_missing_transpose_rejected(ctx):
    rows = [[(0, 0), (1, 0)], [(1, 0)]]   # (0,1) upper without (1,0) lower
    A = chidg_matrix(pt.from_rows(2, rows), ctx)
    with pytest.raises(_lib.CdgError) as e:
        precon_ILU0().init(A)
    assert e.value.code == 4


def test_zero_rhs(ctx):
    p, blocks = synthetic_system(2, 2, 2, 1, seed=5)
    A = upload(ctx, p, blocks)
    x = np.ones(p.vec_len)
    ls = fgmres(nkrylov=30, inner_fgmres=False)
    ls.solve(A, x, np.zeros(p.vec_len), precon_ILU0())
    assert ls.niter == 0 and np.all(x == 0)


def test_device_resident_vectors(ctx):
    import torch
    p, blocks = synthetic_system(3, 3, 3, 1, seed=8)
    A = upload(ctx, p, blocks)
    xh = po.fill_vector(p.vec_len, 2)
    xd = torch.from_numpy(xh).to("cuda:0")
    yd = chidg_mv(A, xd)
    assert yd.is_cuda
    assert np.array_equal(yd.cpu().numpy(), chidg_mv(A, xh))


# --------------------------------------------- full size (16^3 P2) properties
@pytest.fixture(scope="module")
def big(ctx):
    p = pt.hex_box(16, 16, 16, 135)
    A = chidg_matrix(p, ctx)
    A.fill_synthetic(20261202, 1 / np.sqrt(135), 0.1 / np.sqrt(135))
    yield p, A
    A.release()


def test_full_size_mv_linearity(big):
    import torch
    p, A = big
    x = A.synthetic_vector(1, like=torch.empty(0, device="cuda:0"))
    y = A.synthetic_vector(2, like=torch.empty(0, device="cuda:0"))
    lhs = chidg_mv(A, x + 2.0 * y)
    rhs = chidg_mv(A, x) + 2.0 * chidg_mv(A, y)
    assert float(torch.linalg.norm(lhs - rhs) / torch.linalg.norm(rhs)) < 1e-14


def test_full_size_fgmres_fp64(big):
    import torch
    p, A = big
    b = A.synthetic_vector(7, like=torch.empty(0, device="cuda:0"))
    x = torch.zeros_like(b)
    M = precon_ILU0().init(A)
    ls = fgmres(nkrylov=30, tol=1e-30, rtol=1e-8, inner_fgmres=False)
    ls.solve(A, x, b, M)
    true_res = float(torch.linalg.norm(b - chidg_mv(A, x)) / torch.linalg.norm(b))
    assert true_res < 1e-8
    assert ls.residual_norm / float(torch.linalg.norm(b)) == pytest.approx(true_res, rel=1e-6)
    z1 = M.apply(A, b)
    assert torch.equal(z1, M.apply(A, b))


# ------------------------------------------- P3 (320 x 320 blocks, config 3)
@pytest.mark.parametrize("precision,tol", [("fp64", 1e-12), ("fp32", 1e-5), ("bf16", 5e-2)])
def test_p3_factor_apply_fgmres(ctx, precision, tol):
    """n = 320 (P3, 5 fields): factor through the global-memory block inverse,
    apply against the FP64 oracle, FGMRES(30) to rtol 1e-8 in FP64."""
    p, blocks = synthetic_system(3, 2, 2, 3, seed=31, periodic=(True, False, False))
    A = upload(ctx, p, blocks)
    M = precon_ILU0(precision).init(A)
    M.update(A)
    L_ref, D_ref = po.ilu0_factor(p, blocks)
    if precision == "fp64":
        _, D = M.factors(A)
        assert rel(D, D_ref) < 1e-12
    v = po.fill_vector(p.vec_len, 9)
    assert rel(M.apply(A, v), po.ilu0_apply(p, blocks, L_ref, D_ref, v)) < tol
    b = po.fill_vector(p.vec_len, 10)
    x = np.zeros(p.vec_len)
    ls = fgmres(nkrylov=30, tol=1e-30, rtol=1e-8, inner_fgmres=False)
    ls.solve(A, x, b, M)
    x_ref, n_ref, _, _ = po.fgmres(p, blocks, b, nkrylov=30, tol=1e-30, rtol=1e-8)
    assert ls.residual_norm / np.linalg.norm(b) < 1e-8
    assert rel(x, x_ref) < 1e-7
    if precision == "fp64":
        assert abs(ls.niter - n_ref) <= 1


def test_tiny_and_empty_systems(ctx):
    """One block row (n = 1 and n = 40) and an empty matrix."""
    for n in (1, 40):
        p = pt.hex_box(1, 1, 1, n, periodic=(False, False, False))
        blocks = po.fill_synthetic(p, 3, 1 / np.sqrt(n), 0.1 / np.sqrt(n))
        A = upload(ctx, p, blocks)
        b = po.fill_vector(p.vec_len, 2)
        x = np.zeros(p.vec_len)
        ls = fgmres(nkrylov=30, tol=0.0, rtol=1e-12, inner_fgmres=False)
        ls.solve(A, x, b, precon_ILU0())
        assert ls.niter == 1        # ILU(0) of one block is its exact inverse
        assert rel(x, np.linalg.solve(blocks.reshape(n, n).T, b)) < 1e-12
    p = pt.from_rows(5, [])
    A = chidg_matrix(p, ctx)
    A.store(np.zeros(0))
    assert chidg_mv(A, np.zeros(0)).shape == (0,)


def test_controller_refactors_only_when_needed(ctx):
    """fgmres%solve with a solver_controller (type_fgmres.f90:118-122): the
    second solve on an unchanged matrix reuses the factor; a store() that
    re-stamps the matrix, and the force flag, each trigger one refactor."""
    from chidg_amd.linsolve import solver_controller
    p, blocks = synthetic_system(3, 3, 3, 1, seed=44)
    A = upload(ctx, p, blocks)
    M = precon_ILU0().init(A)
    ctl = solver_controller()
    ls = fgmres(nkrylov=30, tol=1e-30, rtol=1e-8, inner_fgmres=False)
    b = po.fill_vector(p.vec_len, 3)
    seen = []
    for action in ("first", "same", "store", "same", "force", "same"):
        if action == "store":
            A.store(blocks)
        elif action == "force":
            ctl.force_update_preconditioner = True
        x = np.zeros(p.vec_len)
        ls.solve(A, x, b, M, solver_controller=ctl)
        assert ls.residual_norm / np.linalg.norm(b) < 1e-8
        seen.append(ls.stats.factored)
    assert seen == [1, 0, 1, 0, 1, 0]


def test_torch_producer_ordered_before_library(ctx):
    """Device inputs written by torch kernels still in flight on torch's
    stream: the library's stream waits for them (cdg_ctx_wait_stream), and
    torch consumers of an enqueue-only apply wait for it."""
    import torch
    p, blocks = synthetic_system(4, 4, 4, 2, seed=45)
    A = upload(ctx, p, blocks)
    M = precon_ILU0().init(A)
    M.update(A)
    xh = po.fill_vector(p.vec_len, 5)
    y_ref = chidg_mv(A, xh)
    z_ref = M.apply(A, xh)
    big = torch.randn(4096, 4096, device="cuda:0", dtype=torch.float64)
    for _ in range(3):
        x0 = torch.zeros(p.vec_len, device="cuda:0", dtype=torch.float64)
        _ = big @ big                      # keeps torch's stream busy
        x = x0 + torch.from_numpy(xh).to("cuda:0")
        assert np.array_equal(chidg_mv(A, x).cpu().numpy(), y_ref)
        _ = big @ big
        v = x * 1.0
        z = torch.empty_like(v)
        M.apply_async(A, v, z)
        assert np.array_equal((z * 1.0).cpu().numpy(), z_ref)
