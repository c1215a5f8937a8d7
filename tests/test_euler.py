"""The DG Euler first-Newton-step generator (csrc/euler_jac.cpp), pinned by
what the mathematics fixes: the Jacobian equals central finite differences of
the residual (every column, small boxes), the uniform freestream is an exact
steady state, the periodic discretization conserves every field, the
pseudo-transient term is mass/dtau on the diagonal sub-blocks only, and the
blocks land in the hex-box pattern (interior neighbours local, wrap
neighbours coupling-only).  CPU only."""
import numpy as np
import pytest

from chidg_amd import euler
from chidg_amd import pattern as pt
from helpers import dense

GAM = 1.4


def test_freestream_is_steady():
    for N, order in ((2, 0), (3, 1), (2, 2)):
        q = euler.state(N, order, mach=0.3, perturb=0.0)
        R = euler.residual(N, order, q)
        assert np.max(np.abs(R)) < 1e-13


@pytest.mark.parametrize("N,order", [(3, 1), (2, 2), (2, 3)])
def test_conservation(N, order):
    """Periodic box: the mode-0 (constant test function) residuals of each
    field sum to zero over the elements (face fluxes cancel pairwise and the
    volume term of a constant test function vanishes)."""
    n = pt.block_size(order)
    nt = n // 5
    q = euler.state(N, order, mach=0.4, seed=3, perturb=0.02)
    R = euler.residual(N, order, q).reshape(N ** 3, 5, nt)
    total = R[:, :, 0].sum(axis=0)
    assert np.max(np.abs(total)) < 1e-13 * max(1.0, np.max(np.abs(R)))


@pytest.mark.parametrize("N,order", [(3, 1), (2, 2), (3, 0)])
def test_jacobian_matches_finite_differences(N, order):
    """Every column of dR/dq (cfl = 0) against central differences of R."""
    q = euler.state(N, order, mach=0.3, seed=7, perturb=0.01)
    p, blocks, R = euler.jacobian(N, order, q, cfl=0.0)
    assert np.allclose(R, euler.residual(N, order, q), rtol=0, atol=1e-14)
    J = dense(p, blocks)
    h = 1e-6
    ncol = q.size
    Jfd = np.empty((ncol, ncol))
    for k in range(ncol):
        qp = q.copy(); qp[k] += h
        qm = q.copy(); qm[k] -= h
        Jfd[:, k] = (euler.residual(N, order, qp, 1) - euler.residual(N, order, qm, 1)) / (2 * h)
    err = np.max(np.abs(J - Jfd)) / np.max(np.abs(J))
    assert err < 1e-7, err


def test_pseudo_transient_term():
    """J(cfl) - J(0) is mass/dtau on the diagonal of every diagonal block,
    the same for the 5 fields of a mode, dtau = cfl h / (mean|u| + mean c)."""
    N, order, cfl = 3, 1, 2.0
    n = pt.block_size(order)
    nt = n // 5
    q = euler.state(N, order, mach=0.3, seed=2, perturb=0.01)
    p, B1, _ = euler.jacobian(N, order, q, cfl=cfl)
    _, B0, _ = euler.jacobian(N, order, q, cfl=0.0)
    D = (B1 - B0).reshape(p.nnzb, n, n)
    rows = np.repeat(np.arange(p.nrows), np.diff(p.row_ptr))
    diag = (p.kind == pt.KIND_LOCAL) & (p.col == rows)
    assert np.all(D[~diag] == 0)
    h = 1.0 / N
    # mass of mode (i,j,k) on a cube of side h: (h/2)^3 prod 2/(2i+1)
    P1 = order + 1
    m = np.array([(h / 2) ** 3 * np.prod([2.0 / (2 * ((t // P1 ** d) % P1) + 1) for d in range(3)])
                  for t in range(nt)])
    for b in np.flatnonzero(diag):
        d = np.diag(D[b])
        assert np.allclose(D[b], np.diag(d))
        ratio = d.reshape(5, nt) / m            # 1/dtau, one value per element
        assert np.allclose(ratio, ratio[0, 0], rtol=1e-12)
        e = rows[b]
        # element-mean wave speed at the quadrature points bounds dtau
        lam = cfl * h * ratio[0, 0]
        assert 1.0 < lam < 1.0 + 0.3 + 0.2     # c ~ 1, |u| ~ 0.3 at this state
        del e


def test_block_placement():
    """Interior-face neighbours sit in local blocks, periodic wrap neighbours
    in coupling blocks; no block of the pattern stays empty (N = 3)."""
    N, order = 3, 1
    q = euler.state(N, order, seed=4)
    p, blocks, _ = euler.jacobian(N, order, q)
    B = blocks.reshape(p.nnzb, p.n, p.n)
    assert p.nnzb == 7 * N ** 3
    assert np.all(np.max(np.abs(B), axis=(1, 2)) > 0)
    # N = 2: interior and wrap neighbours coincide; both blocks are filled
    p2, b2, _ = euler.jacobian(2, 0, euler.state(2, 0, seed=4))
    B2 = b2.reshape(p2.nnzb, p2.n, p2.n)
    assert np.all(np.max(np.abs(B2), axis=(1, 2)) > 0)
    assert np.any(p2.kind == pt.KIND_COUPLING)


def test_first_newton_system_solves():
    """Config 1 (4^3 P1, Mach 0.3): the oracle's FGMRES(30)+ILU(0) reaches
    rtol 1e-8 on the generated system with b = -R."""
    import pyoracle as po
    p, blocks, R, _ = euler.system(4, 1, mach=0.3, seed=1, perturb=0.01, cfl=1.0)
    x, niter, res, _ = po.fgmres(p, blocks, -R, nkrylov=30, tol=0.0, rtol=1e-8)
    assert res / np.linalg.norm(R) < 1e-8
    assert 1 <= niter < 200
    A = dense(p, blocks)
    assert np.linalg.norm(A @ x + R) / np.linalg.norm(R) < 1e-7


def test_invalid_inputs_are_errors():
    """Sizes outside the generator's range, non-finite states, and states
    without a real sound speed (negative pressure: the pseudo-time step is
    undefined) come back as status codes, not NaN blocks."""
    from chidg_amd import _lib
    with pytest.raises(_lib.CdgError):
        euler.state(1, 1)
    with pytest.raises(_lib.CdgError):
        euler.state(3, 4)
    q = euler.state(2, 1, mach=0.3, perturb=0.0)
    n = pt.block_size(1)
    q = q.reshape(8, 5, n // 5)
    q[3, 4, 0] = -1.0            # energy below the kinetic energy: p < 0
    with pytest.raises(_lib.CdgError):
        euler.jacobian(2, 1, q.ravel(), cfl=1.0)
    q[3, 4, 0] = np.nan
    with pytest.raises(_lib.CdgError):
        euler.residual(2, 1, q.ravel())
