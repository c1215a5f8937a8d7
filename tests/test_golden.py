"""Committed fixtures (tests/golden/golden.npz, written by make_golden.py with
numpy dense products, a numpy block ILU(0) and numpy.linalg.solve): the C
oracle against them on the CPU, the gfx950 path against them on the GPU.  The
fixtures' FGMRES iteration counts (from the oracle) are pinned by a textbook
flexible GMRES in numpy."""
import os

import numpy as np
import pytest

import pyoracle as po
from chidg_amd import pattern as pt

HERE = os.path.dirname(os.path.abspath(__file__))
G = np.load(os.path.join(HERE, "golden", "golden.npz"))
NAMES = sorted({k.split("/")[0] for k in G.files})


def case(name):
    f = {k.split("/")[1]: G[k] for k in G.files if k.startswith(name + "/")}
    p = pt.Pattern(len(f["row_ptr"]) - 1, int(f["n"]), f["row_ptr"].astype(np.int32), f["col"].astype(np.int32),
                   f["kind"].astype(np.int32))
    return p, f


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.parametrize("name", NAMES)
def test_oracle_matches_golden(name):
    p, f = case(name)
    assert rel(po.mv(p, f["blocks"], f["x"]), f["Ax"]) < 1e-13
    L, D = po.ilu0_factor(p, f["blocks"])
    lower = np.zeros(p.nnzb, bool)
    rows = np.repeat(np.arange(p.nrows), np.diff(p.row_ptr))
    lower[(p.kind == 0) & (p.col < rows)] = True
    nn = p.n * p.n
    if lower.any():
        assert rel(L.reshape(-1, nn)[lower], f["L"].reshape(-1, nn)[lower]) < 1e-12
    assert rel(D, f["Dinv"]) < 1e-12
    assert rel(po.ilu0_apply(p, f["blocks"], L, D, f["b"]), f["ilu_b"]) < 1e-12
    x, niter, _, _ = po.fgmres(p, f["blocks"], f["b"], nkrylov=30, tol=0.0, rtol=1e-8)
    assert niter == int(f["niter"])
    assert rel(x, f["solution"]) < 1e-7


def textbook_fgmres(A, b, precon, m=30, rtol=1e-8, nmax=1000):
    """Flexible GMRES(m) as in Saad, Iterative Methods for Sparse Linear
    Systems, 2nd ed., Alg. 9.6 (right preconditioning that may change per
    step), with modified Gram-Schmidt run twice, Givens rotations, restarts
    from the current x, x0 = 0, stopping when the residual estimate falls
    below rtol times the first ||r0||.  Dense numpy, independent of the C
    oracle.  Returns (x, iterations)."""
    x = np.zeros_like(b)
    r0_first = None
    it = 0
    while it < nmax:
        r = b - A @ x
        beta = np.linalg.norm(r)
        if r0_first is None:
            r0_first = beta
        if beta / r0_first < rtol:
            break
        V = np.zeros((m + 1, b.size))
        Z = np.zeros((m, b.size))
        H = np.zeros((m + 1, m))
        cs, sn = np.zeros(m), np.zeros(m)
        g = np.zeros(m + 1)
        g[0] = beta
        V[0] = r / beta
        k = 0
        for j in range(m):
            Z[j] = precon(V[j])
            w = A @ Z[j]
            for _ in range(2):
                for i in range(j + 1):
                    h = V[i] @ w
                    H[i, j] += h
                    w = w - h * V[i]
            H[j + 1, j] = np.linalg.norm(w)
            V[j + 1] = w / H[j + 1, j]
            for i in range(j):
                H[i, j], H[i + 1, j] = cs[i] * H[i, j] + sn[i] * H[i + 1, j], -sn[i] * H[i, j] + cs[i] * H[i + 1, j]
            d = np.hypot(H[j, j], H[j + 1, j])
            cs[j], sn[j] = H[j, j] / d, H[j + 1, j] / d
            H[j, j], H[j + 1, j] = d, 0.0
            g[j], g[j + 1] = cs[j] * g[j], -sn[j] * g[j]
            it += 1
            k = j + 1
            if abs(g[j + 1]) / r0_first < rtol or it >= nmax:
                break
        y = np.linalg.solve(np.triu(H[:k, :k]), g[:k])
        x = x + Z[:k].T @ y
        if abs(g[k]) / r0_first < rtol:
            break
    return x, it


@pytest.mark.parametrize("name", NAMES)
def test_golden_iterations_textbook_fgmres(name):
    """The fixtures' FGMRES(30) iteration counts (written from the C oracle)
    against a textbook flexible GMRES in numpy, preconditioned by the
    fixtures' numpy-built ILU(0) factors: the count within 1 (the two
    orthogonalize differently: selective classical Gram-Schmidt against
    modified Gram-Schmidt twice, so rounding can move the last step) and the
    solution within 1e-7 of the numpy.linalg.solve one."""
    import sys
    sys.path.insert(0, os.path.join(HERE, "golden"))
    from helpers import dense
    from make_golden import block_ilu0_apply
    p, f = case(name)
    A = dense(p, f["blocks"])
    x, it = textbook_fgmres(A, f["b"], lambda v: block_ilu0_apply(p, f["blocks"], f["L"], f["Dinv"], v))
    assert abs(it - int(f["niter"])) <= 1, (it, int(f["niter"]))
    assert rel(x, f["solution"]) < 1e-7
    if name == "chain4":        # ILU(0) of a chain is its exact block LU: one iteration
        assert it == 1


@pytest.mark.gpu
@pytest.mark.parametrize("name", NAMES)
def test_gpu_matches_golden(name):
    from chidg_amd.linsolve import Context, chidg_matrix, chidg_mv, fgmres, precon_ILU0
    p, f = case(name)
    ctx = Context(0)
    A = chidg_matrix(p, ctx)
    A.store(f["blocks"])
    assert rel(chidg_mv(A, f["x"]), f["Ax"]) < 1e-13
    for schedule in ("levels", "dataflow", "stream"):
        M = precon_ILU0(schedule=schedule).init(A)
        M.update(A)
        assert rel(M.apply(A, f["b"]), f["ilu_b"]) < 1e-12
        x = np.zeros(p.vec_len)
        ls = fgmres(nkrylov=30, tol=0.0, rtol=1e-8, inner_fgmres=False)
        ls.solve(A, x, f["b"], M)
        assert abs(ls.niter - int(f["niter"])) <= 1
        assert rel(x, f["solution"]) < 1e-7
        M.tear_down()
    A.release()
    ctx.close()
