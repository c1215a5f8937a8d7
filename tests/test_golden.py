"""Committed fixtures (tests/golden/golden.npz, written by make_golden.py with
numpy dense products, a numpy block ILU(0) and numpy.linalg.solve): the C
oracle against them on the CPU, the gfx950 path against them on the GPU."""
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
