"""Writes tests/golden/golden.npz: small block systems with reference outputs
computed independently of the C oracle (numpy dense products, a numpy block
ILU(0) written from precon_ILU0.f90:111-265 / 286-411, numpy.linalg.solve),
plus the FGMRES(30)+ILU(0) iteration count of the oracle at rtol 1e-8.

Cases
  p0_4c   4x4x4 periodic box, P0 (5x5 blocks): chimera wrap blocks in the matvec
  p1_221  2x2x1 box, P1 (40x40 blocks), no periodic faces
  chain4  4x1x1 chain, P1: ILU(0) is the exact block LU (one FGMRES iteration)

Run from the repository root:  python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "tests"))

from chidg_amd import pattern as pt  # noqa: E402
from helpers import dense  # noqa: E402

CASES = {
    "p0_4c": dict(dims=(4, 4, 4), n=5, periodic=(True, True, True), seed=11),
    "p1_221": dict(dims=(2, 2, 1), n=40, periodic=(False, False, False), seed=12),
    "chain4": dict(dims=(4, 1, 1), n=40, periodic=(False, False, False), seed=13),
}


def blocks_for(p, seed):
    """Diagonal kind-0 blocks 2I + G/sqrt(n), others 0.1 G/sqrt(n), G ~ N(0,1)
    from numpy's PCG64 (column-major storage)."""
    rng = np.random.default_rng(seed)
    n = p.n
    out = np.empty((p.nnzb, n, n))
    for i in range(p.nrows):
        for b in range(p.row_ptr[i], p.row_ptr[i + 1]):
            g = rng.standard_normal((n, n))
            if p.kind[b] == 0 and p.col[b] == i:
                out[b] = 2.0 * np.eye(n) + g / np.sqrt(n)
            else:
                out[b] = 0.1 * g / np.sqrt(n)
    return np.ascontiguousarray(out.transpose(0, 2, 1)).reshape(-1)   # column-major per block


def block_ilu0(p, blocks):
    """Natural-order block ILU(0), precon_ILU0.f90:111-265 (row-major numpy)."""
    n = p.n
    B = blocks.reshape(p.nnzb, n, n).transpose(0, 2, 1)
    where = {(i, int(p.col[b])): b for i in range(p.nrows) for b in range(p.row_ptr[i], p.row_ptr[i + 1])
             if p.kind[b] == 0}
    L = np.zeros_like(B)
    Dinv = np.zeros((p.nrows, n, n))
    D = np.array([B[where[(i, i)]] for i in range(p.nrows)])
    Dinv[0] = np.linalg.inv(D[0])
    for i in range(1, p.nrows):
        for b in range(p.row_ptr[i], p.row_ptr[i + 1]):
            j = int(p.col[b])
            if p.kind[b] == 0 and j < i:
                L[b] = B[b] @ Dinv[j]
                D[i] = D[i] - L[b] @ B[where[(j, i)]]
        Dinv[i] = np.linalg.inv(D[i])
    return (np.ascontiguousarray(L.transpose(0, 2, 1)).reshape(-1),
            np.ascontiguousarray(Dinv.transpose(0, 2, 1)).reshape(-1))


def block_ilu0_apply(p, blocks, L, Dinv, v):
    """precon_ILU0.f90:286-411: forward with L, backward with A's upper blocks and Dinv."""
    n = p.n
    B = blocks.reshape(p.nnzb, n, n).transpose(0, 2, 1)
    Lb = L.reshape(p.nnzb, n, n).transpose(0, 2, 1)
    Db = Dinv.reshape(p.nrows, n, n).transpose(0, 2, 1)
    z = v.reshape(p.nrows, n).copy()
    for i in range(p.nrows):
        for b in range(p.row_ptr[i], p.row_ptr[i + 1]):
            j = int(p.col[b])
            if p.kind[b] == 0 and j < i:
                z[i] -= Lb[b] @ z[j]
    for i in range(p.nrows - 1, -1, -1):
        for b in range(p.row_ptr[i], p.row_ptr[i + 1]):
            k = int(p.col[b])
            if p.kind[b] == 0 and k > i:
                z[i] -= B[b] @ z[k]
        z[i] = Db[i] @ z[i]
    return z.reshape(-1)


def main():
    import pyoracle as po
    out = {}
    for name, c in CASES.items():
        p = pt.hex_box(*c["dims"], c["n"], c["periodic"])
        blocks = blocks_for(p, c["seed"])
        rng = np.random.default_rng(c["seed"] + 1000)
        x = rng.standard_normal(p.vec_len)
        b = rng.standard_normal(p.vec_len)
        Ad = dense(p, blocks)
        L, Dinv = block_ilu0(p, blocks)
        z = block_ilu0_apply(p, blocks, L, Dinv, b)
        xs = np.linalg.solve(Ad, b)
        _, niter, _, _ = po.fgmres(p, blocks, b, nkrylov=30, tol=0.0, rtol=1e-8)
        for k, v in dict(row_ptr=p.row_ptr, col=p.col, kind=p.kind, n=np.int32(p.n), blocks=blocks, x=x,
                         Ax=Ad @ x, b=b, L=L, Dinv=Dinv, ilu_b=z, solution=xs, niter=np.int32(niter)).items():
            out[f"{name}/{k}"] = v
        print(f"{name}: nrows {p.nrows} n {p.n} nnzb {p.nnzb} oracle FGMRES niter {niter}")
    np.savez_compressed(os.path.join(HERE, "golden.npz"), **out)


if __name__ == "__main__":
    main()
