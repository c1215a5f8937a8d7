"""Shared test helpers: known-answer systems of the reference tests and dense
assembly (the analogue of ``chidg_matrix_t%to_real``,
``src/containers/type_chidg_matrix.f90:1364-1432``)."""
from __future__ import annotations

import numpy as np

from chidg_amd import pattern as pt


def dense(p: pt.Pattern, blocks: np.ndarray, kinds=(0, 1)) -> np.ndarray:
    """Dense matrix of the given block kinds (blocks column-major)."""
    n = p.n
    A = np.zeros((p.nrows * n, p.nrows * n))
    B = blocks.reshape(p.nnzb, n, n)
    for i in range(p.nrows):
        for b in range(p.row_ptr[i], p.row_ptr[i + 1]):
            if p.kind[b] in kinds:
                j = p.col[b]
                A[i * n:(i + 1) * n, j * n:(j + 1) * n] += B[b].T   # column-major -> row-major
    return A


def set_block(p: pt.Pattern, blocks: np.ndarray, row: int, parent: int, value, kind=0):
    """Write a block addressed like ``lblks(row)%loc(parent)`` (first match)."""
    n = p.n
    for b in range(p.row_ptr[row], p.row_ptr[row + 1]):
        if p.col[b] == parent and p.kind[b] == kind:
            v = np.asarray(value, dtype=np.float64)
            blk = np.broadcast_to(v, (n, n)) if v.ndim == 0 else v
            blocks[b * n * n:(b + 1) * n * n] = np.asarray(blk).T.reshape(-1)  # store column-major
            return b
    raise KeyError((row, parent, kind))


def ilu_test_system():
    """test/unit_testing/preconditioners/test_ILU0_mod.pf:53-124: 3x1x1 mesh,
    1x1 blocks, A = [[1,-2,0],[2,-5,12],[0,2,-10]], b = [1,2,3]."""
    p = pt.hex_box(3, 1, 1, 1, periodic=(False, False, False))
    blocks = np.zeros(p.nnzb)
    A = [[1, -2, 0], [2, -5, 12], [0, 2, -10]]
    for i in range(3):
        for j in range(3):
            if A[i][j] != 0:
                set_block(p, blocks, i, j, float(A[i][j]))
    return p, blocks, np.array([1.0, 2.0, 3.0])


def fgmres_test_system():
    """test/unit_testing/linear_solvers/test_fgmres.pf: tridiagonal 3x3,
    A = [[4,3,0],[3,4,-1],[0,-1,4]], b = [24,30,-24] -> x = [3,4,-5]."""
    p = pt.hex_box(3, 1, 1, 1, periodic=(False, False, False))
    blocks = np.zeros(p.nnzb)
    A = [[4, 3, 0], [3, 4, -1], [0, -1, 4]]
    for i in range(3):
        for j in range(3):
            if A[i][j] != 0:
                set_block(p, blocks, i, j, float(A[i][j]))
    return p, blocks, np.array([24.0, 30.0, -24.0]), np.array([3.0, 4.0, -5.0])


def fgmres_ref_test_system():
    """test/unit_testing/linear_solvers/test_fgmres_ref.pf: A = [[0,1],[-1,0]],
    b = [1,1] -> x = [-1,1]."""
    p = pt.hex_box(2, 1, 1, 1, periodic=(False, False, False))
    blocks = np.zeros(p.nnzb)
    set_block(p, blocks, 0, 0, 0.0)
    set_block(p, blocks, 0, 1, 1.0)
    set_block(p, blocks, 1, 0, -1.0)
    set_block(p, blocks, 1, 1, 0.0)
    return p, blocks, np.array([1.0, 1.0]), np.array([-1.0, 1.0])


def chimera_mv_test_system():
    """test/unit_testing/containers/test_chidg_matrix_operator_mv.pf: two 2x2x2
    domains (rows 0-7 and 8-15), nterms_s = 8 with one field (8x8 blocks).
    lblks(dom1, elem1) diag = 1; chi_blks(dom1, elem2) = {2 (x=0.2), 3 (x=0.3)};
    chi_blks(dom2, elem1) = {4 (x=0.1)}.  Expected b(dom1,elem2) = (2*0.2+3*0.3)*8,
    b(dom2,elem1) = 4*0.1*8."""
    n = 8
    d1 = pt.hex_box(2, 2, 2, n, periodic=(False, False, False))
    rows = []
    for dom in range(2):
        for e in range(8):
            r = [(int(d1.col[b]) + 8 * dom, 0) for b in range(d1.row_ptr[e], d1.row_ptr[e + 1])]
            rows.append(r)
    # donors: dom1 elem2 <- dom2 elems 1 and 3 ; dom2 elem1 <- dom1 elem 2
    rows[1] += [(8, 1), (10, 1)]
    rows[8] += [(1, 1)]
    p = pt.from_rows(n, rows)
    blocks = np.zeros(p.nnzb * n * n)
    set_block(p, blocks, 0, 0, 1.0)
    set_block(p, blocks, 1, 8, 2.0, kind=1)
    set_block(p, blocks, 1, 10, 3.0, kind=1)
    set_block(p, blocks, 8, 1, 4.0, kind=1)
    x = np.zeros(p.vec_len)
    x[8 * n:9 * n] = 0.2
    x[10 * n:11 * n] = 0.3
    x[1 * n:2 * n] = 0.1
    expect = {1: (2 * 0.2 + 3 * 0.3) * 8, 8: (4 * 0.1) * 8}
    return p, blocks, x, expect


def bc_mv_test_system():
    """test/unit_testing/parallel/test_comm_bc_D1E8_mod.pf:850-1185, NRANK = 1:
    2x2x2 mesh, 1x1 blocks, globally coupled BC blocks (2,8)=2.8, (4,6)=4.6,
    (6,4)=6.4, (8,2)=8.2 (1-based), x(e) = e -> b(2,4,6,8) = 22.4, 27.6, 25.6, 16.4."""
    base = pt.hex_box(2, 2, 2, 1, periodic=(False, False, False))
    rows = [[(int(base.col[b]), 0) for b in range(base.row_ptr[e], base.row_ptr[e + 1])]
            for e in range(8)]
    coupling = {1: (7, 2.8), 3: (5, 4.6), 5: (3, 6.4), 7: (1, 8.2)}
    for r, (c, _) in coupling.items():
        rows[r].append((c, 1))
    p = pt.from_rows(1, rows)
    blocks = np.zeros(p.nnzb)
    for r, (c, v) in coupling.items():
        set_block(p, blocks, r, c, v, kind=1)
    x = np.zeros(8)
    for e in (1, 3, 5, 7):
        x[e] = float(e + 1)
    expect = {1: 22.4, 3: 27.6, 5: 25.6, 7: 16.4}
    return p, blocks, x, expect


def synthetic_system(nx, ny, nz, degree, seed, periodic=(True, True, True), nfields=5,
                     s_off_scale=0.1):
    """Seeded synthetic Jacobian at polynomial degree ``degree`` (P1: n=40,
    P2: 135, P3: 320): diag blocks 2I + G/sqrt(n), off-diagonal blocks
    0.1*G/sqrt(n), G ~ approx N(0,1) from the counter hash."""
    import pyoracle as po
    n = pt.block_size(degree, nfields)
    p = pt.hex_box(nx, ny, nz, n, periodic)
    blocks = po.fill_synthetic(p, seed, 1.0 / np.sqrt(n), s_off_scale / np.sqrt(n))
    return p, blocks


def _mix64(z):
    z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
    z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return z ^ (z >> np.uint64(31))


def synth_block(seed: int, b: int, n: int, is_diag: bool, s_diag: float, s_off: float) -> np.ndarray:
    """Block ``b`` of the seeded synthetic generator (oracle/synth.c), n*n
    column-major, computed with numpy's wrapping uint64 arithmetic: every
    step is exact or one IEEE rounding, so the values are bitwise those of
    the generator.  Lets a test check single blocks of a matrix too large to
    generate on the host."""
    G = np.uint64(0x9E3779B97F4A7C15)
    with np.errstate(over="ignore"):
        idx = np.uint64(b) * np.uint64(2097152) + np.arange(n * n, dtype=np.uint64)
        s = np.uint64(seed)
        h1 = _mix64(s + G * (np.uint64(2) * idx + np.uint64(1)))
        h2 = _mix64(s + G * (np.uint64(2) * idx + np.uint64(2)))
    lo = np.uint64(0xFFFFFFFF)
    S = (h1 & lo) + (h1 >> np.uint64(32)) + (h2 & lo) + (h2 >> np.uint64(32))
    g = (S.astype(np.float64) * 2.3283064365386963e-10 - 2.0) * 1.7320508075688772
    out = g * (s_diag if is_diag else s_off)
    if is_diag:
        out[np.arange(n) * (n + 1)] += 2.0
    return out
