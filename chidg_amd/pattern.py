"""Block-sparse Jacobian patterns in ChiDG's own row/block order.

A pattern is the flattened form of ``domain_matrix_t`` (reference
``src/containers/type_domain_matrix.f90``): one block row per element (per
time level), and per row the ``lblks`` entries followed by the coupling-only
entries (``chi_blks``/``bc_blks``/``hb_blks``).

* ``kind == 0``: an ``lblks`` entry, i.e. an interior-face neighbour or the
  diagonal.  Lower/upper status is ``parent < row`` / ``parent > row``
  (``type_domain_matrix.f90:295-302``); ILU(0) reads only these.
* ``kind == 1``: a coupling-only entry.  Periodic faces are Chimera faces in
  ChiDG (``src/boundary_conditions/periodic/bc_periodic.f90:116-125``), so a
  periodic wrap neighbour lands in ``chi_blks`` (``mod_integrate.f90:714-721``):
  the matvec uses it, ILU(0) never does.

Element numbering of a structured hex block is lexicographic with ``i``
fastest (``src/io/mod_plot3d_utilities.f90:140-173``).  ``lblks`` entries of a
row follow the face order ``[XI_MIN, XI_MAX, ETA_MIN, ETA_MAX, ZETA_MIN,
ZETA_MAX, DIAG]`` with missing neighbours skipped
(``type_domain_matrix.f90:134,253-305``); Chimera donors are appended per face
in the same face order and a donor already present is not added twice
(``type_domain_matrix.f90:318-374``).
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

KIND_LOCAL = 0
KIND_COUPLING = 1

# face order of the reference (src/constants/mod_constants.f90:94-100)
XI_MIN, XI_MAX, ETA_MIN, ETA_MAX, ZETA_MIN, ZETA_MAX, DIAG = range(1, 8)


@dataclass
class Pattern:
    """Row structure of a block-sparse matrix (all arrays int32)."""

    nrows: int
    n: int                      # block size = nterms * nfields
    row_ptr: np.ndarray         # nrows + 1
    col: np.ndarray             # parent row of each block
    kind: np.ndarray            # KIND_LOCAL / KIND_COUPLING
    face: np.ndarray = field(default=None)  # face id of each block (DIAG for diagonal); info only

    @property
    def nnzb(self) -> int:
        return int(self.row_ptr[-1])

    @property
    def vec_len(self) -> int:
        return self.nrows * self.n

    def block_bytes(self, itemsize: int = 8) -> int:
        return self.nnzb * self.n * self.n * itemsize

    def counts(self) -> dict:
        rows = np.repeat(np.arange(self.nrows), np.diff(self.row_ptr))
        loc = self.kind == KIND_LOCAL
        return {
            "diag": int(np.sum(loc & (self.col == rows))),
            "lower": int(np.sum(loc & (self.col < rows))),
            "upper": int(np.sum(loc & (self.col > rows))),
            "coupling": int(np.sum(~loc)),
            "total": self.nnzb,
        }

    def validate(self) -> None:
        assert self.row_ptr.shape == (self.nrows + 1,)
        assert self.row_ptr[0] == 0 and np.all(np.diff(self.row_ptr) >= 0)
        assert self.col.shape == (self.nnzb,) and self.kind.shape == (self.nnzb,)
        assert np.all((self.col >= 0) & (self.col < self.nrows))


def from_rows(n: int, rows: list[list[tuple[int, int]]]) -> Pattern:
    """Pattern from explicit per-row lists of ``(parent, kind)``."""
    row_ptr = np.zeros(len(rows) + 1, dtype=np.int32)
    col, kind = [], []
    for i, r in enumerate(rows):
        row_ptr[i + 1] = row_ptr[i] + len(r)
        for parent, k in r:
            col.append(parent)
            kind.append(k)
    p = Pattern(len(rows), n, row_ptr, np.asarray(col, dtype=np.int32),
                np.asarray(kind, dtype=np.int32), None)
    p.validate()
    return p


def hex_box(nx: int, ny: int, nz: int, n: int, periodic=(True, True, True)) -> Pattern:
    """Pattern of a uniform structured hex box.

    ``periodic[d]`` makes the two boundary faces normal to direction ``d``
    periodic (coupling-only wrap blocks); otherwise they are plain boundary
    faces with no coupling block.
    """
    dims = (nx, ny, nz)
    stride = (1, nx, nx * ny)
    nrows = nx * ny * nz
    row_ptr = np.zeros(nrows + 1, dtype=np.int32)
    col, kind, face = [], [], []
    for e in range(nrows):
        idx = (e % nx, (e // nx) % ny, e // (nx * ny))
        start = len(col)
        chi = []
        for d in range(3):
            for side, fid in ((-1, XI_MIN + 2 * d), (+1, XI_MAX + 2 * d)):
                q = idx[d] + side
                if 0 <= q < dims[d]:
                    col.append(e + side * stride[d])
                    kind.append(KIND_LOCAL)
                    face.append(fid)
                elif periodic[d]:
                    wrap = (q + dims[d]) % dims[d]
                    donor = e + (wrap - idx[d]) * stride[d]
                    if donor not in [c for c, _ in chi]:
                        chi.append((donor, fid))
        col.append(e)
        kind.append(KIND_LOCAL)
        face.append(DIAG)
        for donor, fid in chi:
            col.append(donor)
            kind.append(KIND_COUPLING)
            face.append(fid)
        row_ptr[e + 1] = row_ptr[e] + (len(col) - start)
    p = Pattern(nrows, n, row_ptr, np.asarray(col, dtype=np.int32),
                np.asarray(kind, dtype=np.int32), np.asarray(face, dtype=np.int32))
    p.validate()
    return p


def block_size_order(order: int, nfields: int = 5) -> int:
    """Block size for ChiDG's 'Solution Order' k: ``nterms = k**3``
    (type_chidg.f90:570-572)."""
    return order ** 3 * nfields


def block_size(p: int, nfields: int = 5) -> int:
    """Block size for polynomial degree ``p`` (P0, P1, P2, P3): the tensor
    Legendre basis has ``(p+1)**3`` terms, i.e. Solution Order ``p+1``;
    P1/P2/P3 with 5 fields give 40/135/320."""
    return block_size_order(p + 1, nfields)


def levels(p: Pattern) -> tuple[np.ndarray, np.ndarray]:
    """Forward / backward wavefront level of each row (kind-0 lower/upper DAG)."""
    fl = np.zeros(p.nrows, dtype=np.int32)
    bl = np.zeros(p.nrows, dtype=np.int32)
    for i in range(p.nrows):
        s, e = p.row_ptr[i], p.row_ptr[i + 1]
        c, k = p.col[s:e], p.kind[s:e]
        lo = c[(k == KIND_LOCAL) & (c < i)]
        fl[i] = fl[lo].max() + 1 if lo.size else 0
    for i in range(p.nrows - 1, -1, -1):
        s, e = p.row_ptr[i], p.row_ptr[i + 1]
        c, k = p.col[s:e], p.kind[s:e]
        up = c[(k == KIND_LOCAL) & (c > i)]
        bl[i] = bl[up].max() + 1 if up.size else 0
    return fl, bl
