"""The block-inverse kernels are chosen by block size (and CDG_INV, read once
per process): the global-memory blocked inverse serves n > 192 by default and
any n under CDG_INV=2.  This runs the factor and apply in a child process per
setting at block sizes that are not multiples of 16 (n = 5, 40, 135, 243:
partial last panels, lanes outside the block) and that are (n = 64, 192), with
diagonal blocks that need row pivots on every other row, against the oracle
(precon_ILU0.f90:111-265, mod_inv.f90:99-150).  n = 243 also takes the factor
GEMMs' 64-edge tiles with padding (4 x 64 = 256)."""
import os
import subprocess
import sys

import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)

CHILD = r'''
import sys
sys.path[:0] = [ROOT, ROOT + "/oracle", ROOT + "/tests"]
import numpy as np
import pyoracle as po
from chidg_amd import pattern as pt
from chidg_amd.linsolve import Context, chidg_matrix, precon_ILU0
from helpers import synthetic_system

def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))

ctx = Context(0)
for degree, nfields in ((0, 5), (1, 5), (2, 5), (1, 8), (3, 3), (2, 9)):   # n = 5, 40, 135, 64, 192, 243
    for pivots in (False, True):
        p, blocks = synthetic_system(3, 2, 2, degree, seed=41 + degree, nfields=nfields)
        n = p.n
        if pivots:
            B = blocks.reshape(p.nnzb, n, n)
            rows = np.repeat(np.arange(p.nrows), np.diff(p.row_ptr))
            shift = np.zeros((n, n))
            shift[np.arange(n), (np.arange(n) + 1) % n] = 2.0
            rng = np.random.default_rng(9)
            for b in np.flatnonzero((p.kind == pt.KIND_LOCAL) & (p.col == rows) & (rows % 2 == 0)):
                B[b] = (shift + 0.1 * rng.standard_normal((n, n)) / np.sqrt(n)).T
        A = chidg_matrix(p, ctx)
        A.store(blocks)
        L_ref, D_ref = po.ilu0_factor(p, blocks)
        v = po.fill_vector(p.vec_len, 3)
        z_ref = po.ilu0_apply(p, blocks, L_ref, D_ref, v)
        for precision, tol in (("fp64", 1e-12), ("fp32", 1e-5)):
            M = precon_ILU0(precision).init(A)
            M.update(A)
            _, D = M.factors(A)
            eD, ez = rel(D, D_ref), rel(M.apply(A, v), z_ref)
            print(f"n={n} pivots={pivots} {precision}: Dinv {eD:.1e}, apply {ez:.1e}", flush=True)
            assert eD < tol and ez < tol, (n, pivots, precision, eD, ez)
            M.tear_down()
        A.release()
ctx.close()
print("OK", flush=True)
'''.replace("ROOT", repr(ROOT))


@pytest.mark.parametrize("env", [{"CDG_INV": "2"}, {"CDG_INV": "2", "CDG_GINV_TILES": "4"},
                                 {"CDG_INV": "2", "CDG_GINV_TILES": "8"}, {"CDG_INV": "0"}])
def test_inverse_kernels_vs_oracle(env):
    e = dict(os.environ, **env)
    r = subprocess.run([sys.executable, "-c", CHILD], env=e, capture_output=True, text=True, timeout=240)
    print(r.stdout[-3000:], r.stderr[-2000:])
    assert r.returncode == 0 and r.stdout.rstrip().endswith("OK")
