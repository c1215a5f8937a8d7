"""The allocation-free apply (cdg_prec_apply_noalloc) through the C ABI:
the same bits as the existing apply entry point (cdg_prec_apply_async under
the row-slice schedule) on calls in a row, the second on a smaller system, in
every ILU(0) precision; calls it cannot serve without allocating, and sizes
past its scratch array, are refused."""
import ctypes as C

import numpy as np
import pytest

from chidg_amd import _lib
from chidg_amd import pattern as pt
from chidg_amd.linsolve import Context, chidg_matrix, precon_ILU0

pytestmark = pytest.mark.gpu

CDG_ERR_ARG, CDG_ERR_STATE = 1, 7
SCRATCH_BYTES = 32 << 20          # CDG_NOALLOC_SCRATCH_BYTES


def _system(ctx, dims, order, seed):
    import torch
    n = pt.block_size(order)
    p = pt.hex_box(*dims, n)
    A = chidg_matrix(p, ctx)
    A.fill_synthetic(seed, 1 / np.sqrt(n), 0.1 / np.sqrt(n))
    v = A.synthetic_vector(seed + 1, like=torch.empty(0, dtype=torch.float64, device="cuda:0"))
    return A, v


def _noalloc(ctx, M, A, v, z):
    return ctx.lib.cdg_prec_apply_noalloc(ctx.h, M.h, A.h, C.c_void_p(v.data_ptr()), C.c_void_p(z.data_ptr()))


@pytest.mark.parametrize("precision", ["fp64", "fp32", "bf16"])
def test_noalloc_apply_matches_apply(precision):
    import torch
    ctx = Context(0)
    big = _system(ctx, (8, 8, 6), 2, 31)
    small = _system(ctx, (4, 3, 3), 1, 37)
    torch.cuda.synchronize()
    for A, v in (big, small, big):
        Mref = precon_ILU0(precision, schedule="rowslice").init(A)
        M = precon_ILU0(precision, schedule="rowslice").init(A)
        Mref.update(A)
        M.update(A)
        z_ref = torch.empty_like(v)
        Mref.apply_async(A, v, z_ref, order=False)
        torch.cuda.synchronize()
        for _ in range(2):           # the first call forms W and the copies, the second reuses them
            z = torch.full_like(v, float("nan"))
            torch.cuda.synchronize()
            assert _noalloc(ctx, M, A, v, z) == _lib.CDG_OK, ctx.lib.cdg_last_error(ctx.h)
            torch.cuda.synchronize()
            assert torch.equal(z, z_ref), precision
        Mref.tear_down()
        M.tear_down()
    for A, _ in (big, small):
        A.release()
    ctx.close()


def _chain(nrows, n):
    """1-D chain pattern: row i couples to i-1 and i+1 (no wrap)."""
    cnt = np.full(nrows, 3, dtype=np.int32)
    cnt[0] = cnt[-1] = 2
    row_ptr = np.zeros(nrows + 1, dtype=np.int32)
    row_ptr[1:] = np.cumsum(cnt)
    idx = np.arange(nrows, dtype=np.int32)
    col = np.empty(int(row_ptr[-1]), dtype=np.int32)
    has_lo = idx > 0
    col[row_ptr[:-1][has_lo]] = idx[has_lo] - 1
    dpos = row_ptr[:-1] + has_lo.astype(np.int32)
    col[dpos] = idx
    has_up = idx < nrows - 1
    col[dpos[has_up] + 1] = idx[has_up] + 1
    return pt.Pattern(nrows=nrows, n=n, row_ptr=row_ptr, col=col, kind=np.zeros_like(col))


def test_noalloc_apply_refusals():
    """Without row-slice storage (the call would have to allocate it) and
    before an update the call is refused with CDG_ERR_STATE; a system whose
    FP64 forward vector is one block row past the scratch array, with
    CDG_ERR_ARG."""
    import torch
    ctx = Context(0)
    A, v = _system(ctx, (4, 4, 4), 1, 41)
    z = torch.empty_like(v)
    M = precon_ILU0("fp64", schedule="levels").init(A)
    M.update(A)
    assert _noalloc(ctx, M, A, v, z) == CDG_ERR_STATE
    M.tear_down()
    M = precon_ILU0("fp64", schedule="rowslice").init(A)
    assert _noalloc(ctx, M, A, v, z) == CDG_ERR_STATE          # not updated
    M.tear_down()
    A.release()
    n = 4
    nrows = SCRATCH_BYTES // (8 * n) + 1
    p = _chain(nrows, n)
    A = chidg_matrix(p, ctx)
    M = precon_ILU0("fp64").init(A)
    vb = torch.zeros(p.vec_len, dtype=torch.float64, device="cuda:0")
    zb = torch.empty_like(vb)
    assert _noalloc(ctx, M, A, vb, zb) == CDG_ERR_ARG
    assert b"scratch" in ctx.lib.cdg_last_error(ctx.h)
    M.tear_down()
    A.release()
    ctx.close()
