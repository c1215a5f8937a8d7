"""The launch-shape autotuner (cdg_prec_tune / cdg_mat_tune): the tuned
shapes give the oracle's results to the same tolerances as the defaults, and
objects created afterwards with the same block size (and precision) start
from the tuned shape."""
import numpy as np
import pytest

import pyoracle as po
from chidg_amd import _lib
from chidg_amd.linsolve import Context, chidg_matrix, chidg_mv, precon_ILU0
from helpers import dense, synthetic_system

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


@pytest.mark.parametrize("dims,degree,precision,schedule,tol", [
    ((4, 3, 2), 2, "fp64", "levels", 1e-12),
    ((4, 3, 2), 2, "fp32", "folded", 1e-5),
    ((4, 4, 4), 1, "fp32", "levels", 1e-5),
    ((2, 3, 2), 3, "fp64", "folded", 1e-12),
    ((3, 3, 3), 1, "bf16", "levels", 5e-2),
])
def test_prec_tune_keeps_parity(ctx, dims, degree, precision, schedule, tol):
    p, blocks = synthetic_system(*dims, degree, seed=23)
    A = upload(ctx, p, blocks)
    M = precon_ILU0(precision, schedule=schedule).init(A)
    M.update(A)
    L_ref, D_ref = po.ilu0_factor(p, blocks)
    v = po.fill_vector(p.vec_len, 4)
    z_ref = po.ilu0_apply(p, blocks, L_ref, D_ref, v)
    t = M.tune(A, reps=2)
    assert t > 0
    shape = M.launch_shape()
    assert shape["slices"] * shape["slice_width"] >= p.n > (shape["slices"] - 1) * shape["slice_width"]
    assert shape["slice_width"] % 8 == 0 and shape["loads_in_flight"] in (4, 8) and shape["threads"] > 0
    z = M.apply(A, v)
    assert rel(z, z_ref) < tol
    assert np.array_equal(M.apply(A, v), z)
    # a new preconditioner of this block size and precision starts from the tuned shape
    M2 = precon_ILU0(precision).init(A)
    assert M2.launch_shape() == shape
    M2.update(A)
    assert rel(M2.apply(A, v), z_ref) < tol
    M.tear_down()
    M2.tear_down()


@pytest.mark.parametrize("dims,degree", [((4, 4, 4), 1), ((3, 3, 3), 2), ((2, 3, 2), 3)])
def test_mat_tune_keeps_parity(ctx, dims, degree):
    p, blocks = synthetic_system(*dims, degree, seed=29)
    A = upload(ctx, p, blocks)
    x = np.random.default_rng(3).standard_normal(p.vec_len)
    y_ref = dense(p, blocks) @ x
    t = A.tune(reps=3)
    assert t > 0
    shape = A.launch_shape()
    assert shape["threads"] > 0 and shape["loads_in_flight"] in (4, 8, 16)
    assert rel(chidg_mv(A, x), y_ref) < 1e-13
    A2 = upload(ctx, p, blocks)
    assert A2.launch_shape() == shape
    assert rel(chidg_mv(A2, x), y_ref) < 1e-13


def test_tune_before_update_is_an_error(ctx):
    p, blocks = synthetic_system(3, 2, 2, 1, seed=31)
    A = upload(ctx, p, blocks)
    M = precon_ILU0("fp64").init(A)
    with pytest.raises(_lib.CdgError):
        M.tune(A)
    M.tear_down()


@pytest.mark.parametrize("degree,sw,precision,tol", [
    (2, 16, "fp32", 1e-5),     # 9 slices: a gather batch of 8 and one of 1
    (2, 16, "fp64", 1e-12),
    (3, 16, "fp32", 1e-5),     # 20 slices: batches of 8, 8, 4
    (1, 8, "bf16", 5e-2),      # 5 slices of 8 columns
])
def test_many_slices_match_oracle(ctx, monkeypatch, degree, sw, precision, tol):
    """Narrow column slices (CDG_APPLY_SLICE_W) exercise the multi-batch
    gather and the accumulation-type partial sums; level and folded schedules
    against the oracle."""
    monkeypatch.setenv("CDG_APPLY_SLICE_W", str(sw))
    p, blocks = synthetic_system(3, 2, 2, degree, seed=41)
    A = upload(ctx, p, blocks)
    L_ref, D_ref = po.ilu0_factor(p, blocks)
    v = po.fill_vector(p.vec_len, 5)
    z_ref = po.ilu0_apply(p, blocks, L_ref, D_ref, v)
    for schedule in (("levels",) if precision == "bf16" else ("levels", "folded")):
        M = precon_ILU0(precision, schedule=schedule).init(A)
        shape = M.launch_shape()
        assert shape["slice_width"] == sw and shape["slices"] == -(-p.n // sw)
        M.update(A)
        z = M.apply(A, v)
        assert rel(z, z_ref) < tol, schedule
        assert np.array_equal(M.apply(A, v), z)
        M.tear_down()
