"""solver_controller_t%update_preconditioner (type_solver_controller.f90:89-124):
the refactor decision on the host, without a GPU."""
from chidg_amd.linsolve import solver_controller


class _Stamped:
    def __init__(self, stamp):
        self.stamp = stamp


def test_first_call_updates_then_skips():
    c, A, M = solver_controller(), _Stamped(5), _Stamped(0)
    assert c.update_preconditioner(A, M) and c.preconditioner_updated
    assert c.lhs_stamp == 5
    M.stamp = 5                      # the solve factored with this matrix version
    assert not c.update_preconditioner(A, M) and not c.preconditioner_updated


def test_changed_lhs_updates():
    c, A, M = solver_controller(), _Stamped(5), _Stamped(5)
    c.lhs_stamp = 5
    A.stamp = 6                      # chidg_matrix%store re-stamps the lhs
    assert c.update_preconditioner(A, M)
    assert c.lhs_stamp == 6


def test_force_flag_is_one_shot():
    c, A, M = solver_controller(), _Stamped(5), _Stamped(5)
    c.lhs_stamp = 5
    c.force_update_preconditioner = True
    assert c.update_preconditioner(A, M)
    assert not c.force_update_preconditioner
    assert not c.update_preconditioner(A, M)


def test_preconditioner_from_another_matrix_version_updates():
    """lhs unchanged since the last call, but the preconditioner was updated
    with an older version (e.g. a solve without the controller in between)."""
    c, A, M = solver_controller(), _Stamped(7), _Stamped(6)
    c.lhs_stamp = 7
    assert c.update_preconditioner(A, M)


def test_update_lhs_default_always():
    c = solver_controller()
    assert c.update_lhs(_Stamped(1), 3, 0.5) and c.lhs_updated
