"""Host logic of the inexact Newton's refactorisation threshold
(NewtonSolver reuse_max_iters="auto"): the number of modified-Newton
iterations a factorisation is worth, from running means of the measured
iteration times."""
import pytest

from ve_flow_around_cylinder_fortran_amd.newton import NewtonError, NewtonRecord, NewtonSolver


def _policy(limit):
    s = object.__new__(NewtonSolver)  # the timing logic only: no GPU objects
    s.reuse_max_iters = limit
    s._t_nrp = s._t_mnr = None
    return s


def test_fixed_limit_is_kept():
    s = _policy(8.0)
    s._time_iteration(NewtonRecord(1, "NRP", 1.0, 1.0, 0.30))
    s._time_iteration(NewtonRecord(2, "MNR", 1.0, 1.0, 0.05))
    assert s._reuse_limit() == 8.0


def test_auto_limit_from_measured_times():
    s = _policy("auto")
    assert s._reuse_limit() == 8.0  # nothing measured yet
    s._time_iteration(NewtonRecord(1, "NRP", 1.0, 1.0, 0.25))
    s._time_iteration(NewtonRecord(2, "MNR", 1.0, 1.0, 0.04))
    assert s._reuse_limit() == pytest.approx((0.25 - 0.04) / 0.04)
    # running means: the newest iteration weighs 0.3
    s._time_iteration(NewtonRecord(3, "MNR", 1.0, 1.0, 0.05))
    t_mnr = 0.7 * 0.04 + 0.3 * 0.05
    assert s._reuse_limit() == pytest.approx((0.25 - t_mnr) / t_mnr)
    # a factorisation cheaper than a modified-Newton iteration: at least one
    s._time_iteration(NewtonRecord(4, "NRP", 1.0, 1.0, 0.0))
    s._time_iteration(NewtonRecord(5, "NRP", 1.0, 1.0, 0.0))
    s._time_iteration(NewtonRecord(6, "NRP", 1.0, 1.0, 0.0))
    s._time_iteration(NewtonRecord(7, "NRP", 1.0, 1.0, 0.0))
    assert s._reuse_limit() == 1.0


@pytest.mark.parametrize("limit", [0, -1.0, "fast"])
def test_bad_limit_rejected(limit):
    with pytest.raises(NewtonError, match="reuse_max_iters"):  # checked before anything touches the GPU
        NewtonSolver.__init__(object.__new__(NewtonSolver), None, None, reuse_max_iters=limit)


def test_krylov_options_checked():
    with pytest.raises(NewtonError, match="reuse_factors"):
        NewtonSolver.__init__(object.__new__(NewtonSolver), None, None, krylov=True)
    for tol, its in ((0.0, 10), (1.0, 10), (1e-3, 0)):
        with pytest.raises(NewtonError, match="krylov_tol"):
            NewtonSolver.__init__(object.__new__(NewtonSolver), None, None, reuse_factors=True, krylov=True,
                                  krylov_tol=tol, krylov_max_its=its)


def test_refine_tol_checked():
    """refine_tol sets the FP32 factorisation's GMRES target: only for FP32,
    within (0, 1), checked before any GPU work."""
    for prec, tol in (("f64", 1e-2), ("f32", 0.0), ("f32", 1.0)):
        with pytest.raises(NewtonError, match="refine_tol"):
            NewtonSolver.__init__(object.__new__(NewtonSolver), None, None, linear_precision=prec, refine_tol=tol)


def test_krylov_iterations_time_as_carried_factor_iterations():
    s = _policy("auto")
    s._time_iteration(NewtonRecord(1, "NRP", 1.0, 1.0, 0.20))
    s._time_iteration(NewtonRecord(2, "KRY", 1.0, 1.0, 0.05))
    assert s._reuse_limit() == pytest.approx((0.20 - 0.05) / 0.05)
