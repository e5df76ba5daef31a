"""Newton failure recovery of the CPU restatement, the checker of the GPU
driver's: checkNewtonConvergence + emergencyNewtonRaphson
(src/Newton_Raphson_mod.F90:453-584).  Too many iterations or a correction
above the divergence norm restart the solve from the step's prediction with
the relaxation xF and the step dS halved and the iteration cap raised to 50x;
below xF = 1e-2 the run stops.  Both conditions are forced from the host (a
small cap, a tiny divergence norm) on the 240-element mesh."""
import numpy as np
import pytest

from newton_cpu import newton_cpu_arc
from ve_flow_around_cylinder_fortran_amd import cylinder_channel
from ve_flow_around_cylinder_fortran_amd.newton import initial_state

OLDROYD_B = dict(ReN=0.0, BvN=0.59, VeN=0.0, DfN=1e-3, WiN0=1.0, model="L-PTT")


def test_iteration_cap_restarts_relaxed_and_converges(oracle):
    mesh = cylinder_channel(8, 4, 6, 8)
    x0 = initial_state(mesh, 0.1)
    em = []
    x, hist, _ = newton_cpu_arc(oracle, mesh, OLDROYD_B, x0, 0.1, 1, 0.0, None, None, 1.0, max_iter=2,
                                emergencies=em)
    assert len(em) == 1 and em[0][:3] == (3, "iterations", 0.5) and em[0][3] == 0.5
    # after the restart: a fresh Jacobian, then relaxed iterations to convergence
    assert hist[3][1] == "NRP" and hist[3][0] == 1
    assert hist[-1][3] < 1e-8 and hist[-1][2] < 1e-8
    # the same converged state as the unrelaxed solve
    xr, hr, _ = newton_cpu_arc(oracle, mesh, OLDROYD_B, x0, 0.1, 1, 0.0, None, None, 1.0)
    assert np.abs(x - xr).max() <= 1e-7 * np.abs(xr).max()
    assert len(hist) > len(hr) + 3


def test_divergence_halves_relaxation_until_stop(oracle):
    mesh = cylinder_channel(8, 4, 6, 8)
    x0 = initial_state(mesh, 0.1)
    em = []
    with pytest.raises(RuntimeError, match="relaxation factor below 1e-2"):
        newton_cpu_arc(oracle, mesh, OLDROYD_B, x0, 0.1, 1, 0.0, None, None, 1.0, cor_norm_div=1e-30,
                       emergencies=em)
    # xF = 1/2 ... 1/128 < 1e-2: seven restarts, each after one iteration
    assert [e[2] for e in em] == [0.5 ** k for k in range(1, 8)]
    assert [e[0] for e in em] == list(range(1, 8)) and all(e[1] == "divergence" for e in em)
