"""f3: the stability kernel's arithmetic (csrc/vef_stability_math.h) built for
the host and run over whole elements, bit for bit against the oracle's
storeBulkStability element matrices (tests/stability/stab_host_check.hip
repeats the kernel's loop).  Geometry and basis tables come from the oracle,
whose tables and basis2D the GPU reproduces bit for bit (tests/test_gpu_parity.py).
Entries that are zero are compared as +0: the matrix receives them by
addition into zeroed storage."""
import ctypes
import os
import shutil
import subprocess

import numpy as np
import pytest

from test_stability_oracle import PARAM_SETS, affine_xy, div_free_state

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(os.path.dirname(HERE), "ve_flow_around_cylinder_fortran_amd", "csrc")
ORACLE = os.path.join(os.path.dirname(HERE), "oracle")


@pytest.fixture(scope="module")
def host_lib(tmp_path_factory):
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(hipcc):
        pytest.skip("hipcc not available")
    out = str(tmp_path_factory.mktemp("stab") / "libstab_host.so")
    subprocess.run([hipcc, "-O2", "-std=c++17", "-fPIC", "-shared", "--offload-arch=gfx950", "-ffp-contract=off",
                    "-fno-fast-math", "-I", CSRC, "-I", ORACLE, "-o", out,
                    os.path.join(HERE, "stability", "stab_host_check.hip")],
                   check=True, capture_output=True)
    L = ctypes.CDLL(out)
    L.stab_host_element.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_double] + \
        [ctypes.c_void_p] * 9
    L.stab_count_node_coef.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_double, ctypes.c_double,
                                       ctypes.c_double, ctypes.c_void_p]
    return L


def _p(a):
    return ctypes.c_void_p(a.ctypes.data)


def host_element(L, oracle, pset, xy, TLb, EU):
    bf = np.zeros((9, 9))
    dFx = np.zeros((9, 9))
    dFy = np.zeros((9, 9))
    d2x = np.zeros((9, 9))
    d2y = np.zeros((9, 9))
    H2 = np.zeros(9)
    WET = np.zeros(9)
    tab = oracle.tables()
    for g in range(9):
        geo = oracle.basis2d(xy, g)
        bf[g] = tab["bfn"][g]
        dFx[g], dFy[g] = geo["dFdXi"][0], geo["dFdXi"][1]
        d2x[g], d2y[g] = geo["d2FdXi2"][0], geo["d2FdXi2"][1]
        H = 0.0
        for n in range(9):
            H = dFx[g, n] * dFx[g, n] + dFy[g, n] * dFy[g, n] if n == 0 else H + dFx[g, n] * dFx[g, n] + dFy[g, n] * dFy[g, n]
        He = 1.0 / np.sqrt(H)
        H2[g] = He * He
        WET[g] = geo["WET"]
    prm5 = np.array([pset["ReN"], pset["BvN"], pset["VeN"], pset["DfN"], pset["WiN0"]])
    A = np.zeros((54, 54))
    B = np.zeros((54, 54))
    u = np.ascontiguousarray(TLb, dtype=np.float64)
    L.stab_host_element(_p(prm5), 0 if pset["model"] == "L-PTT" else 1, _p(u), float(EU), _p(bf), _p(dFx), _p(dFy),
                        _p(d2x), _p(d2y), _p(H2), _p(WET), _p(A), _p(B))
    return A, B


def canon(M):
    return (M + 0.0).view(np.uint64)


@pytest.mark.parametrize("pset", PARAM_SETS)
def test_host_math_matches_oracle_bitwise(host_lib, oracle, pset):
    prm = oracle.params(**pset)
    rng = np.random.default_rng(21)
    for trial in range(3):
        xy, _ = affine_xy(rng)
        if trial:
            xy[4:9] += 0.03 * rng.standard_normal((5, 2))
        TLb = div_free_state(rng, xy, 0.3, 1.2) + 0.02 * rng.standard_normal((9, 6))
        EU = 0.9 + 0.1 * trial
        Ao, Bo = oracle.stab_element(prm, xy, TLb, EU)
        Ah, Bh = host_element(host_lib, oracle, pset, xy, TLb, EU)
        na = int(np.count_nonzero(canon(Ao) != canon(Ah)))
        nb = int(np.count_nonzero(canon(Bo) != canon(Bh)))
        assert na == 0 and nb == 0, (trial, na, nb, np.abs(Ao - Ah).max(), np.abs(Bo - Bh).max())


def test_node_coefficient_operation_count(host_lib):
    """The per-(Gauss point, test node) coefficients computed once per element
    instead of once per disturbance: their operation count (the oracle's
    counting rules) is the constant the stability roofline subtracts for the
    53 disturbances that reuse them."""
    import importlib.util

    spec = importlib.util.spec_from_file_location("bench", os.path.join(os.path.dirname(HERE), "bench.py"))
    bench = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(bench)
    prm5 = np.array([0.3, 0.59, 0.25, 1e-3, 1.0])
    base10 = np.array([0.7, -0.2, 1.1, 0.05, 0.9, 0.6, 0.3, 0.02, 0.4, 1.2])
    out = np.zeros(4, dtype=np.uint64)
    host_lib.stab_count_node_coef(_p(prm5), _p(base10), 0.25, 1.5, -0.75, _p(out))
    add, mul, div, total = map(int, out)
    assert div == 11 and total == add + mul + div
    assert total == bench.STAB_COEF_FLOPS
