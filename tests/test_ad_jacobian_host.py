"""The analytic Jacobian's Gauss-point function (csrc/vef_ad_math.h, the
coefficients T of the test functions as functions of the interpolated
fields) built for the host in double (tests/jacobian/ad_host_check.hip):

* the element residual it implies equals the oracle's residualBulk to 1e-12
  of the residual's scale (pins T against the reference restatement: every
  term, model and inertia set);
* its forward-mode Jacobian equals the oracle's forward-difference Jacobian
  (storeBulk, eps_jac = 1e-8) and dR/dEU to the difference-quotient error
  (1e-6 of the matrix scale, norm-wise), and dropping a single derivative
  (one column) is caught at that tolerance;
* the GPU kernel's scheme (DualV tangents, 19 directions, the second
  derivatives of U from -BvN times the pressure-gradient direction) equals
  differentiating every one of the 23 directions with Dual to rounding, and
  a wrong factor in that shortcut is caught."""
import ctypes
import os
import shutil
import subprocess

import numpy as np
import pytest

from test_oracle_element import PARAM_SETS as ELEMENT_SETS
from test_oracle_element import curved_xy, random_state

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(os.path.dirname(HERE), "ve_flow_around_cylinder_fortran_amd", "csrc")
PSETS = [p for p in ELEMENT_SETS if p["DfN"] > 1e-12]


@pytest.fixture(scope="module")
def ad_lib(tmp_path_factory):
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(hipcc):
        pytest.skip("hipcc not available")
    out = str(tmp_path_factory.mktemp("ad") / "libad_host.so")
    subprocess.run([hipcc, "-O2", "-std=c++17", "-fPIC", "-shared", "--offload-arch=gfx950", "-I", CSRC, "-I",
                    os.path.join(os.path.dirname(HERE), "oracle"), "-o", out,
                    os.path.join(HERE, "jacobian", "ad_host_check.hip")], check=True, capture_output=True)
    L = ctypes.CDLL(out)
    L.ad_host_element.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_double] + \
        [ctypes.c_void_p] * 10 + [ctypes.c_int]
    L.ad_count_gauss_terms.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                                       ctypes.c_void_p, ctypes.c_double, ctypes.c_int, ctypes.c_void_p]
    return L


def _p(a):
    return ctypes.c_void_p(a.ctypes.data)


def ad_element(L, oracle, pset, xy, TL, EU, scheme=1, bvn_factor=1.0):
    tab = oracle.tables()
    bf = np.ascontiguousarray(tab["bfn"])
    dFx, dFy, d2x, d2y = (np.zeros((9, 9)) for _ in range(4))
    H2, WET = np.zeros(9), np.zeros(9)
    for g in range(9):
        geo = oracle.basis2d(xy, g)
        dFx[g], dFy[g] = geo["dFdXi"]
        d2x[g], d2y[g] = geo["d2FdXi2"]
        H2[g] = 1.0 / np.sum(dFx[g] ** 2 + dFy[g] ** 2)
        WET[g] = geo["WET"]
    prm5 = np.array([pset["ReN"], pset["BvN"], pset["VeN"], pset["DfN"], pset["WiN0"]])
    R, J, Jeu = np.zeros(54), np.zeros((54, 54)), np.zeros(54)
    u = np.ascontiguousarray(TL, dtype=np.float64)
    L.ad_host_element(_p(prm5), 0 if pset["model"] == "L-PTT" else 1, _p(u), float(EU), _p(bf), _p(dFx), _p(dFy),
                      _p(d2x), _p(d2y), _p(H2), _p(WET), _p(R), _p(J), _p(Jeu), int(scheme))
    return R, J, Jeu


@pytest.mark.parametrize("scheme", [0, 1])
@pytest.mark.parametrize("pset", PSETS)
def test_residual_and_jacobian_match_oracle(ad_lib, oracle, pset, scheme):
    prm = oracle.params(**pset)
    rng = np.random.default_rng(31)
    for trial in range(3):
        xy = curved_xy(rng)
        TL = random_state(rng, xy)
        EU = 0.6 + 0.2 * trial
        Res, A, Ac = oracle.element(prm, xy, TL, EU)
        R, J, Jeu = ad_element(ad_lib, oracle, pset, xy, TL, EU, scheme=scheme)
        assert np.abs(R - Res.reshape(-1)).max() <= 1e-12 * np.abs(Res).max()
        assert np.linalg.norm(J - A) <= 1e-6 * np.linalg.norm(A), np.linalg.norm(J - A) / np.linalg.norm(A)
        assert np.linalg.norm(Jeu - Ac) <= 1e-6 * np.linalg.norm(Ac)


@pytest.mark.parametrize("pset", PSETS)
def test_kernel_scheme_equals_all_directions(ad_lib, oracle, pset):
    """19 directions + the second-derivative shortcut = all 23 directions, to
    rounding (1e-13 of the matrix scale entry-wise)."""
    rng = np.random.default_rng(33)
    for trial in range(3):
        xy = curved_xy(rng)
        TL = random_state(rng, xy)
        _, J0, E0 = ad_element(ad_lib, oracle, pset, xy, TL, 0.7, scheme=0)
        _, J1, E1 = ad_element(ad_lib, oracle, pset, xy, TL, 0.7, scheme=1)
        assert np.abs(J1 - J0).max() <= 1e-13 * np.abs(J0).max()
        assert np.abs(E1 - E0).max() <= 1e-13 * np.abs(E0).max()


def test_second_derivative_columns_are_visible(ad_lib, oracle):
    """The second-derivative part of the velocity columns is large enough for
    the tolerance above to catch a wrong factor in the shortcut: removing it
    (the columns of scheme 0 without the d2U contributions, i.e. with the
    d2x/d2y basis set to zero) moves the Jacobian by more than 1e-6 of its
    scale."""
    pset = PSETS[0]
    rng = np.random.default_rng(34)
    xy = curved_xy(rng)
    TL = random_state(rng, xy)
    _, J0, _ = ad_element(ad_lib, oracle, pset, xy, TL, 0.7, scheme=0)
    orig = oracle.basis2d

    def flat(xy_, g):
        geo = dict(orig(xy_, g))
        geo["d2FdXi2"] = np.zeros_like(np.asarray(geo["d2FdXi2"]))
        return geo

    oracle.basis2d = flat
    try:
        _, Jn, _ = ad_element(ad_lib, oracle, pset, xy, TL, 0.7, scheme=0)
    finally:
        oracle.basis2d = orig
    assert np.abs(Jn - J0).max() > 1e-6 * np.abs(J0).max()


def test_fd_tolerance_catches_a_dropped_derivative(ad_lib, oracle):
    """Dropping a single derivative (the S11 column of the first node) fails
    the tolerance above by more than a factor of ten, so it is not vacuous."""
    pset = PSETS[0]
    prm = oracle.params(**pset)
    rng = np.random.default_rng(32)
    xy = curved_xy(rng)
    TL = random_state(rng, xy)
    _, A, _ = oracle.element(prm, xy, TL, 0.7)
    _, J, _ = ad_element(ad_lib, oracle, pset, xy, TL, 0.7)
    J2 = J.copy()
    J2[:, 3] = 0.0
    assert np.linalg.norm(J2 - A) > 1e-5 * np.linalg.norm(A)


def test_fp32_jacobian_operation_count(ad_lib):
    """The FP32 Jacobian's work per element that bench.py reports against the
    FP32 peak: 171 dual-number evaluations of the Gauss-point terms (9 Gauss
    points x 19 directions; every executed operation counted with the
    oracle's instrumented scalar, the same for every direction), the 22 field
    interpolations per Gauss point, the 18 products by WET per evaluation,
    the forming of E (37 operations per (k, q, m): 5 multiplications and 3
    additions for each velocity unknown, one shared sum of the second-
    derivative basis values, 3 multiplications and 2 additions for each of
    the other four) and the outer contraction (3 products and sums per test
    node and column of E, 330 columns)."""
    import importlib.util

    spec = importlib.util.spec_from_file_location("bench", os.path.join(os.path.dirname(HERE), "bench.py"))
    bench = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(bench)
    prm5 = np.array([0.0, 0.59, 0.0, 1e-3, 1.0])  # the bench workload: Oldroyd-B, Re = 0, DfN = 1e-3
    rng = np.random.default_rng(9)
    F = rng.uniform(0.2, 1.2, 23)
    dFx, dFy = rng.standard_normal(9), rng.standard_normal(9)
    counts = set()
    for direction in (0, 5, 11, 22):  # (the counting pass does not depend on the direction)
        out = np.zeros(5, dtype=np.uint64)
        ad_lib.ad_count_gauss_terms(_p(prm5), 0, _p(F), _p(dFx), _p(dFy), 0.3, direction, _p(out))
        counts.add(int(out[4]))
    assert counts == {bench.AD32_DUAL_FLOPS}
    e_forming = 18 * 9 * (2 * (5 + 3) + 1 + 4 * (3 + 2))
    assert e_forming == 162 * 37
    contraction = 9 * (e_forming + (2916 + 54) * 6)
    assert contraction == 214326
    assert bench.AD32_FLOPS_PER_ELEMENT == 171 * bench.AD32_DUAL_FLOPS + 9 * 22 * 9 * 2 + 171 * 18 + contraction
