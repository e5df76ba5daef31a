"""a7-a10, a12: element residuals and forward-difference Jacobians of the
oracle, pinned by properties the equations fix:

* frame invariance: rotating and translating an element and its state rotates
  momentum rows as vectors and leaves continuity rows and the trace of the
  stress rows unchanged; quarter turns permute the stress rows exactly
  (catches transposed operands and wrong index pairs in every tensor term,
  stabilization included);
* mirror symmetry y -> -y (odd terms keep their signs);
* test-function partition: sum over test nodes of the momentum rows is the
  integral of the convective term (all other momentum terms are linear in
  grad phi), of the continuity rows the integral of div u;
* the FD columns against central differences of the residual (index/transpose
  of the element Jacobian) and, bit for bit, against residuals of explicitly
  built perturbed states with the restore drift fl(fl(u+h)-h) of the entries
  perturbed before (storeBulk, src/Equations_mod.F90:267-306);
* the outlet traction against its closed form for a uniform stress.
"""
import numpy as np
import pytest

from conftest import reference_nodes

NODES = reference_nodes()
H_JAC = 1.0e-8


def curved_xy(rng, scale=0.4):
    A = scale * (np.eye(2) + 0.2 * rng.standard_normal((2, 2)))
    xy = NODES @ A.T + rng.standard_normal(2)
    xy[4:9] += 0.03 * scale * rng.standard_normal((5, 2))
    return xy


def random_state(rng, xy, amp=0.3):
    """Smooth nodal fields with S symmetric positive definite."""
    x, y = xy[:, 0], xy[:, 1]
    TL = np.zeros((9, 6))
    c = rng.standard_normal((6, 3))
    for q in range(6):
        TL[:, q] = amp * (c[q, 0] + c[q, 1] * x + c[q, 2] * y) + 0.05 * rng.standard_normal(9)
    TL[:, 3] += 1.2
    TL[:, 5] += 1.0
    TL[:, 4] *= 0.3
    return TL


def rotate_state(TL, Q):
    out = TL.copy()
    out[:, 0:2] = TL[:, 0:2] @ Q.T
    for n in range(9):
        S = np.array([[TL[n, 3], TL[n, 4]], [TL[n, 4], TL[n, 5]]])
        Sr = Q @ S @ Q.T
        out[n, 3:6] = Sr[0, 0], Sr[0, 1], Sr[1, 1]
    return out


PARAM_SETS = [
    dict(ReN=0.0, BvN=0.59, VeN=0.0, DfN=1e-3, WiN0=1.0, model="L-PTT"),
    dict(ReN=0.8, BvN=0.59, VeN=0.25, DfN=1e-3, WiN0=1.0, model="L-PTT"),
    dict(ReN=0.3, BvN=0.4, VeN=0.1, DfN=1e-3, WiN0=2.0, model="GIESEKUS"),
    dict(ReN=0.0, BvN=0.59, VeN=0.0, DfN=0.0, WiN0=1.0, model="L-PTT"),  # DCS on
]


@pytest.mark.parametrize("pset", range(len(PARAM_SETS)))
def test_frame_invariance(oracle, pset):
    rng = np.random.default_rng(20 + pset)
    prm = oracle.params(**PARAM_SETS[pset])
    for trial in range(4):
        xy = curved_xy(rng)
        TL = random_state(rng, xy)
        th = rng.uniform(0, 2 * np.pi)
        Q = np.array([[np.cos(th), -np.sin(th)], [np.sin(th), np.cos(th)]])
        t = rng.standard_normal(2) * 3
        R0, _, _ = oracle.element(prm, xy, TL, 0.6, jac=False)
        R1, _, _ = oracle.element(prm, xy @ Q.T + t, rotate_state(TL, Q), 0.6, jac=False)
        scale = np.abs(R0).max()
        np.testing.assert_allclose(R1[:, 0:2], R0[:, 0:2] @ Q.T, atol=1e-11 * scale)
        np.testing.assert_allclose(R1[:, 2], R0[:, 2], atol=1e-11 * scale)
        # Stress rows: the xy row carries the single tensor component in the
        # Galerkin, LSCE, DfN and DCS terms but both off-diagonal contributions
        # in the LSME term (src/Residuals_mod.F90:449-453), so under a general
        # rotation only the trace row is invariant ...
        np.testing.assert_allclose(R1[:, 3] + R1[:, 5], R0[:, 3] + R0[:, 5], atol=1e-11 * scale)
        # ... while quarter turns map rows exactly: xx <-> yy, xy -> -xy.
        k = int(rng.integers(1, 4))
        Qk = np.linalg.matrix_power(np.array([[0.0, -1.0], [1.0, 0.0]]), k)
        R2, _, _ = oracle.element(prm, xy @ Qk.T + t, rotate_state(TL, Qk), 0.6, jac=False)
        np.testing.assert_allclose(R2[:, 0:2], R0[:, 0:2] @ Qk.T, atol=1e-11 * scale)
        np.testing.assert_allclose(rotate_state(np.c_[np.zeros((9, 3)), R0[:, 3:6]], Qk)[:, 3:6], R2[:, 3:6],
                                   atol=1e-11 * scale)


@pytest.mark.parametrize("pset", range(len(PARAM_SETS)))
def test_mirror_symmetry(oracle, pset):
    rng = np.random.default_rng(40 + pset)
    prm = oracle.params(**PARAM_SETS[pset])
    M = np.diag([1.0, -1.0])
    sgn = np.array([1, -1, 1, 1, -1, 1.0])
    for trial in range(3):
        xy = curved_xy(rng)
        TL = random_state(rng, xy)
        R0, _, _ = oracle.element(prm, xy, TL, 0.6, jac=False)
        R1, _, _ = oracle.element(prm, xy @ M, TL * sgn, 0.6, jac=False)
        np.testing.assert_allclose(R1, R0 * sgn, atol=1e-12 * np.abs(R0).max())


def test_test_function_partition(oracle):
    """Affine element, linear velocity: sum_n R_cont(n) = div(u) * area and,
    with Re = 0, sum_n R_mom(n) = 0; with Re > 0 it is the integral of the
    convection, computed here from the linear field by 3x3 Gauss."""
    rng = np.random.default_rng(5)
    A = 0.5 * (np.eye(2) + 0.2 * rng.standard_normal((2, 2)))
    xy = NODES @ A.T + 0.3
    area = 4 * abs(np.linalg.det(A))
    G = rng.standard_normal((2, 2))  # u = G x + c
    cvec = rng.standard_normal(2)
    TL = random_state(rng, xy)
    TL[:, 0:2] = xy @ G.T + cvec
    for ReN in (0.0, 0.9):
        prm = oracle.params(ReN=ReN, VeN=0.2)
        R, _, _ = oracle.element(prm, xy, TL, 0.7, jac=False)
        assert R[:, 2].sum() == pytest.approx(np.trace(G) * area, rel=1e-12)
        # integral of Re (u.grad)u = Re G u over the element (linear integrand: exact with the centroid)
        T = oracle.tables()
        conv = np.zeros(2)
        for g in range(9):
            xg = A @ T["P2"][g] + 0.3
            conv += T["W2"][g] * abs(np.linalg.det(A)) * (G @ (G @ xg + cvec))
        np.testing.assert_allclose(R[:, 0:2].sum(axis=0), ReN * conv, atol=1e-12 * np.abs(R).max())


def test_fd_columns_match_central_differences(oracle):
    rng = np.random.default_rng(6)
    prm = oracle.params(ReN=0.4, VeN=0.25)
    xy = curved_xy(rng)
    TL = random_state(rng, xy)
    EU = 0.6
    _, A, Ac = oracle.element(prm, xy, TL, EU)
    h = 1e-5
    for col in list(range(0, 54, 5)) + [53]:
        n, q = divmod(col, 6)
        Tp, Tm = TL.copy(), TL.copy()
        Tp[n, q] += h
        Tm[n, q] -= h
        Rp, _, _ = oracle.element(prm, xy, Tp, EU, jac=False)
        Rm, _, _ = oracle.element(prm, xy, Tm, EU, jac=False)
        cd = ((Rp - Rm) / (2 * h)).ravel()
        scale = np.abs(cd).max()
        np.testing.assert_allclose(A[col], cd, atol=2e-6 * scale + 1e-9)
    Rp, _, _ = oracle.element(prm, xy, TL, EU + h, jac=False)
    Rm, _, _ = oracle.element(prm, xy, TL, EU - h, jac=False)
    cd = ((Rp - Rm) / (2 * h)).ravel()
    np.testing.assert_allclose(Ac, cd, atol=2e-6 * np.abs(cd).max())


def drift(u):
    return (u + H_JAC) - H_JAC


@pytest.mark.parametrize("ied", [0, 2])
def test_fd_columns_follow_perturb_restore_sequence(oracle, ied):
    """Column p is ((R(state_p) - R(u))*h)/(h*h) with state_p = drifted entries
    before p, u_p + h at p, original entries after p; the EU column sees all 54
    entries drifted.  A near-zero state makes the drift frequent."""
    rng = np.random.default_rng(7 + ied)
    prm = oracle.params()
    xy = curved_xy(rng)
    TL = random_state(rng, xy)
    TL[:, 0:2] *= 1e-9  # velocities small enough that fl(fl(u+h)-h) != u
    TL[:, 4] *= 1e-8
    u = TL.ravel()
    assert np.count_nonzero(drift(u) != u) > 10
    R0, A, Ac = oracle.element(prm, xy, TL, 0.5, ied=ied)
    for p in range(54):
        st = np.where(np.arange(54) < p, drift(u), u)
        st[p] = u[p] + H_JAC
        Rp, _, _ = oracle.element(prm, xy, st.reshape(9, 6), 0.5, ied=ied, jac=False)
        col = ((Rp - R0).ravel() * H_JAC) / (H_JAC * H_JAC)
        assert np.array_equal(A[p], col), p
    Rp, _, _ = oracle.element(prm, xy, drift(u).reshape(9, 6), 0.5 + H_JAC, ied=ied, jac=False)
    assert np.array_equal(Ac, ((Rp - R0).ravel() * H_JAC) / (H_JAC * H_JAC))


def test_outlet_traction_closed_form(oracle):
    """U = 0, uniform S and P on a straight outlet edge x = const (edge 2):
    R(n, i) = -(T n + 0)_i ... = -sum_k n_k (T_ki) * integral(phi_n), with
    integral(phi) = L/6, L/6, 4L/6 for the corner, corner, mid nodes."""
    xy = NODES * np.array([0.5, 0.8]) + np.array([14.5, 0.3])  # edge 2 at x = 15
    S = np.array([[1.3, 0.2], [0.2, 0.9]])
    TL = np.zeros((9, 6))
    TL[:, 2] = 0.7
    TL[:, 3], TL[:, 4], TL[:, 5] = S[0, 0], S[0, 1], S[1, 1]
    EU, WiN0, beta = 0.5, 1.0, 0.59
    prm = oracle.params(BvN=beta, WiN0=WiN0)
    R, _, _ = oracle.element(prm, xy, TL, EU, ied=2, jac=False)
    T = (1 - beta) / (WiN0 * EU) * (S @ S - np.eye(2))
    Lg = 2 * 0.8
    w = {1: Lg / 6, 2: Lg / 6, 5: 4 * Lg / 6}
    expected = np.zeros((9, 6))
    for n, wn in w.items():
        expected[n, 0:2] = -np.array([1.0, 0.0]) @ T * wn  # -(n_k T_ki) with n = (1, 0)
    np.testing.assert_allclose(R, expected, atol=1e-8)
