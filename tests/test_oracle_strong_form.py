"""a5/a6: the Gauss-point strong form of the oracle (strongFormConEQ_SQRT,
strongFormMomentum, strongFormContinuity, setStressTensorSQRT), pinned
without restating it:

* the square-root identity: with C = S^2, S R_S + R_S S equals the textbook
  conformation residual R_C = Wi (u.grad C - L C - C L^T) + f2 C - f1 I + f3 C^2
  (L the traceless velocity gradient, L_ij = du_i/dx_j) for Oldroyd-B, linear
  PTT and Giesekus — this holds only if the omega tensor makes
  S.GU_tr + W.S symmetric, so a wrong sign, index or transpose in UCD, W,
  S^-1 or the model functions fails it;
* the exact Oldroyd-B Poiseuille solution has zero momentum, continuity and
  constitutive residuals;
* the momentum residual against finite differences of T = (1-beta)/Wi (S^2 - I)
  evaluated from analytic fields.
"""
import numpy as np
import pytest
from scipy.linalg import sqrtm


def random_point(rng, spd_scale=0.3):
    U = rng.standard_normal(2)
    GU = rng.standard_normal((2, 2))  # GU[j][i] = dU_i/dx_j
    G2U = rng.standard_normal((2, 2))
    M = rng.standard_normal((2, 2)) * spd_scale
    S = np.eye(2) + 0.5 * (M + M.T)
    S = S @ S.T * 0.5 + np.eye(2) * 0.5  # symmetric positive definite
    dS = rng.standard_normal((3, 2)) * 0.3  # (xx, xy, yy) x (d/dx, d/dy)
    return U, GU, G2U, rng.standard_normal(), rng.standard_normal(2), S, dS


def sym(v):
    return np.array([[v[0], v[1]], [v[1], v[2]]])


def conformation_residual(U, GU, S, dS, Wi, model, VeN):
    """Textbook residual of the conformation equation for C = S^2."""
    L = GU.T.copy()  # L_ij = du_i/dx_j
    L -= np.trace(L) / 2.0 * np.eye(2)  # traceless velocity gradient
    C = S @ S
    dC = [sym(dS[:, d]) @ S + S @ sym(dS[:, d]) for d in range(2)]
    adv = U[0] * dC[0] + U[1] * dC[1]
    ucd = adv - L @ C - C @ L.T
    I = np.eye(2)
    if model == "L-PTT":
        f1 = f2 = 1.0 + VeN * (np.trace(C) - 2.0)
        f3 = 0.0
    else:
        f1, f2, f3 = 1.0 - VeN, 1.0 - 2.0 * VeN, VeN
    return Wi * ucd + f2 * C - f1 * I + f3 * C @ C


@pytest.mark.parametrize("model,VeN", [("L-PTT", 0.0), ("L-PTT", 0.25), ("GIESEKUS", 0.1), ("GIESEKUS", 0.3)])
def test_sqrt_conformation_identity(oracle, model, VeN):
    rng = np.random.default_rng(10)
    for trial in range(50):
        U, GU, G2U, P, dP, S, dS = random_point(rng)
        EU, WiN0 = 0.3 + rng.random(), 1.0 + rng.random()
        prm = oracle.params(VeN=VeN, model=model, WiN0=WiN0)
        out = oracle.point_strong_form(prm, U, GU, G2U, P, dP, [S[0, 0], S[0, 1], S[1, 1]], dS, EU)
        RS = out["Con_EQ"]
        Wi = WiN0 * EU
        RC = conformation_residual(U, GU, S, dS, Wi, model, VeN)
        lhs = S @ RS + RS @ S
        scale = np.abs(RC).max() + Wi * np.abs(S).max() ** 2 * (np.abs(GU).max() + np.abs(dS).max())
        np.testing.assert_allclose(lhs, RC, rtol=0, atol=1e-13 * scale * 10)
        # R_S itself is symmetric up to rounding
        assert abs(RS[0, 1] - RS[1, 0]) <= 1e-13 * scale * 10
        # C = S^2 and the stress T = (1-beta)/Wi (C - I), T_zz = -(1-beta)/Wi
        np.testing.assert_allclose(out["C"], S @ S, atol=1e-14)
        a = (1 - 0.59) / Wi
        np.testing.assert_allclose(out["T_ve"][:2, :2], a * (S @ S - np.eye(2)), atol=1e-13)
        assert out["T_ve"][2, 2] == pytest.approx(-a, rel=1e-15)


def test_rest_state_zero_residual(oracle):
    prm = oracle.params()
    out = oracle.point_strong_form(prm, [0, 0], np.zeros((2, 2)), np.zeros((2, 2)), 0.0, [0, 0], [1, 0, 1],
                                   np.zeros((3, 2)), 0.5)
    assert np.abs(out["Con_EQ"]).max() == 0.0
    assert np.abs(out["Momentum"]).max() == 0.0


def poiseuille(y, Wi, H=2.0, Ubar=1.0):
    """Oldroyd-B channel flow u = 1.5 Ubar (1 - y^2/H^2) with C from the
    steady constitutive law: C = [[1 + 2(Wi g)^2, Wi g], [Wi g, 1]], g = du/dy."""
    u = 1.5 * Ubar * (1 - y * y / (H * H))
    g = -3.0 * Ubar * y / (H * H)
    upp = -3.0 * Ubar / (H * H)
    C = np.array([[1 + 2 * (Wi * g) ** 2, Wi * g], [Wi * g, 1.0]])
    return u, g, upp, C


@pytest.mark.parametrize("Wi", [0.1, 0.5, 0.9])
def test_oldroyd_b_poiseuille_is_exact(oracle, Wi):
    beta = 0.59
    prm = oracle.params(BvN=beta, WiN0=1.0)
    for y in (-1.7, -0.4, 0.3, 1.2):
        u, g, upp, C = poiseuille(y, Wi)
        S = np.real(sqrtm(C))
        h = 1e-5
        dSdy = (np.real(sqrtm(poiseuille(y + h, Wi)[3])) - np.real(sqrtm(poiseuille(y - h, Wi)[3]))) / (2 * h)
        # momentum balance: dP/dx = beta u'' + d(tau_xy)/dy = u''
        x = 0.37
        GU = np.array([[0.0, 0.0], [g, 0.0]])  # dUx/dy = g
        G2U = np.array([[0.0, 0.0], [upp, 0.0]])
        dS = np.array([[0.0, dSdy[0, 0]], [0.0, dSdy[0, 1]], [0.0, dSdy[1, 1]]])
        out = oracle.point_strong_form(prm, [u, 0.0], GU, G2U, upp * x, [upp, 0.0], [S[0, 0], S[0, 1], S[1, 1]], dS, Wi)
        assert np.abs(out["Con_EQ"]).max() < 1e-12
        assert np.abs(out["Momentum"]).max() < 5e-8  # central-difference error of dS/dy
        assert out["Continuity"] == 0.0


def test_momentum_against_finite_differences(oracle):
    """Momentum = Re u.grad u + grad P - div T - beta lap u at a point, with the
    derivatives of T = (1-beta)/Wi (S^2 - I) taken by central differences of
    analytic fields."""
    rng = np.random.default_rng(11)
    beta, Re, WiN0, EU = 0.59, 0.7, 1.3, 0.6
    Wi = WiN0 * EU
    prm = oracle.params(ReN=Re, BvN=beta, WiN0=WiN0)
    k = rng.uniform(0.5, 1.5, (8, 2))
    ph = rng.uniform(0, 6.28, 8)

    def fields(x, y):
        s = np.sin(k[:, 0] * x + k[:, 1] * y + ph)
        u = np.array([s[0] + 0.3 * s[1], s[2] - 0.2 * s[3]])
        P = s[4]
        Sm = np.array([[1.2 + 0.2 * s[5], 0.1 * s[6]], [0.1 * s[6], 0.9 + 0.2 * s[7]]])
        return u, P, Sm

    x0, y0, h = 0.3, -0.2, 1e-4

    def d(fun, axis):
        e = np.array([h, 0.0]) if axis == 0 else np.array([0.0, h])
        return (fun(x0 + e[0], y0 + e[1]) - fun(x0 - e[0], y0 - e[1])) / (2 * h)

    def d2(fun, axis):
        e = np.array([h, 0.0]) if axis == 0 else np.array([0.0, h])
        return (fun(x0 + e[0], y0 + e[1]) - 2 * fun(x0, y0) + fun(x0 - e[0], y0 - e[1])) / (h * h)

    u, P, S = fields(x0, y0)
    GU = np.stack([d(lambda x, y: fields(x, y)[0], 0), d(lambda x, y: fields(x, y)[0], 1)])
    G2U = np.stack([d2(lambda x, y: fields(x, y)[0], 0), d2(lambda x, y: fields(x, y)[0], 1)])
    dP = np.array([d(lambda x, y: np.array(fields(x, y)[1]), a) for a in (0, 1)])
    dSm = [d(lambda x, y: fields(x, y)[2], a) for a in (0, 1)]
    dS = np.array([[dSm[0][0, 0], dSm[1][0, 0]], [dSm[0][0, 1], dSm[1][0, 1]], [dSm[0][1, 1], dSm[1][1, 1]]])
    out = oracle.point_strong_form(prm, u, GU, G2U, P, dP, [S[0, 0], S[0, 1], S[1, 1]], dS, EU)

    def T(x, y):
        Sm = fields(x, y)[2]
        return (1 - beta) / Wi * (Sm @ Sm - np.eye(2))

    dT = [d(T, a) for a in (0, 1)]
    divT = np.array([dT[0][0, 0] + dT[1][1, 0], dT[0][0, 1] + dT[1][1, 1]])  # (div T)_i = dT_ji/dx_j
    conv = GU.T @ u  # (u.grad) u
    lap = G2U.sum(axis=0)
    expected = Re * conv + dP - divT - beta * lap
    np.testing.assert_allclose(out["Momentum"], expected, atol=1e-6)
    assert out["Continuity"] == pytest.approx(GU[0, 0] + GU[1, 1], rel=1e-14)
    # P_tot = T + beta (L_tr + L_tr^T) - P I
    L = GU.T - np.trace(GU) / 2 * np.eye(2)
    Ptot = (1 - beta) / Wi * (S @ S - np.eye(2)) + beta * (L + L.T) - P * np.eye(2)
    np.testing.assert_allclose(out["P_tot"][:2, :2], Ptot, atol=1e-12)
