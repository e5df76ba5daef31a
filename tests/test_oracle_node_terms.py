"""a8: the test-node terms of the oracle's element residual (momentumMainSQRT,
continuityBulkMain, stressesBulkMainSQRT and the LSME/LSIC/LSCE/DCS functions,
src/Residuals_mod.F90:90-110,161-234,295-490), pinned term by term through
laws of the weak form rather than by restating the formulas.

The oracle hands out the summands of TERM_RES for each test node at a Gauss
point (`oracle.node_terms`); they are tied to the product residual by
re-adding them in the oracle's order, which must give `oracle.element`'s
residual bit for bit.  Each summand is then checked with quantities pinned
elsewhere: the stabilization coefficients (`gauss_tau`, tests/
test_oracle_stabilization.py), the strong-form momentum, constitutive
residual and its upper-convected part (`point_strong_form`, tests/
test_oracle_strong_form.py), and polynomial fields whose values, gradients
and Laplacians are known in closed form, with an independent numpy Q2 basis.
The laws:

* isoparametric moments: the test space holds 1, x and y, so for any summand
  c(n) that is linear in grad(phi_n), sum_n c(n) and sum_n X_n c(n) (X_n the
  node coordinates) equal the term evaluated with grad v = 0 and grad v = e_d
  — on curved elements too.  That pins LSME (momentum, continuity, stress),
  LSIC, LSCE (stress), the DfN and DCS diffusion and the Galerkin parts;
* integration by parts at the interior (bubble) node, whose basis function
  vanishes on the element boundary: on affine elements and polynomial fields
  integrated exactly by the 3x3 rule, int (ReN Conv_i phi + P_tot(i,:).grad phi)
  = int Momentum_i phi with the strong-form momentum (the sign and every
  factor of the Galerkin stress term), and int DfN grad S.grad phi =
  -int DfN lap(S) phi;
* linearisation consistency of LSCE-momentum: with a linear, traceless test
  velocity v = L x, sum_{n,i} v_i(X_n) LSCE_i(n) = tau_LSCE (1-beta)/Wi
  UCD_v : Con_EQ, UCD_v being the velocity part of the strong form's
  upper-convected derivative at grad u = grad v;
* divergence of a test stress: sum_{n,s} sigma_s(X_n) LSME_s(n) =
  -tau_LSME Wi Momentum . div(sigma) for a linear symmetric test stress
  (pins the j /= k term and the sign), and its Wi scaling at a state whose
  momentum does not depend on Wi;
* closed-form states of the assembled element residual: rest with a linear
  pressure (only the pressure Galerkin term and LSME survive) and rest with a
  uniform stretched conformation (only Con_EQ (1 + tau_LSCE f_T/2) phi
  survives on the stress rows).

Eighteen planted mistakes in the oracle, one or two per summand (a dropped
factor ReN, Wi, (1-beta) or 1/2, a wrong sign or component, the LSCE test
gradient untransposed, the j /= k term dropped, ...), each fail at least one
law.  The GPU's node terms are bitwise equal to the oracle's (test_gpu_parity),
so these pins carry over to the product.
"""
import numpy as np
import pytest

from conftest import reference_nodes
from test_oracle_stabilization import basis, elements

RTOL = 1e-11
COMP = ((0, 0), (0, 1), (1, 1))


# ---------------------------------------------------------------------------
# polynomial fields with closed-form derivatives
# ---------------------------------------------------------------------------
class Field:
    """u (2), p, S (xx, xy, yy) as polynomials of total degree <= 2:
    f(x) = c + g.x + x.H.x / 2 per component, with gradient g + H x and
    Laplacian tr H."""

    def __init__(self, rng, u_deg=2, s_deg=1, p_deg=1, u_scale=1.0, s0=(1.3, 0.2, 0.8), traceless=False,
                 s_amp=0.15, p_amp=1.0):
        def poly(deg, c, amp):
            g = amp * rng.standard_normal(2) if deg >= 1 else np.zeros(2)
            H = amp * rng.standard_normal((2, 2)) if deg >= 2 else np.zeros((2, 2))
            return np.asarray(c, float), g, 0.5 * (H + H.T)

        self.u = [poly(u_deg, u_scale * rng.standard_normal(), u_scale) for _ in range(2)]
        if traceless:  # du/dx + dv/dy = 0 for linear u
            c, g, H = self.u[1]
            self.u[1] = (c, np.array([g[0], -self.u[0][1][0]]), H)
        self.p = poly(p_deg, p_amp * rng.standard_normal(), p_amp)
        self.S = [poly(s_deg, s0[s], s_amp) for s in range(3)]

    @staticmethod
    def ev(f, x):
        c, g, H = f
        return float(c + g @ x + 0.5 * x @ H @ x), g + H @ x, float(np.trace(H))

    def at(self, x):
        U = np.array([self.ev(f, x)[0] for f in self.u])
        GU = np.array([self.ev(f, x)[1] for f in self.u]).T  # GU[j][i] = du_i/dx_j
        lapU = np.array([self.ev(f, x)[2] for f in self.u])
        G2U = np.array([[self.u[i][2][j, j] for i in range(2)] for j in range(2)])  # G2U[j][i] = d2u_i/dx_j2
        P, dP, _ = self.ev(self.p, x)
        S = np.array([self.ev(f, x)[0] for f in self.S])
        dS = np.array([self.ev(f, x)[1] for f in self.S])
        lapS = np.array([self.ev(f, x)[2] for f in self.S])
        return dict(U=U, GU=GU, G2U=G2U, lapU=lapU, P=P, dP=dP, S=S, dS=dS, lapS=lapS)

    def nodal(self, xy):
        TL = np.zeros((9, 6))
        for n in range(9):
            f = self.at(xy[n])
            TL[n, 0:2] = f["U"]
            TL[n, 2] = f["P"]
            TL[n, 3:6] = f["S"]
        return TL


def gauss_point(xy, T, g):
    """Independent isoparametric Q2 basis at Gauss point g: phi (9), grad phi
    (2 x 9), the point x_g and the weight w_g |detJ|."""
    xi, eta = T["P2"][g]
    N, grad = basis(xy, xi, eta)

    def lagrange(t):  # values and derivatives of the 1D quadratics on -1, 0, 1
        return ({-1: t * (t - 1) / 2, 0: 1 - t * t, 1: t * (t + 1) / 2}, {-1: t - 0.5, 0: -2 * t, 1: t + 0.5})
    vx, dx = lagrange(xi)
    ve, de = lagrange(eta)
    ref = reference_nodes().astype(int)
    dNx = np.array([dx[a] * ve[b] for a, b in ref])
    dNe = np.array([vx[a] * de[b] for a, b in ref])
    detJ = (dNx @ xy[:, 0]) * (dNe @ xy[:, 1]) - (dNx @ xy[:, 1]) * (dNe @ xy[:, 0])
    return N, grad, N @ xy, T["W2"][g] * abs(detJ)


def strong(oracle, prm, f, EU):
    return oracle.point_strong_form(prm, f["U"], f["GU"], f["G2U"], f["P"], f["dP"], f["S"], f["dS"], EU)


def close(a, b, scale):
    return bool(np.all(np.abs(np.asarray(a) - np.asarray(b)) <= RTOL * scale))


def sym(s):
    return np.array([[s[0], s[1]], [s[1], s[2]]])


PARAM_SETS = [
    dict(ReN=0.7, BvN=0.59, VeN=0.25, DfN=1e-3, WiN0=1.0, model="L-PTT"),
    dict(ReN=2.0, BvN=0.3, VeN=0.1, DfN=4e-3, WiN0=2.0, model="GIESEKUS"),
]
EU = 0.6


def cases(rng, affine_only=False, **field_kw):
    """(parameters, element nodes, field, affine?) over the test elements: a
    rectangle and a parallelogram (affine) and a curved element."""
    els = elements()
    for pset in PARAM_SETS:
        for e, xy in enumerate(els):
            if affine_only and e == 2:
                continue
            yield pset, xy, Field(rng, **field_kw), e < 2


def gp_fields(oracle, xy, TL, g):
    """The interpolated fields at Gauss point g (setFlowUnknowns), as the
    strong form takes them."""
    f = oracle.gauss_fields(xy, TL, g)
    return dict(U=f["U"], GU=f["GU"], G2U=f["G2U"], P=f["P"], dP=f["dP"], S=f["S"], dS=f["dS"])


# ---------------------------------------------------------------------------
# laws
# ---------------------------------------------------------------------------
def law_recombination(oracle, T):
    """Re-adding the summands in the oracle's order, times WET, over the Gauss
    points gives the element residual bit for bit."""
    rng = np.random.default_rng(11)
    ok = True
    for pset, xy, fld, affine in cases(rng, s_deg=2):
        prm = oracle.params(**pset)
        TL = fld.nodal(xy)
        Res = np.zeros((9, 6))
        for g in range(9):
            t = oracle.node_terms(prm, xy, TL, EU, g)
            for n in range(9):
                m, c, s = t["mom"][n], t["cont"][n], t["str"][n]
                term = [m[i][0] + m[i][1] + m[i][2] + m[i][3] + m[i][4] for i in range(2)]
                term.append(c[0] + c[1])
                term += [s[k][0] + s[k][1] + s[k][2] + s[k][3] + s[k][4] for k in range(3)]
                Res[n] = Res[n] + np.array(term) * t["WET"]
        ok &= np.array_equal(Res, oracle.element(prm, xy, TL, EU, jac=False)[0])
    return ok


def law_galerkin_momentum_by_parts(oracle, T):
    """Bubble node, affine elements: int (ReN Conv_i phi + P_tot(i,:).grad phi)
    = int Momentum_i phi (strong form), and per Gauss point the convection
    summands add up to ReN (U.grad) u_i over the test nodes."""
    rng = np.random.default_rng(12)
    ok = True
    for pset, xy, fld, affine in cases(rng, affine_only=True):
        prm = oracle.params(**pset)
        TL = fld.nodal(xy)
        lhs = np.zeros(2)
        rhs = np.zeros(2)
        scale = np.zeros(2)
        for g in range(9):
            t = oracle.node_terms(prm, xy, TL, EU, g)
            N, grad, x, w = gauss_point(xy, T, g)
            f = gp_fields(oracle, xy, TL, g)
            M = strong(oracle, prm, f, EU)["Momentum"]
            gal = t["mom"][8, :, 0] + t["mom"][8, :, 1]
            lhs += w * gal
            rhs += w * M * N[8]
            scale += w * (np.abs(t["mom"][8, :, 0]) + np.abs(t["mom"][8, :, 1]))
            conv = pset["ReN"] * (f["GU"].T @ f["U"])  # ReN (U.grad) u
            ok &= close(t["mom"][:, :, 0].sum(0), conv, np.abs(t["mom"][:, :, 0]).sum() + 1e-300)
            ok &= close(t["mom"][:, :, 1].sum(0), 0.0, np.abs(t["mom"][:, :, 1]).sum())
        ok &= bool(np.all(np.abs(lhs - rhs) <= 1e-6 * scale))
    return ok


def moments(c, xy):
    """sum_n c(n), sum_n x_n c(n), sum_n y_n c(n) over the leading axis."""
    return c.sum(0), np.tensordot(xy[:, 0], c, 1), np.tensordot(xy[:, 1], c, 1)


def moment_scale(c, xy):
    a = np.abs(c)
    return a.sum() * (1.0 + np.abs(xy).max())


def law_lsme_momentum(oracle, T):
    """LSME momentum: sum_n = 0, sum_n X_n,d = tau_LSME ReN U_d Momentum_i."""
    rng = np.random.default_rng(13)
    ok = True
    for pset, xy, fld, affine in cases(rng):
        prm = oracle.params(**pset)
        TL = fld.nodal(xy)
        for g in (0, 4, 7):
            t = oracle.node_terms(prm, xy, TL, EU, g)
            _, _, x, _ = gauss_point(xy, T, g)
            f = gp_fields(oracle, xy, TL, g)
            M = strong(oracle, prm, f, EU)["Momentum"]
            tau = oracle.gauss_tau(prm, xy, TL, EU, g)["tlsme"]
            c = t["mom"][:, :, 2]
            m0, mx, my = moments(c, xy)
            sc = moment_scale(c, xy)
            ok &= close(m0, 0.0, sc)
            ok &= close(mx, tau * pset["ReN"] * f["U"][0] * M, sc)
            ok &= close(my, tau * pset["ReN"] * f["U"][1] * M, sc)
    return ok


def law_lsic(oracle, T):
    """LSIC: sum_n X_n,d LSIC_i(n) = tau_LSIC delta_id div u."""
    rng = np.random.default_rng(14)
    ok = True
    for pset, xy, fld, affine in cases(rng):
        prm = oracle.params(**pset)
        TL = fld.nodal(xy)
        for g in (1, 5, 8):
            t = oracle.node_terms(prm, xy, TL, EU, g)
            _, _, x, _ = gauss_point(xy, T, g)
            div = np.trace(gp_fields(oracle, xy, TL, g)["GU"])
            tau = oracle.gauss_tau(prm, xy, TL, EU, g)["tlsic"]
            c = t["mom"][:, :, 3]
            m0, mx, my = moments(c, xy)
            sc = moment_scale(c, xy)
            ok &= close(m0, 0.0, sc) and close(mx, [tau * div, 0.0], sc) and close(my, [0.0, tau * div], sc)
    return ok


def law_lsce_momentum(oracle, T):
    """LSCE momentum = tau_LSCE (1-beta)/Wi times the derivative of the
    constitutive residual's velocity part in the direction of the test
    function: for v = L x with tr L = 0, sum_{n,i} v_i(X_n) LSCE_i(n) =
    tau (1-beta)/Wi sum_ab UCD_v[a][b] Con_EQ[b][a]."""
    rng = np.random.default_rng(15)
    ok = True
    for pset, xy, fld, affine in cases(rng, s_deg=1):
        prm = oracle.params(**pset)
        TL = fld.nodal(xy)
        Wi = pset["WiN0"] * EU
        for g in (2, 4, 6):
            t = oracle.node_terms(prm, xy, TL, EU, g)
            _, _, x, _ = gauss_point(xy, T, g)
            f = gp_fields(oracle, xy, TL, g)
            CE = strong(oracle, prm, f, EU)["Con_EQ"]
            tau = oracle.gauss_tau(prm, xy, TL, EU, g)["tlsce"]
            c = t["mom"][:, :, 4]
            ok &= close(c.sum(0), 0.0, moment_scale(c, xy))
            for trial in range(2):
                L = rng.standard_normal((2, 2))
                L[1, 1] = -L[0, 0]
                v = xy @ L.T  # v_i(X_n)
                lhs = float(np.sum(v * c))
                # the strong form's UCD at rest (U = 0) with grad u = L: GU[j][i] = L[i][j]
                UCDv = strong(oracle, prm, dict(f, U=np.zeros(2), GU=L.T.copy()), EU)["UCD"]
                rhs = tau * (1.0 - pset["BvN"]) / Wi * float(np.sum(UCDv * CE.T))
                ok &= close(lhs, rhs, float(np.sum(np.abs(v * c))) * 10)
    return ok


def law_continuity(oracle, T):
    """Continuity row: the Galerkin summands add up to div u (and to div u x_d
    against X_n,d); LSME: sum_n X_n,d LSME(n) = tau_LSME Momentum_d."""
    rng = np.random.default_rng(16)
    ok = True
    for pset, xy, fld, affine in cases(rng):
        prm = oracle.params(**pset)
        TL = fld.nodal(xy)
        for g in (0, 3, 8):
            t = oracle.node_terms(prm, xy, TL, EU, g)
            _, _, x, _ = gauss_point(xy, T, g)
            f = gp_fields(oracle, xy, TL, g)
            div = np.trace(f["GU"])
            M = strong(oracle, prm, f, EU)["Momentum"]
            tau = oracle.gauss_tau(prm, xy, TL, EU, g)["tlsme"]
            c0 = t["cont"][:, 0]
            m0, mx, my = moments(c0, xy)
            sc = moment_scale(c0, xy)
            ok &= close(m0, div, sc) and close(mx, div * x[0], sc) and close(my, div * x[1], sc)
            c1 = t["cont"][:, 1]
            m0, mx, my = moments(c1, xy)
            sc = moment_scale(c1, xy)
            ok &= close(m0, 0.0, sc) and close(mx, tau * M[0], sc) and close(my, tau * M[1], sc)
    return ok


def law_stress_galerkin(oracle, T):
    """Stress rows: sum_n Con_EQ(j,k) phi_n = Con_EQ(j,k), and x_d Con_EQ(j,k)
    against X_n,d."""
    rng = np.random.default_rng(17)
    ok = True
    for pset, xy, fld, affine in cases(rng, s_deg=2):
        prm = oracle.params(**pset)
        TL = fld.nodal(xy)
        for g in (1, 4, 6):
            t = oracle.node_terms(prm, xy, TL, EU, g)
            _, _, x, _ = gauss_point(xy, T, g)
            CE = strong(oracle, prm, gp_fields(oracle, xy, TL, g), EU)["Con_EQ"]
            want = np.array([CE[j, k] for j, k in COMP])
            c = t["str"][:, :, 0]
            m0, mx, my = moments(c, xy)
            sc = moment_scale(c, xy)
            ok &= close(m0, want, sc) and close(mx, want * x[0], sc) and close(my, want * x[1], sc)
    return ok


def law_diffusion(oracle, T):
    """DfN term: sum_n X_n,d = DfN dS/dx_d; at the bubble node on affine
    elements int DfN grad S.grad phi = -int DfN lap(S) phi."""
    rng = np.random.default_rng(18)
    ok = True
    for pset, xy, fld, affine in cases(rng, s_deg=2):
        prm = oracle.params(**pset)
        TL = fld.nodal(xy)
        lhs = np.zeros(3)
        rhs = np.zeros(3)
        scale = np.zeros(3)
        for g in range(9):
            t = oracle.node_terms(prm, xy, TL, EU, g)
            N, _, x, w = gauss_point(xy, T, g)
            f = gp_fields(oracle, xy, TL, g)
            c = t["str"][:, :, 1]
            m0, mx, my = moments(c, xy)
            sc = moment_scale(c, xy)
            D = pset["DfN"]
            ok &= close(m0, 0.0, sc) and close(mx, D * f["dS"][:, 0], sc) and close(my, D * f["dS"][:, 1], sc)
            lhs += w * c[8]
            rhs += -w * D * fld.at(x)["lapS"] * N[8]
            scale += w * np.abs(c[8])
        if affine:
            ok &= bool(np.all(np.abs(lhs - rhs) <= 1e-6 * scale))
    return ok


def law_lsme_stress(oracle, T):
    """sum_{n,s} sigma_s(X_n) LSME_s(n) = -tau_LSME Wi Momentum . div(sigma)
    for a linear symmetric test stress sigma."""
    rng = np.random.default_rng(19)
    ok = True
    for pset, xy, fld, affine in cases(rng, s_deg=2, p_deg=2):
        prm = oracle.params(**pset)
        TL = fld.nodal(xy)
        Wi = pset["WiN0"] * EU
        for g in (0, 2, 7):
            t = oracle.node_terms(prm, xy, TL, EU, g)
            _, _, x, _ = gauss_point(xy, T, g)
            M = strong(oracle, prm, gp_fields(oracle, xy, TL, g), EU)["Momentum"]
            tau = oracle.gauss_tau(prm, xy, TL, EU, g)["tlsme"]
            c = t["str"][:, :, 2]
            ok &= close(c.sum(0), 0.0, moment_scale(c, xy))
            for trial in range(3):
                c0 = rng.standard_normal(3)
                D = rng.standard_normal((3, 2))  # sigma_s = c0_s + D_s . x
                sig = c0[None, :] + xy @ D.T
                div = np.array([D[0, 0] + D[1, 1], D[1, 0] + D[2, 1]])
                lhs = float(np.sum(sig * c))
                ok &= close(lhs, -tau * Wi * float(M @ div), float(np.sum(np.abs(sig * c))) * 10)
    return ok


def law_lsme_stress_wi_scaling(oracle, T):
    """At rest with S = I the momentum residual is grad P whatever Wi, so
    LSME_s / tau_LSME scales as Wi."""
    rng = np.random.default_rng(20)
    ok = True
    for xy in elements():
        fld = Field(rng, u_deg=-1, s_deg=0, s0=(1.0, 0.0, 1.0), s_amp=0.0)
        fld.u = [(np.array(0.0), np.zeros(2), np.zeros((2, 2)))] * 2
        TL = fld.nodal(xy)
        ratios = []
        for Wi0 in (0.5, 1.5):
            prm = oracle.params(ReN=0.0, BvN=0.59, VeN=0.0, DfN=1e-3, WiN0=Wi0)
            g = 5
            t = oracle.node_terms(prm, xy, TL, 1.0, g)
            tau = oracle.gauss_tau(prm, xy, TL, 1.0, g)["tlsme"]
            ratios.append(t["str"][:, :, 2] / tau)
        a, b = ratios
        ok &= close(b, 3.0 * a, np.abs(b).max())
        ok &= bool(np.abs(a).max() > 0)
    return ok


def law_lsce_stress(oracle, T):
    """LSCE stress: sum_n = tau_LSCE f_T/2 Con_EQ(j,k); against X_n,d:
    tau_LSCE (Wi U_d + f_T x_d / 2) Con_EQ(j,k), f_T the model function
    (1 + eps (tr C - 2) for linear PTT, 1 - 2 alpha for Giesekus)."""
    rng = np.random.default_rng(21)
    ok = True
    for pset, xy, fld, affine in cases(rng, s_deg=2):
        prm = oracle.params(**pset)
        TL = fld.nodal(xy)
        Wi = pset["WiN0"] * EU
        for g in (3, 4, 5):
            t = oracle.node_terms(prm, xy, TL, EU, g)
            _, _, x, _ = gauss_point(xy, T, g)
            f = gp_fields(oracle, xy, TL, g)
            sf = strong(oracle, prm, f, EU)
            CE = np.array([sf["Con_EQ"][j, k] for j, k in COMP])
            S = sym(f["S"])
            if pset["model"] == "L-PTT":
                fT = 1.0 + pset["VeN"] * (np.trace(S @ S) - 2.0)
            else:
                fT = 1.0 - 2.0 * pset["VeN"]
            tau = oracle.gauss_tau(prm, xy, TL, EU, g)["tlsce"]
            c = t["str"][:, :, 3]
            m0, mx, my = moments(c, xy)
            sc = moment_scale(c, xy)
            ok &= close(m0, tau * fT / 2 * CE, sc)
            ok &= close(mx, tau * (Wi * f["U"][0] + fT * x[0] / 2) * CE, sc)
            ok &= close(my, tau * (Wi * f["U"][1] + fT * x[1] / 2) * CE, sc)
    return ok


def law_discontinuity_capturing(oracle, T):
    """DfN = 0: the DCS summand is tdcs |Con_EQ| grad S(j,k).grad phi, with
    |Con_EQ| = sqrt(Con_EQ:Con_EQ/2) from the strong form: sum_n X_n,d DCS(n)
    = tdcs |Con_EQ| dS(j,k)/dx_d; with DfN > 1e-12 it is zero."""
    rng = np.random.default_rng(22)
    ok = True
    for pset, xy, fld, affine in cases(rng, s_deg=2):
        for DfN in (0.0, pset["DfN"]):
            prm = oracle.params(**dict(pset, DfN=DfN))
            TL = fld.nodal(xy)
            for g in (0, 6):
                t = oracle.node_terms(prm, xy, TL, EU, g)
                f = gp_fields(oracle, xy, TL, g)
                CE = strong(oracle, prm, f, EU)["Con_EQ"]
                mag = np.sqrt(0.5 * np.sum(CE * CE))
                tdcs = oracle.gauss_tau(prm, xy, TL, EU, g)["tdcs"]
                c = t["str"][:, :, 4]
                if DfN > 1e-12:
                    ok &= bool(np.all(c == 0.0))
                    continue
                m0, mx, my = moments(c, xy)
                sc = moment_scale(c, xy)
                ok &= tdcs > 0
                ok &= close(m0, 0.0, sc)
                ok &= close(mx, tdcs * mag * f["dS"][:, 0], sc) and close(my, tdcs * mag * f["dS"][:, 1], sc)
    return ok


def _element_rows(oracle, prm, xy, TL, EU_):
    return oracle.element(prm, xy, TL, EU_, jac=False)[0]


def law_pressure_state(oracle, T):
    """U = 0, S = I, P = p0 + a x + b y: the assembled element residual is
    -int P d_i phi on the momentum rows, int tau_LSME grad P.grad phi on the
    continuity row and -int tau_LSME Wi (d_k phi dP/dx_j + [j /= k] d_j phi
    dP/dx_k) on the stress rows."""
    rng = np.random.default_rng(23)
    ok = True
    for xy in elements():
        for pset in PARAM_SETS:
            prm = oracle.params(**pset)
            Wi = pset["WiN0"] * EU
            p0, a, b = rng.standard_normal(3)
            TL = np.zeros((9, 6))
            TL[:, 2] = p0 + a * xy[:, 0] + b * xy[:, 1]
            TL[:, 3] = TL[:, 5] = 1.0
            R = _element_rows(oracle, prm, xy, TL, EU)
            want = np.zeros((9, 6))
            dP = np.array([a, b])
            for g in range(9):
                N, grad, x, w = gauss_point(xy, T, g)
                tau = oracle.gauss_tau(prm, xy, TL, EU, g)["tlsme"]
                P = p0 + a * x[0] + b * x[1]
                want[:, 0] += w * (-P * grad[0])
                want[:, 1] += w * (-P * grad[1])
                want[:, 2] += w * tau * (dP @ grad)
                for s, (j, k) in enumerate(COMP):
                    v = grad[k] * dP[j] + (grad[j] * dP[k] if j != k else 0.0)
                    want[:, 3 + s] += -w * tau * Wi * v
            ok &= close(R, want, np.abs(want).max())
    return ok


def law_stretched_rest_state(oracle, T):
    """U = 0, uniform S = S0 (not I), uniform P: stress rows of the assembled
    residual = int Con_EQ(j,k) (1 + tau_LSCE f_T / 2) phi, continuity rows 0."""
    ok = True
    for xy in elements():
        for pset in PARAM_SETS:
            for S0 in ((1.4, 0.3, 0.8), (0.9, -0.2, 1.6)):
                prm = oracle.params(**pset)
                TL = np.zeros((9, 6))
                TL[:, 2] = 0.4
                TL[:, 3:6] = S0
                R = _element_rows(oracle, prm, xy, TL, EU)
                want = np.zeros((9, 3))
                for g in range(9):
                    N, _, x, w = gauss_point(xy, T, g)
                    f = dict(U=np.zeros(2), GU=np.zeros((2, 2)), G2U=np.zeros((2, 2)), P=0.4, dP=np.zeros(2),
                             S=np.array(S0), dS=np.zeros((3, 2)))
                    sf = strong(oracle, prm, f, EU)
                    CE = np.array([sf["Con_EQ"][j, k] for j, k in COMP])
                    tau = oracle.gauss_tau(prm, xy, TL, EU, g)["tlsce"]
                    want += w * np.outer(N, CE * (1.0 + tau * sf["f_T"] / 2))
                ok &= close(R[:, 3:], want, np.abs(want).max())
                ok &= close(R[:, 2], 0.0, np.abs(want).max())
    return ok


LAWS = [law_recombination, law_galerkin_momentum_by_parts, law_lsme_momentum, law_lsic, law_lsce_momentum,
        law_continuity, law_stress_galerkin, law_diffusion, law_lsme_stress, law_lsme_stress_wi_scaling,
        law_lsce_stress, law_discontinuity_capturing, law_pressure_state, law_stretched_rest_state]

N_PLANTS = 18


@pytest.fixture(scope="module")
def tables(oracle):
    return oracle.tables()


@pytest.mark.parametrize("law", LAWS, ids=lambda f: f.__name__)
def test_node_term_law(oracle, tables, law):
    assert law(oracle, tables)


@pytest.mark.parametrize("plant", range(1, N_PLANTS + 1))
def test_planted_node_term_mistake_breaks_a_law(oracle, tables, plant):
    oracle.term_plant(plant)
    try:
        broken = [law.__name__ for law in LAWS if not law(oracle, tables)]
    finally:
        oracle.term_plant(0)
    assert broken, f"plant {plant} passes every law"


def test_node_terms_argument_checks(oracle):
    xy = elements()[0]
    with pytest.raises(ValueError):
        oracle.node_terms(oracle.params(), xy, np.zeros((9, 6)), 1.0, 9)
    with pytest.raises(ValueError):
        oracle.term_plant(N_PLANTS + 1)
