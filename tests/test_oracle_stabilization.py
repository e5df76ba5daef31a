"""a7: the stabilization coefficients of the oracle (setStabilizationTerms,
Element_Calculations_mod.F90:701-773), pinned through laws that isolate one
term at a time rather than by restating the formulas:

* geometry: Helem**2 = 1 / sum_n |grad phi_n|^2 and tsugn = sum_n |u . grad phi_n|
  against an independent numpy Q2 isoparametric basis, on affine and curved
  elements at every Gauss point;
* tau_LSIC * tau_LSME = Helem**2 for any state;
* viscous limit: BvN = 1 makes the polymer stress vanish, so tau_LSME =
  Helem**2 exactly at ReN = 0 whatever the flow;
* polymer limit: BvN = 0, ReN = 0 gives tau_LSME = Helem**2 sqrt(1 + G:G/2) /
  (a sqrt(1 + |C - I|^2 / 2)) with a = (1-beta)/Wi and G the rate of strain of
  the traceless velocity gradient; C = S^2 is 3x3 with C33 = 0 in the planar
  problem (S33 is never set), so at rest |C - I|^2 = 1 and tau_LSME =
  Helem**2 / sqrt(BvN^2 + 3 a^2 / 2) — a quirk of the source that an
  independent derivation would miss, kept because the reference has it;
* inertia: tau_LSME^-2(ReN) - tau_LSME^-2(0) = ReN^2 (1/dt^2 + tsugn^2);
* constitutive time scale: tau_LSCE^-2 - tau_LSCE^-2 at rest-like terms grows as
  Wi^2 (1/dt^2 + tsugn^2 + G:G/2) over several Wi;
* model functions at U = 0: tau_LSCE^-2 = f2^2 + (f3 sqrt(C:C/2))^2 +
  (DfN/Helem**2)^2 for linear PTT (f2 = 1 + eps (tr C - 2), f3 = 0) and
  Giesekus (f2 = 1 - 2 alpha, f3 = alpha);
* DfN enters as (DfN/Helem**2)^2: differences over two DfN;
* discontinuity capturing (DfN = 0): tdcs = Helem**2 / sqrt(S:S/2) and it adds
  Con_EQ:Con_EQ / S:S to tau_LSCE^-2, with Con_EQ from the strong form at the
  same point.

The definitions themselves exist only in the reference's source (it has no
tests or published values for them); fourteen planted mistakes in the oracle,
one per term (a dropped square root, factor 1/2, abs() or Weissenberg number,
a swapped model function, ...), each fail at least one law.
"""
import numpy as np
import pytest

from conftest import reference_nodes

RTOL = 1e-11


def basis(xy, xi, eta):
    """Q2 values and physical gradients at (xi, eta), from the 1D quadratic
    Lagrange polynomials on the nodes -1, 0, 1."""
    def l(t):
        return {-1: (t * (t - 1) / 2, t - 0.5), 0: (1 - t * t, -2 * t), 1: (t * (t + 1) / 2, t + 0.5)}

    ref = reference_nodes()
    lx, le = l(xi), l(eta)
    N = np.array([lx[int(a)][0] * le[int(b)][0] for a, b in ref])
    dNx = np.array([lx[int(a)][1] * le[int(b)][0] for a, b in ref])
    dNe = np.array([lx[int(a)][0] * le[int(b)][1] for a, b in ref])
    J = np.array([[dNx @ xy[:, 0], dNx @ xy[:, 1]], [dNe @ xy[:, 0], dNe @ xy[:, 1]]])  # J[a][b] = dx_b/dxi_a
    return N, np.linalg.solve(J, np.vstack([dNx, dNe]))


def elements():
    ref = reference_nodes()
    rect = np.column_stack([0.2 + (ref[:, 0] + 1) * 0.15, -0.1 + (ref[:, 1] + 1) * 0.06])
    th = 0.7
    A = np.array([[np.cos(th), -np.sin(th)], [np.sin(th), np.cos(th)]]) @ np.array([[0.11, 0.03], [0.0, 0.07]])
    para = ref @ A.T + np.array([1.3, 0.4])
    curved = para.copy()
    curved[4:8] += np.array([[0.004, -0.006], [0.007, 0.002], [-0.003, 0.005], [-0.006, -0.004]])
    curved[0] += np.array([0.01, -0.005])
    curved[8] += np.array([0.003, 0.002])
    return [rect, para, curved]


def nodal(xy, L=np.zeros((2, 2)), u0=(0.0, 0.0), S0=(1.0, 0.0, 1.0), dS=np.zeros((3, 2)), P=(0.3, 0.1, -0.2)):
    """TL[9][6] of linear fields: u = u0 + L x (L_ij = du_i/dx_j), S = S0 + dS x,
    P linear; Q2 reproduces them exactly."""
    TL = np.zeros((9, 6))
    for n in range(9):
        x = xy[n]
        TL[n, 0:2] = np.asarray(u0) + L @ x
        TL[n, 2] = P[0] + P[1] * x[0] + P[2] * x[1]
        TL[n, 3:6] = np.asarray(S0) + dS @ x
    return TL


def half_GG(L):
    """G:G/2 with G the symmetric part (times 2) of the traceless gradient."""
    D = L - np.trace(L) / 2 * np.eye(2)
    G = D + D.T
    return 0.5 * np.sum(G * G)


def geometry_at(xy, TL, g, T):
    N, grad = basis(xy, *T["P2"][g])
    U = N @ TL[:, 0:2]
    return 1.0 / np.sum(grad * grad), np.sum(np.abs(U @ grad))


def close(a, b, scale=None):
    scale = max(abs(a), abs(b)) if scale is None else scale
    return abs(a - b) <= RTOL * scale


def law_geometry(oracle, T):
    rng = np.random.default_rng(3)
    prm = oracle.params()
    ok = True
    for xy in elements():
        for trial in range(2):
            TL = nodal(xy, L=rng.standard_normal((2, 2)), u0=rng.standard_normal(2))
            for g in range(9):
                out = oracle.gauss_tau(prm, xy, TL, 1.0, g)
                H2, ts = geometry_at(xy, TL, g, T)
                ok &= close(out["H2"], H2) and close(out["tsugn"], ts)
    return ok


def law_lsic(oracle, T):
    rng = np.random.default_rng(4)
    ok = True
    for xy in elements():
        prm = oracle.params(ReN=0.7, BvN=0.4, VeN=0.2, WiN0=1.3)
        M = rng.standard_normal((2, 2)) * 0.2
        S = np.eye(2) + 0.5 * (M + M.T)
        TL = nodal(xy, L=rng.standard_normal((2, 2)), u0=rng.standard_normal(2), S0=(S[0, 0], S[0, 1], S[1, 1]))
        for g in (0, 4, 8):
            o = oracle.gauss_tau(prm, xy, TL, 0.8, g)
            ok &= close(o["tlsic"] * o["tlsme"], o["H2"])
    return ok


def law_viscous_limit(oracle, T):
    """BvN = 1: a = 0, T_ve = 0, Ha = 0 -> tau_LSME = Helem**2 at ReN = 0."""
    rng = np.random.default_rng(5)
    prm = oracle.params(ReN=0.0, BvN=1.0, WiN0=0.9)
    ok = True
    for xy in elements():
        TL = nodal(xy, L=rng.standard_normal((2, 2)), u0=rng.standard_normal(2), S0=(1.4, 0.3, 0.8))
        for g in (1, 5):
            o = oracle.gauss_tau(prm, xy, TL, 1.0, g)
            ok &= close(o["tlsme"], geometry_at(xy, TL, g, T)[0])
    return ok


def law_polymer_limit(oracle, T):
    """BvN = 0, ReN = 0: tau_LSME = H2 sqrt(1 + G:G/2) / (a sqrt(1 + |C-I|^2/2)),
    C = S^2 with C33 = 0; at rest this is H2 / sqrt(BvN^2 + 3a^2/2)."""
    ok = True
    cases = [(np.zeros((2, 2)), (1.0, 0.0, 1.0)),          # rest
             (np.array([[0.0, 2.5], [0.0, 0.0]]), (1.0, 0.0, 1.0)),   # shear
             (np.array([[1.7, 0.0], [0.0, -1.7]]), (1.3, 0.2, 0.9)),  # extension, stressed
             (np.array([[0.4, -0.9], [1.2, 0.3]]), (0.8, -0.25, 1.6))]
    for WiN0, EU in ((0.7, 1.0), (2.0, 0.6)):
        a = 1.0 / (WiN0 * EU)
        prm = oracle.params(ReN=0.0, BvN=0.0, WiN0=WiN0)
        for xy in elements()[1:]:
            for L, S0 in cases:
                TL = nodal(xy, L=L, S0=S0)
                S = np.array([[S0[0], S0[1]], [S0[1], S0[2]]])
                CmI = np.zeros((3, 3))
                CmI[:2, :2] = S @ S - np.eye(2)
                CmI[2, 2] = -1.0
                for g in (2, 6):
                    H2 = geometry_at(xy, TL, g, T)[0]
                    want = H2 * np.sqrt(1 + half_GG(L)) / (a * np.sqrt(1 + 0.5 * np.sum(CmI * CmI)))
                    ok &= close(oracle.gauss_tau(prm, xy, TL, EU, g)["tlsme"], want)
    # at rest with both viscosities
    for BvN in (0.3, 0.59):
        prm = oracle.params(ReN=0.0, BvN=BvN, WiN0=1.1)
        xy = elements()[0]
        TL = nodal(xy)
        H2 = geometry_at(xy, TL, 3, T)[0]
        a = (1 - BvN) / 1.1
        ok &= close(oracle.gauss_tau(prm, xy, TL, 1.0, 3)["tlsme"], H2 / np.sqrt(BvN ** 2 + 1.5 * a * a))
    return ok


def law_inertia(oracle, T):
    """tau_LSME^-2(ReN) - tau_LSME^-2(0) = ReN^2 (1/dt^2 + tsugn^2)."""
    ok = True
    L = np.array([[0.3, 1.1], [-0.4, -0.2]])
    for xy in elements():
        TL = nodal(xy, L=L, u0=(4.0, -2.5))
        for dt in (0.02, None):
            for ReN in (0.5, 3.0):
                p1 = oracle.params(ReN=ReN, BvN=0.59, WiN0=1.0)
                p0 = oracle.params(ReN=0.0, BvN=0.59, WiN0=1.0)
                if dt is not None:
                    p1.dt = p0.dt = dt
                g = 7
                t1 = oracle.gauss_tau(p1, xy, TL, 1.0, g)["tlsme"] ** -2
                t0 = oracle.gauss_tau(p0, xy, TL, 1.0, g)["tlsme"] ** -2
                ts = geometry_at(xy, TL, g, T)[1]
                want = ReN ** 2 * ((1.0 / dt ** 2 if dt else 0.0) + ts * ts)
                ok &= close(t1 - t0, want, scale=t1)
    return ok


def law_constitutive_time_scale(oracle, T):
    """tau_LSCE^-2 = Wi^2 (1/dt^2 + tsugn^2 + G:G/2) + f2^2 + (DfN/H2)^2 for
    Oldroyd-B (f2 = 1), DfN > 1e-12; checked as the Wi^2 slope over three Wi
    and the intercept."""
    ok = True
    for L, u0 in ((np.array([[0.0, 2.0], [0.0, 0.0]]), (0.5, 0.2)), (np.array([[1.5, 0.0], [0.0, -1.5]]), (0.0, 0.0)),
                  (np.array([[0.3, -0.6], [0.9, 0.1]]), (-1.0, 0.7))):
        for dt in (0.05, None):
            xy = elements()[2]
            TL = nodal(xy, L=L, u0=u0)
            g = 4
            H2, ts = geometry_at(xy, TL, g, T)
            DfN = 1e-3
            vals = []
            for Wi in (0.5, 1.0, 2.5):
                prm = oracle.params(ReN=0.0, BvN=0.59, VeN=0.0, DfN=DfN, WiN0=Wi)
                if dt is not None:
                    prm.dt = dt
                vals.append((Wi, oracle.gauss_tau(prm, xy, TL, 1.0, g)["tlsce"] ** -2))
            rate = (1.0 / dt ** 2 if dt else 0.0) + ts * ts + half_GG(L)
            for Wi, v in vals:
                ok &= close(v, Wi * Wi * rate + 1.0 + (DfN / H2) ** 2)
    return ok


def law_model_functions(oracle, T):
    """U = 0: tau_LSCE^-2 = f2^2 + (f3 sqrt(C:C/2))^2 + (DfN/H2)^2."""
    ok = True
    xy = elements()[1]
    for S0 in ((1.0, 0.0, 1.0), (1.3, 0.25, 0.85), (0.7, -0.1, 1.9)):
        S = np.array([[S0[0], S0[1]], [S0[1], S0[2]]])
        C = S @ S
        TL = nodal(xy, S0=S0)
        H2 = geometry_at(xy, TL, 0, T)[0]
        DfN = 2e-3
        for model, VeN in (("L-PTT", 0.0), ("L-PTT", 0.25), ("GIESEKUS", 0.1), ("GIESEKUS", 0.3)):
            prm = oracle.params(ReN=0.0, BvN=0.59, VeN=VeN, DfN=DfN, WiN0=1.2, model=model)
            if model == "L-PTT":
                f2, f3 = 1.0 + VeN * (np.trace(C) - 2.0), 0.0
            else:
                f2, f3 = 1.0 - 2.0 * VeN, VeN
            want = f2 ** 2 + f3 ** 2 * 0.5 * np.sum(C * C) + (DfN / H2) ** 2
            ok &= close(oracle.gauss_tau(prm, xy, TL, 1.0, 0)["tlsce"] ** -2, want)
    return ok


def law_diffusivity(oracle, T):
    """tau_LSCE^-2(DfN1) - tau_LSCE^-2(DfN2) = (DfN1^2 - DfN2^2) / H2^2."""
    ok = True
    for xy in elements():
        TL = nodal(xy, L=np.array([[0.2, 0.5], [0.1, -0.2]]), u0=(0.3, 0.1), S0=(1.2, 0.1, 0.9))
        g = 8
        H2 = geometry_at(xy, TL, g, T)[0]
        D1, D2 = 0.02 * H2, 0.005 * H2
        v = [oracle.gauss_tau(oracle.params(DfN=D, VeN=0.1), xy, TL, 1.0, g)["tlsce"] ** -2 for D in (D1, D2)]
        ok &= close(v[0] - v[1], (D1 * D1 - D2 * D2) / (H2 * H2), scale=v[0])
    return ok


def law_discontinuity_capturing(oracle, T):
    """DfN = 0: tdcs = H2 / sqrt(S:S/2) and tau_LSCE^-2 gains Con_EQ:Con_EQ / S:S
    over its value with DfN just above the 1e-12 switch."""
    ok = True
    tiny = 2e-12
    for xy in elements():
        TL = nodal(xy, L=np.array([[0.4, 1.3], [-0.2, -0.4]]), u0=(0.6, -0.3), S0=(1.5, 0.3, 0.7),
                   dS=np.array([[0.4, -0.2], [0.1, 0.3], [-0.3, 0.2]]))
        for g in (0, 4):
            p0 = oracle.params(DfN=0.0, VeN=0.2, WiN0=1.4)
            pt = oracle.params(DfN=tiny, VeN=0.2, WiN0=1.4)
            o0 = oracle.gauss_tau(p0, xy, TL, 0.9, g)
            ot = oracle.gauss_tau(pt, xy, TL, 0.9, g)
            f = oracle.gauss_fields(xy, TL, g)
            Sv = f["S"]
            SS = Sv[0] ** 2 + 2 * Sv[1] ** 2 + Sv[2] ** 2
            ps = oracle.point_strong_form(p0, f["U"], f["GU"], f["G2U"], f["P"], f["dP"], Sv, f["dS"], 0.9)
            CE2 = np.sum(ps["Con_EQ"] ** 2)
            H2 = geometry_at(xy, TL, g, T)[0]
            ok &= ot["tdcs"] == 0.0
            ok &= close(o0["tdcs"], H2 / np.sqrt(0.5 * SS))
            lhs = o0["tlsce"] ** -2 - ot["tlsce"] ** -2 + (tiny / H2) ** 2
            ok &= close(lhs, CE2 / SS, scale=o0["tlsce"] ** -2)
    return ok


LAWS = [law_geometry, law_lsic, law_viscous_limit, law_polymer_limit, law_inertia, law_constitutive_time_scale,
        law_model_functions, law_diffusivity, law_discontinuity_capturing]


@pytest.fixture(scope="module")
def tables(oracle):
    return oracle.tables()


@pytest.mark.parametrize("law", LAWS, ids=lambda f: f.__name__)
def test_stabilization_law(oracle, tables, law):
    assert law(oracle, tables)


@pytest.mark.parametrize("plant", range(1, 15))
def test_planted_mistake_breaks_a_law(oracle, tables, plant):
    oracle.tau_plant(plant)
    try:
        broken = [law.__name__ for law in LAWS if not law(oracle, tables)]
    finally:
        oracle.tau_plant(0)
    assert broken, f"plant {plant} passes every law"
    assert all(law(oracle, tables) for law in LAWS)


def test_gauss_point_index_checked(oracle):
    xy = elements()[0]
    with pytest.raises(ValueError):
        oracle.gauss_tau(oracle.params(), xy, nodal(xy), 1.0, 9)
    with pytest.raises(ValueError):
        oracle.tau_plant(15)
