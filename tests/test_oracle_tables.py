"""a1: Gauss points, weights and Q2 basis tables of the oracle, pinned by closed
forms (Lagrange products of the 1D quadratics), polynomial reproduction and
the quadrature rule's moments, including the reference's shifted centre point
(src/GaussIntegration/Gauss_Quadrangle_mod.F90:69-72)."""
import numpy as np

from conftest import reference_nodes

NODES = reference_nodes()


def lagrange_1d(s):
    """Quadratic Lagrange polynomials on nodes -1, 0, 1 and their derivatives."""
    L = {-1: s * (s - 1) / 2, 0: 1 - s * s, 1: s * (s + 1) / 2}
    dL = {-1: s - 0.5, 0: -2 * s, 1: s + 0.5}
    d2L = {-1: 1.0, 0: -2.0, 1: 1.0}
    return L, dL, d2L


def closed_form(c, e):
    Lc, dLc, d2Lc = lagrange_1d(c)
    Le, dLe, d2Le = lagrange_1d(e)
    out = {k: np.zeros(9) for k in ("bfn", "dfdc", "dfde", "d2fdc2", "d2fde2", "d2fdce")}
    for n, (a, b) in enumerate(NODES.astype(int)):
        out["bfn"][n] = Lc[a] * Le[b]
        out["dfdc"][n] = dLc[a] * Le[b]
        out["dfde"][n] = Lc[a] * dLe[b]
        out["d2fdc2"][n] = d2Lc[a] * Le[b]
        out["d2fde2"][n] = Lc[a] * d2Le[b]
        out["d2fdce"][n] = dLc[a] * dLe[b]
    return out


def test_points_and_weights(oracle):
    T = oracle.tables()
    assert T["P1"][0] == 0.0 + 1.0e-8  # shifted centre abscissa
    assert T["P1"][1] == np.sqrt(3.0 / 5.0) and T["P1"][2] == -np.sqrt(3.0 / 5.0)
    assert list(T["W1"]) == [8.0 / 9.0, 5.0 / 9.0, 5.0 / 9.0]
    # 2D: xi outer, eta inner
    for i in range(3):
        for j in range(3):
            g = 3 * i + j
            assert T["P2"][g, 0] == T["P1"][i] and T["P2"][g, 1] == T["P1"][j]
            assert T["W2"][g] == T["W1"][i] * T["W1"][j]


def test_rule_moments(oracle):
    """Exact for polynomials up to degree 5 except the O(1e-8) odd moments the
    shifted centre point introduces."""
    T = oracle.tables()
    p, w = T["P1"], T["W1"]
    for k in range(6):
        exact = 0.0 if k % 2 else 2.0 / (k + 1)
        got = np.sum(w * p ** k)
        if k == 1:
            assert abs(got - (8.0 / 9.0) * 1e-8) < 1e-16  # odd moment = w_1 * 1e-8
        else:
            assert abs(got - exact) < 1e-14 + (1e-15 if k else 0)


def test_bulk_tables_closed_form(oracle):
    T = oracle.tables()
    for g in range(9):
        cf = closed_form(*T["P2"][g])
        for k in cf:
            np.testing.assert_allclose(T[k][g], cf[k], rtol=0, atol=4e-16, err_msg=f"{k} at gp {g}")


def test_edge_tables_closed_form(oracle):
    T = oracle.tables()
    for ed in range(4):
        for g in range(3):
            s = T["P1"][g]
            c, e = [(s, -1.0), (1.0, s), (s, 1.0), (-1.0, s)][ed]
            cf = closed_form(c, e)
            for k in cf:
                np.testing.assert_allclose(T[k + "E"][ed, g], cf[k], rtol=0, atol=4e-16)


def test_polynomial_reproduction(oracle):
    """sum_n phi_n p(node_n) = p for the biquadratic monomials; derivative tables
    reproduce the derivatives (pins node order, signs and factors)."""
    T = oracle.tables()
    xi, eta = NODES[:, 0], NODES[:, 1]
    monos = [(0, 0), (1, 0), (0, 1), (1, 1), (2, 0), (0, 2), (2, 1), (1, 2), (2, 2)]
    for g in range(9):
        c, e = T["P2"][g]
        for a, b in monos:
            vals = xi ** a * eta ** b
            assert abs(T["bfn"][g] @ vals - c ** a * e ** b) < 1e-15
            d_c = a * c ** (a - 1) * e ** b if a else 0.0
            d_e = b * c ** a * e ** (b - 1) if b else 0.0
            assert abs(T["dfdc"][g] @ vals - d_c) < 1e-14
            assert abs(T["dfde"][g] @ vals - d_e) < 1e-14
            d_cc = a * (a - 1) * c ** (a - 2) * e ** b if a > 1 else 0.0
            d_ee = b * (b - 1) * c ** a * e ** (b - 2) if b > 1 else 0.0
            d_ce = a * b * c ** (a - 1) * e ** (b - 1) if a and b else 0.0
            assert abs(T["d2fdc2"][g] @ vals - d_cc) < 1e-14
            assert abs(T["d2fde2"][g] @ vals - d_ee) < 1e-14
            assert abs(T["d2fdce"][g] @ vals - d_ce) < 1e-14


def test_edge_basis_vanishes_off_edge(oracle):
    """On edge ied only its three nodes carry basis weight (exactly zero elsewhere)."""
    from ve_flow_around_cylinder_fortran_amd.mesh import EDGE_NODES

    T = oracle.tables()
    for ied in (1, 2, 3, 4):
        off = [n for n in range(9) if n not in EDGE_NODES[ied]]
        assert np.all(T["bfnE"][ied - 1][:, off] == 0.0)
