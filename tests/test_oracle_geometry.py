"""a3/a4: isoparametric geometry (basis2D) and Gauss-point interpolation
(setFlowUnknowns) of the oracle, pinned by closed forms on affine elements,
isoparametric reproduction on curved elements, exact areas, and exact
reproduction of quadratic fields."""
import numpy as np
import pytest

from conftest import reference_nodes
from test_oracle_tables import closed_form

NODES = reference_nodes()


def affine_element(rng):
    A = np.array([[0.7, 0.2], [-0.1, 0.5]]) + 0.1 * rng.standard_normal((2, 2))
    c = rng.standard_normal(2)
    return NODES @ A.T + c, A, c


def test_affine_geometry_closed_form(oracle):
    rng = np.random.default_rng(1)
    T = oracle.tables()
    for _ in range(5):
        xy, A, _ = affine_element(rng)
        Ainv = np.linalg.inv(A)
        for g in range(9):
            G = oracle.basis2d(xy, g)
            assert G["detJ"] == pytest.approx(np.linalg.det(A), rel=1e-13)
            assert G["WET"] == pytest.approx(T["W2"][g] * abs(np.linalg.det(A)), rel=1e-13)
            cf = closed_form(*T["P2"][g])
            grad_ref = np.stack([cf["dfdc"], cf["dfde"]])  # (2, 9)
            np.testing.assert_allclose(G["dFdXi"], Ainv.T @ grad_ref, rtol=0, atol=1e-13)
            for n in range(9):
                H = np.array([[cf["d2fdc2"][n], cf["d2fdce"][n]], [cf["d2fdce"][n], cf["d2fde2"][n]]])
                Hx = Ainv.T @ H @ Ainv
                np.testing.assert_allclose(G["d2FdXi2"][:, n], np.diag(Hx), rtol=0, atol=1e-12)


def curved_element(rng):
    xy, _, _ = affine_element(rng)
    xy = xy.copy()
    xy[4:9] += 0.05 * rng.standard_normal((5, 2))  # move mid-edge and centre nodes
    return xy


def test_curved_geometry_reproduces_linears(oracle):
    rng = np.random.default_rng(2)
    for _ in range(5):
        xy = curved_element(rng)
        for g in range(9):
            G = oracle.basis2d(xy, g)
            dF = G["dFdXi"]
            np.testing.assert_allclose(dF @ xy[:, 0], [1.0, 0.0], atol=1e-13)
            np.testing.assert_allclose(dF @ xy[:, 1], [0.0, 1.0], atol=1e-13)
            np.testing.assert_allclose(dF.sum(axis=1), [0.0, 0.0], atol=1e-13)


def test_curved_second_derivatives_omit_map_curvature(oracle):
    """d2FdXi2 = diag(J^-T H_xi J^-1): the map's own second derivatives are not
    included (src/Element_Calculations_mod.F90:352-355)."""
    rng = np.random.default_rng(3)
    T = oracle.tables()
    xy = curved_element(rng)
    for g in range(9):
        cf = closed_form(*T["P2"][g])
        J = np.array([[cf["dfdc"] @ xy[:, 0], cf["dfde"] @ xy[:, 0]], [cf["dfdc"] @ xy[:, 1], cf["dfde"] @ xy[:, 1]]])
        Jinv = np.linalg.inv(J)
        G = oracle.basis2d(xy, g)
        for n in range(9):
            H = np.array([[cf["d2fdc2"][n], cf["d2fdce"][n]], [cf["d2fdce"][n], cf["d2fde2"][n]]])
            np.testing.assert_allclose(G["d2FdXi2"][:, n], np.diag(Jinv.T @ H @ Jinv), rtol=1e-12, atol=1e-12)


def test_area_of_straight_sided_quad(oracle):
    corners = np.array([[0.1, 0.0], [1.3, 0.2], [1.0, 1.1], [-0.2, 0.9]])
    mids = 0.5 * (corners + np.roll(corners, -1, axis=0))
    centre = corners.mean(axis=0)
    xy = np.vstack([corners, mids, centre])
    area = 0.5 * abs(np.dot(corners[:, 0], np.roll(corners[:, 1], -1)) - np.dot(corners[:, 1], np.roll(corners[:, 0], -1)))
    total = sum(oracle.basis2d(xy, g)["WET"] for g in range(9))
    assert total == pytest.approx(area, rel=1e-14)


def test_interpolation_reproduces_quadratics(oracle):
    """GU(j,i) = dU_i/dx_j, G2U(j,i) = d2U_i/dx_j^2, dS(s,d) = dS_s/dx_d,
    exact for quadratic fields on an affine element."""
    rng = np.random.default_rng(4)
    xy, A, c = affine_element(rng)
    coef = rng.standard_normal((6, 6))  # field q: a + b x + c y + d x^2 + e x y + f y^2

    def f(q, x, y):
        a, b, cc, d, e, ff = coef[q]
        return a + b * x + cc * y + d * x * x + e * x * y + ff * y * y

    def grad(q, x, y):
        a, b, cc, d, e, ff = coef[q]
        return np.array([b + 2 * d * x + e * y, cc + e * x + 2 * ff * y])

    def hess_diag(q):
        return np.array([2 * coef[q, 3], 2 * coef[q, 5]])

    TL = np.stack([f(q, xy[:, 0], xy[:, 1]) for q in range(6)], axis=1)
    T = oracle.tables()
    for g in range(9):
        xg = A @ T["P2"][g] + c
        F = oracle.gauss_fields(xy, TL, g)
        for i in range(2):
            assert F["U"][i] == pytest.approx(f(i, *xg), abs=1e-13)
            np.testing.assert_allclose(F["GU"][:, i], grad(i, *xg), atol=1e-12)
            np.testing.assert_allclose(F["G2U"][:, i], hess_diag(i), atol=1e-11)
        assert F["P"] == pytest.approx(f(2, *xg), abs=1e-13)
        np.testing.assert_allclose(F["dP"], grad(2, *xg), atol=1e-12)
        for s in range(3):
            assert F["S"][s] == pytest.approx(f(3 + s, *xg), abs=1e-13)
            np.testing.assert_allclose(F["dS"][s], grad(3 + s, *xg), atol=1e-12)
