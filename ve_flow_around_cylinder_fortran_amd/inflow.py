"""Fully developed inflow: the reference's Inflow_1D problem, whose solution
gives the inlet's Dirichlet values for any constitutive model.

The reference solves it inside every main-problem assembly of an arclength
continuation (loopNewtonRaphsonInflow, src/Newton_Raphson_mod.F90:176-277,
called at :369-372 with WiN = WiN0*EU of the main state,
src/Element_Calculations_mod.F90:592-597) and inserts the result as Dirichlet
rows of Ux, Uy, Sxx, Sxy, Syy at the inlet nodes
(applyDirichletBCFullyDeveloped, src/Dirichlet_mod.F90:149-190).  The problem
(src/Physical_mod.F90:31-52): quadratic line elements along y over the inlet
nodes, unknowns Ux, Uy and the three square-root-conformation components per
node plus the pressure gradient dPL, walls (Ux = Uy = 0) at both ends, and
the flow rate 2*Height.  At its Gauss points every x-derivative, the Uy
gradient, P and all stress gradients are zero
(src/Element_Calculations_mod.F90:516-525), so the equations are
    momentum x:  (dPL phi + P_tot(1,2) dphi/dy) WET        (momentumInflow,
                 src/Residuals_mod.F90:494-540; ReN dUf/dt = 0 when steady)
    Uy:          Uy phi WET
    stresses:    Con_EQ(j,k) phi WET                        (stressesInflow,
                 :544-595; the DfN term multiplies dT/dXi = 0)
    flow rate:   sum of Ux WET = 2*Height                    (extraEquations-
                 ResidualsBulkInflow, src/Extra_Equations_mod.F90:133-150)
with Con_EQ and P_tot those of the square-root formulation in simple shear.

Host code on the inlet's few hundred unknowns (a Newton iteration with a
grouped forward-difference Jacobian and a sparse LU: the node rows are banded,
plus the dPL column and the flow-rate row; as small as the reference's own
PETSc solve of it); the main problem's assembly stays on the GPU.
"""
from __future__ import annotations

import math

import numpy as np

from .mesh import MESH_TOLERANCE, Mesh

# 3-point Gauss rule with the centre abscissa shifted by 1e-8, as the
# reference's line and quadrangle rules have it (Gauss_Quadrangle_mod.F90:69-72).
_XI = np.array([0.0 + 1.0e-8, math.sqrt(3.0 / 5.0), -math.sqrt(3.0 / 5.0)])
_W = np.array([8.0 / 9.0, 5.0 / 9.0, 5.0 / 9.0])
# The reference stops at 1e-8 on both norms (Cor_Norm_eps); iterating on to
# 1e-12 makes the values independent of the starting point (the previous
# solution), so every caller at the same WiN gets the same Dirichlet values.
INFLOW_TOL = 1.0e-12


class InflowError(RuntimeError):
    pass


def simple_shear_terms(g, S00, S01, S11, WiN, BvN, VeN, model):
    """Con_EQ (xx, xy, yy) and P_tot(1,2) of the square-root formulation for
    U = (u(y), 0) with du/dy = g and spatially uniform S (strongFormConEQ_SQRT,
    src/Constitutive_Models_SQRT_mod.F90:120-509, with GU(2,1) = g the only
    velocity gradient and dS/dXi = 0).  Arrays broadcast."""
    # C = S.S (the 2x2 block; C33 = 0 because S33 is never set)
    C00 = S00 * S00 + S01 * S01
    C01 = S00 * S01 + S01 * S11
    C11 = S01 * S01 + S11 * S11
    if model == 0:  # L-PTT: f(1) = f(2) = 1 + eps (tr C - 2), f(3) = 0
        fT = 1.0 + VeN * ((C00 + C11) - 2.0)
        f0, f1, f2 = fT, fT, 0.0
    else:  # Giesekus
        f0, f1, f2 = 1.0 - VeN, 1.0 - 2.0 * VeN, VeN
    # omega(1,2) from S and the (traceless, here plain) transposed gradient
    w = (S11 * g) / (S00 + S11)
    # UCD = U.grad S (= 0) - S.GU_tr - W.S, GU_tr = [[0, 0], [g, 0]]
    U00 = -(S01 * g) - w * S01
    U01 = -(w * S11)
    U11 = w * S01
    det = S00 * S11 - S01 * S01
    SI00, SI01, SI11 = S11 / det, -S01 / det, S00 / det
    t00 = f1 * S00 - f0 * SI00
    t01 = f1 * S01 - f0 * SI01
    t11 = f1 * S11 - f0 * SI11
    if model != 0:
        t00 = t00 + f2 * (S00 * C00 + S01 * C01)
        t01 = t01 + f2 * (S00 * C01 + S01 * C11)
        t11 = t11 + f2 * (S01 * C01 + S11 * C11)
    CE00 = WiN * U00 + t00 / 2.0
    CE01 = WiN * U01 + t01 / 2.0
    CE11 = WiN * U11 + t11 / 2.0
    a = (1.0 - BvN) / WiN
    P01 = a * C01 + BvN * g  # T_ve(1,2) + BvN*G_dot_tr(1,2)
    return CE00, CE01, CE11, P01


# half-bandwidth of the inflow problem's node block (5 unknowns per node,
# rows of node k coupled to nodes k-2 .. k+2)
_KL = 14


class _BorderedBandLU:
    """LU of [[A, c], [f, d]] with A banded (kl sub- and superdiagonals; LAPACK
    gbtrf with partial pivoting), c and f one dense column and row, d a scalar:
    A y = c once, then the Schur complement s = d - f.y; a solve is
    x1 = A^-1 b1, xn = (bn - f.x1) / s, x1 -= xn y."""

    def __init__(self, lu, piv, kl, y, f, s):
        self.lu, self.piv, self.kl, self.y, self.f, self.s = lu, piv, kl, y, f, s

    @classmethod
    def factor(cls, rows, cols, vals, n, kl):
        """From the (row, column, value) entries of an n x n matrix (each entry
        once); None when the node block is singular, the entries leave the
        band, or the Schur complement is zero."""
        from scipy.linalg.lapack import dgbtrf, dgbtrs

        m = n - 1
        inner = (rows < m) & (cols < m)
        ri, ci = rows[inner], cols[inner]
        if np.any(np.abs(ri - ci) > kl):
            return None
        ab = np.zeros((3 * kl + 1, m))
        ab[2 * kl + ri - ci, ci] = vals[inner]
        lu, piv, info = dgbtrf(ab, kl, kl)
        if info != 0:
            return None
        c = np.zeros(m)
        f = np.zeros(m)
        d = 0.0
        col_b = (cols == m) & (rows < m)
        c[rows[col_b]] = vals[col_b]
        row_b = (rows == m) & (cols < m)
        f[cols[row_b]] = vals[row_b]
        corner = (rows == m) & (cols == m)
        if np.any(corner):
            d = float(vals[corner][0])
        y, info = dgbtrs(lu, kl, kl, c, piv)
        s = d - float(f @ y)
        if info != 0 or s == 0.0 or not math.isfinite(s):
            return None
        return cls(lu, piv, kl, y, f, s)

    def solve(self, b):
        from scipy.linalg.lapack import dgbtrs

        x1, info = dgbtrs(self.lu, self.kl, self.kl, b[:-1], self.piv)
        xn = (b[-1] - float(self.f @ x1)) / self.s
        return np.concatenate([x1 - xn * self.y, [xn]])


class FullyDevelopedInflow:
    """The Inflow_1D problem on given node positions y (ascending; an odd
    count, consecutive triples form the quadratic elements) of a channel of
    half-width H, for Params (ReN is irrelevant: the flow is steady)."""

    def __init__(self, y: np.ndarray, H: float, params):
        y = np.asarray(y, dtype=np.float64)
        if y.ndim != 1 or len(y) < 3 or len(y) % 2 == 0 or np.any(np.diff(y) <= 0):
            raise InflowError("inflow nodes: an odd number >= 3 of ascending y positions")
        self.y, self.H, self.params = y, float(H), params
        self.model = {"L-PTT": 0, "GIESEKUS": 1}[params.model]
        nn = len(y)
        self.nn = nn
        self.el = np.stack([np.arange(0, nn - 2, 2), np.arange(1, nn - 1, 2), np.arange(2, nn, 2)], axis=1)
        xi = _XI[:, None]
        phi = np.concatenate([xi * (xi - 1) / 2, 1 - xi * xi, xi * (xi + 1) / 2], axis=1)  # [g][3]
        dphi = np.concatenate([xi - 0.5, -2 * xi, xi + 0.5], axis=1)
        ye = y[self.el]                                   # [e][3]
        J = dphi @ ye.T                                   # [g][e] dy/dxi
        self.phi = phi                                    # [g][n]
        self.dphi_dy = dphi[:, None, :] / J[:, :, None]   # [g][e][n]
        self.WET = _W[:, None] * np.abs(J)                # [g][e]
        self.walls = np.array([0, nn - 1])
        self._groups = self._jacobian_groups()
        self.z = None  # last solution (warm start)
        self._lu = None  # factorised Jacobian, kept across calls
        self.newton_iterations = []
        self.jacobians = 0

    @classmethod
    def from_mesh(cls, mesh: Mesh, params):
        """The inlet nodes of a cylinder mesh (x = -L_up), in ascending y."""
        g = mesh.geometry
        nodes = np.nonzero(np.abs(mesh.coords[:, 0] + g["L_up"]) < MESH_TOLERANCE)[0]
        order = np.argsort(mesh.coords[nodes, 1])
        obj = cls(mesh.coords[nodes[order], 1], g["H"], params)
        obj.nodes = nodes[order]
        return obj

    def residual(self, z: np.ndarray, WiN: float) -> np.ndarray:
        """Residual of the unknowns z = [node-major (Ux, Uy, Sxx, Sxy, Syy)..., dPL]
        with the wall rows replaced by Ux and Uy themselves."""
        return self._residuals(np.asarray(z, dtype=np.float64)[None, :], WiN)[0]

    def _residuals(self, Z: np.ndarray, WiN: float) -> np.ndarray:
        """residual() of each row of Z [B][n], in one pass over the batch.
        Every sum is written out in a fixed order (over the 3 element nodes,
        the 3 Gauss points, the 2 elements of a shared node), so a row's
        residual does not depend on the batch it is evaluated in."""
        p = self.params
        B = Z.shape[0]
        nod = Z[:, :-1].reshape(B, self.nn, 5)
        dPL = Z[:, -1][:, None, None, None]                        # [B][1][1][1]
        v = nod[:, self.el]                                        # [B][e][3][5]
        ph = self.phi[None, :, None, :, None]                      # [1][g][1][3][1]
        vv = v[:, None]                                            # [B][1][e][3][5]
        Fg = (ph[..., 0, :] * vv[..., 0, :] + ph[..., 1, :] * vv[..., 1, :]) + ph[..., 2, :] * vv[..., 2, :]  # [B][g][e][5]
        dy = self.dphi_dy[None]                                    # [1][g][e][3]
        u = v[:, None, :, :, 0]                                    # [B][1][e][3]
        gU = (dy[..., 0] * u[..., 0] + dy[..., 1] * u[..., 1]) + dy[..., 2] * u[..., 2]  # du/dy [B][g][e]
        CE00, CE01, CE11, P01 = simple_shear_terms(gU, Fg[..., 2], Fg[..., 3], Fg[..., 4], WiN, p.BvN, p.VeN,
                                                   self.model)
        W = self.WET[None, :, :, None]                             # [1][g][e][1]
        phn = self.phi[None, :, None, :]                           # [1][g][1][3]
        t = np.empty((B, 3, self.el.shape[0], 3, 5))               # [B][g][e][n][5]
        t[..., 0] = (dPL * phn + P01[..., None] * dy) * W          # momentum x
        t[..., 1] = (Fg[..., 1:2] * phn) * W                       # Uy
        t[..., 2] = (CE00[..., None] * phn) * W
        t[..., 3] = (CE01[..., None] * phn) * W
        t[..., 4] = (CE11[..., None] * phn) * W
        T = (t[:, 0] + t[:, 1]) + t[:, 2]                          # [B][e][n][5]
        R = np.zeros((B, self.nn, 5))
        # element order; a node shared by two elements gets (0 + a) + b,
        # which equals (0 + b) + a
        R[:, self.el[:, 0]] += T[:, :, 0]
        R[:, self.el[:, 1]] += T[:, :, 1]
        R[:, self.el[:, 2]] += T[:, :, 2]
        R[:, self.walls, 0] = nod[:, self.walls, 0]
        R[:, self.walls, 1] = nod[:, self.walls, 1]
        flow = (Fg[..., 0] * self.WET[None]).reshape(B, -1).sum(axis=1) - 2.0 * self.H
        return np.concatenate([R.reshape(B, -1), flow[:, None]], axis=1)

    def initial_guess(self, WiN: float) -> np.ndarray:
        """The Oldroyd-B closed form: U = 1.5 (1 - (y/H)^2), S = sqrt(C), dPL = -3/H^2."""
        from .newton import oldroyd_b_inflow

        U, Uy, Sxx, Sxy, Syy = oldroyd_b_inflow(self.y, self.H, WiN)
        return np.concatenate([np.stack([U, Uy, Sxx, Sxy, Syy], axis=1).ravel(), [-3.0 / self.H ** 2]])

    def _factor(self, z: np.ndarray, WiN: float, r: np.ndarray):
        """LU of the forward-difference Jacobian at z (r = residual(z)): the
        node block is banded (14 sub- and superdiagonals: node k's rows hold
        the unknowns of nodes k-2 .. k+2), bordered by the dPL column and the
        flow-rate row, so it is factorised by LAPACK's banded LU with partial
        pivoting and the border by its Schur complement; SuperLU of the whole
        matrix when the node block is singular.  ValueError for non-finite
        entries, as a dense factorisation's finiteness check gives."""
        rr, cc, vv = self._jacobian_parts(z, WiN, r)
        if not np.all(np.isfinite(vv)):
            raise ValueError("array must not contain infs or NaNs")
        self.jacobians += 1
        n = len(z)
        lu = _BorderedBandLU.factor(rr, cc, vv, n, _KL)
        if lu is not None:
            return lu
        import scipy.sparse as sp
        from scipy.sparse.linalg import splu

        try:
            return splu(sp.csc_matrix((vv, (rr, cc)), shape=(n, n)), permc_spec="NATURAL")
        except RuntimeError as e:  # an exactly singular Jacobian
            raise ValueError(str(e)) from None

    def _jacobian_groups(self):
        """For each group of columns perturbed together (node unknown q of the
        nodes k with k % 5 == c): the columns, and each (row, column) pair of
        their entries once (the rows of nodes k-2 .. k+2 within the nodes)."""
        nn = self.nn
        node = np.arange(nn)
        groups = []
        for c in range(5):
            members = node[node % 5 == c]
            near = members[:, None] + np.arange(-2, 3)[None, :]             # [m][5] nodes
            ok = (near >= 0) & (near < nn)
            mi, k = np.nonzero(ok)
            rows_n = near[mi, k]                                             # node of each (member, near) pair
            for q in range(5):
                cols = members * 5 + q
                rr = (rows_n[:, None] * 5 + np.arange(5)[None, :]).ravel()
                cc = np.repeat(cols[mi], 5)
                groups.append((cols, rr, cc))
        return groups

    def _jacobian_parts(self, z: np.ndarray, WiN: float, r: np.ndarray):
        """Entries (rows, cols, values) of the node columns, the dPL column
        and the flow-rate row of the forward-difference Jacobian at z."""
        n = len(z)
        h = 1.0e-7 * np.maximum(1.0, np.abs(z))
        ng = len(self._groups)
        Zp = np.repeat(z[None, :], ng + 1, axis=0)  # one perturbed state per group, dPL last
        for i, (cols, _, _) in enumerate(self._groups):
            Zp[i, cols] += h[cols]
        Zp[ng, -1] += h[-1]
        D = self._residuals(Zp, WiN) - r[None, :]
        R, C, V = [], [], []
        for i, (cols, rr, cc) in enumerate(self._groups):
            R.append(rr)
            C.append(cc)
            V.append(D[i, rr] / h[cc])
        R.append(np.arange(n - 1))
        C.append(np.full(n - 1, n - 1))
        V.append((D[ng] / h[-1])[:-1])
        flow = np.zeros(self.nn)
        np.add.at(flow, self.el, np.einsum("gn,ge->en", self.phi, self.WET))
        R.append(np.full(self.nn, n - 1))
        C.append(np.arange(self.nn) * 5)
        V.append(flow)
        return np.concatenate(R), np.concatenate(C), np.concatenate(V)

    def jacobian(self, z: np.ndarray, WiN: float, r: np.ndarray) -> np.ndarray:
        """The forward-difference Jacobian at z (r = residual(z)), formed
        with grouped columns: the rows of node k depend only on the nodes of
        the elements that hold k (k-2 .. k+2), so the columns of nodes five
        apart perturb disjoint rows and one residual evaluation serves a whole
        group (25 groups of node unknowns, and dPL alone); each node row keeps
        exactly the quotient a one-column perturbation gives it.  The
        flow-rate row is linear in Ux and taken exactly (sum of phi*WET per
        node).  Dense; the factorisation takes the same entries sparse."""
        n = len(z)
        J = np.zeros((n, n))
        rr, cc, vv = self._jacobian_parts(z, WiN, r)
        J[rr, cc] = vv
        return J

    def solve(self, WiN: float, max_iter: int = 100):
        """Newton on the inflow problem at WiN from the last solution (or the
        Oldroyd-B closed form) until both norms are below INFLOW_TOL; returns
        (Ux, Uy, Sxx, Sxy, Syy) at the nodes and dPL.  The factorised Jacobian
        is kept across calls (WiN moves little between them) and formed again
        when a correction shrinks by less than a factor 0.3."""
        z = self.z.copy() if self.z is not None else self.initial_guess(WiN)
        prev = None
        for it in range(1, max_iter + 1):
            r = self.residual(z, WiN)
            if not np.all(np.isfinite(r)):
                # e.g. a negative conformation under a square root at a wild WiN
                self._lu = None
                raise InflowError(f"inflow Newton: non-finite residual at WiN = {WiN}")
            if self._lu is None:
                try:
                    self._lu = self._factor(z, WiN, r)
                except ValueError as e:  # non-finite Jacobian entries
                    raise InflowError(f"inflow Newton: {e} at WiN = {WiN}") from None
            dz = self._lu.solve(r)
            z = z - dz
            nd = float(np.linalg.norm(dz))
            if not np.all(np.isfinite(z)):
                raise InflowError(f"inflow Newton diverged at WiN = {WiN}")
            if nd < INFLOW_TOL and np.linalg.norm(r) < INFLOW_TOL:
                break
            if prev is not None and nd > 0.3 * prev:
                self._lu = None
            prev = nd
        else:
            raise InflowError(f"inflow Newton: no convergence in {max_iter} iterations at WiN = {WiN}")
        self.z = z
        self.WiN = WiN
        self.newton_iterations.append(it)
        nod = z[:-1].reshape(self.nn, 5)
        return tuple(nod[:, q].copy() for q in range(5)), float(z[-1])

    def sensitivity(self, WiN: float):
        """d(Ux, Uy, Sxx, Sxy, Syy)/dWiN at the nodes for the solution at WiN
        (solved first if it is not the last one): J dz/dWiN = -dR/dWiN with J
        the Jacobian at the solution and dR/dWiN a forward difference in WiN."""
        if self.z is None or getattr(self, "WiN", None) != WiN:
            self.solve(WiN)
        z = self.z
        r = self.residual(z, WiN)
        h = 1.0e-7 * max(1.0, abs(WiN))
        dR = (self.residual(z, WiN + h) - r) / h
        try:
            lu = self._factor(z, WiN, r)
        except ValueError as e:  # non-finite entries next to the solution
            raise InflowError(f"inflow sensitivity: {e} at WiN = {WiN}") from None
        dz = -lu.solve(dR)
        self._lu = lu  # the Jacobian at this solution: the next solves start from it
        nod = dz[:-1].reshape(self.nn, 5)
        return tuple(nod[:, q].copy() for q in range(5))
