"""TEST INFRASTRUCTURE ONLY: ctypes access to the CPU oracle (libvef_oracle.so).

Loaded by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg, as
the checker.  See vef_oracle.cpp for what it restates.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
DBL_MAX = float(np.finfo(np.float64).max)


class OracleParams(ctypes.Structure):
    _fields_ = [("ReN", ctypes.c_double), ("BvN", ctypes.c_double), ("VeN", ctypes.c_double),
                ("DfN", ctypes.c_double), ("WiN0", ctypes.c_double), ("dt", ctypes.c_double),
                ("model", ctypes.c_int32), ("pad", ctypes.c_int32)]


def params(ReN=0.0, BvN=0.59, VeN=0.0, DfN=1.0e-3, WiN0=1.0, model="L-PTT") -> OracleParams:
    m = {"L-PTT": 0, "GIESEKUS": 1}[model]
    return OracleParams(ReN, BvN, VeN, DfN, WiN0, DBL_MAX, m, 0)


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


_libs = {}


def lib(flops: bool = False):
    name = "libvef_oracle_flops.so" if flops else "libvef_oracle.so"
    if name not in _libs:
        path = os.path.join(_HERE, name)
        if not os.path.exists(path):
            build()
        L = ctypes.CDLL(path)
        vp = ctypes.c_void_p
        if flops:
            L.vo_count_flops.argtypes = [ctypes.POINTER(OracleParams), vp, vp, ctypes.c_double, ctypes.c_int, vp]
            L.vo_count_flops_stability.argtypes = [ctypes.POINTER(OracleParams), vp, vp, ctypes.c_double, vp, vp]
            L.vo_count_flops_geometry.argtypes = [ctypes.POINTER(OracleParams), vp, vp]
        else:
            L.vo_tables.argtypes = [vp]
            L.vo_basis2d.argtypes = [vp, ctypes.c_int, ctypes.c_int, vp]
            L.vo_gauss_fields.argtypes = [vp, vp, ctypes.c_int, vp]
            L.vo_gauss_tau.argtypes = [ctypes.POINTER(OracleParams), vp, vp, ctypes.c_double, ctypes.c_int, vp]
            L.vo_tau_plant.argtypes = [ctypes.c_int]
            L.vo_term_plant.argtypes = [ctypes.c_int]
            L.vo_node_terms.argtypes = [ctypes.POINTER(OracleParams), vp, vp, ctypes.c_double, ctypes.c_int, vp]
            L.vo_point_strong_form.argtypes = [ctypes.POINTER(OracleParams), vp, vp, vp, ctypes.c_double, vp, vp, vp,
                                               ctypes.c_double, vp]
            L.vo_element.argtypes = [ctypes.POINTER(OracleParams), vp, vp, ctypes.c_double, ctypes.c_int, ctypes.c_int,
                                     vp, vp, vp]
            L.vo_arclength.argtypes = [ctypes.c_int, ctypes.c_int, vp, vp, vp, ctypes.c_int, ctypes.c_double,
                                       ctypes.c_double, ctypes.c_double, ctypes.c_int, vp, vp, vp, vp]
            L.vo_drag.argtypes = [ctypes.POINTER(OracleParams), ctypes.c_int, ctypes.c_int, vp, vp, vp, ctypes.c_int,
                                  vp, vp, vp]
            L.vo_nodal_stress.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_int, vp, vp]
            L.vo_pattern_nnz.argtypes = [ctypes.c_int, ctypes.c_int, vp]
            L.vo_pattern_nnz.restype = ctypes.c_int64
            L.vo_pattern.argtypes = [ctypes.c_int, ctypes.c_int, vp, vp, vp]
            L.vo_assemble.argtypes = [ctypes.POINTER(OracleParams), ctypes.c_int, ctypes.c_int, vp, vp, ctypes.c_int, vp,
                                      vp, ctypes.c_double, ctypes.c_int, vp, vp, vp, vp, ctypes.c_int, ctypes.c_int,
                                      ctypes.c_int]
            L.vo_stab_plant.argtypes = [ctypes.c_int]
            L.vo_stab_element.argtypes = [ctypes.POINTER(OracleParams), vp, vp, ctypes.c_double, vp, vp]
            L.vo_stab_residual.argtypes = [ctypes.POINTER(OracleParams), vp, vp, ctypes.c_double, vp, vp, vp]
            L.vo_stab_linearization.argtypes = [ctypes.POINTER(OracleParams), vp, vp, ctypes.c_double, vp, ctypes.c_int,
                                                ctypes.c_double, vp, vp]
            L.vo_stab_pattern_nnz.argtypes = [ctypes.c_int, ctypes.c_int, vp]
            L.vo_stab_pattern_nnz.restype = ctypes.c_int64
            L.vo_stab_pattern.argtypes = [ctypes.c_int, ctypes.c_int, vp, vp, vp]
            L.vo_stab_assemble.argtypes = [ctypes.POINTER(OracleParams), ctypes.c_int, ctypes.c_int, vp, vp, vp, vp, vp,
                                           vp, vp, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        _libs[name] = L
    return _libs[name]


def _p(a):
    return ctypes.c_void_p(a.ctypes.data) if a is not None else None


def tables():
    out = np.zeros(6 + 18 + 9 + 6 * 81 + 6 * 108)
    lib().vo_tables(_p(out))
    o = 0
    res = {}
    for name, size in (("P1", 3), ("W1", 3)):
        res[name] = out[o:o + size]
        o += size
    res["P2"] = out[o:o + 18].reshape(9, 2)
    o += 18
    res["W2"] = out[o:o + 9]
    o += 9
    for name in ("bfn", "dfdc", "dfde", "d2fdc2", "d2fde2", "d2fdce"):
        res[name] = out[o:o + 81].reshape(9, 9)  # [gp][node]
        o += 81
    for name in ("bfnE", "dfdcE", "dfdeE", "d2fdc2E", "d2fde2E", "d2fdceE"):
        res[name] = out[o:o + 108].reshape(4, 3, 9)
        o += 108
    return res


def basis2d(xy, gp, edge=0):
    xy = np.ascontiguousarray(xy, dtype=np.float64)
    out = np.zeros(38)
    lib().vo_basis2d(_p(xy), gp, edge, _p(out))
    return dict(detJ=out[0], WET=out[1], dFdXi=out[2:20].reshape(2, 9), d2FdXi2=out[20:38].reshape(2, 9))


def gauss_fields(xy, TL, gp):
    xy = np.ascontiguousarray(xy, dtype=np.float64)
    TL = np.ascontiguousarray(TL, dtype=np.float64)
    out = np.zeros(22)
    lib().vo_gauss_fields(_p(xy), _p(TL), gp, _p(out))
    return dict(U=out[0:2], GU=out[2:6].reshape(2, 2), G2U=out[6:10].reshape(2, 2), P=out[10], dP=out[11:13],
                S=out[13:16], dS=out[16:22].reshape(3, 2))


def point_strong_form(prm, U, GU, G2U, P, dP, S, dS, EU):
    """Strong-form quantities at a point from explicit field values.

    GU[j][i] = dU_i/dx_j, G2U[j][i] = d2U_i/dx_j^2, S = (Sxx, Sxy, Syy),
    dS[s][d] = dS_s/dx_d."""
    a = [np.ascontiguousarray(v, dtype=np.float64) for v in (U, GU, G2U, dP, S, dS)]
    out = np.zeros(40)
    lib().vo_point_strong_form(ctypes.byref(prm), _p(a[0]), _p(a[1]), _p(a[2]), float(P), _p(a[3]), _p(a[4]),
                               _p(a[5]), float(EU), _p(out))
    return dict(Con_EQ=out[0:4].reshape(2, 2), Momentum=out[4:6], Continuity=out[6], T_ve=out[7:16].reshape(3, 3),
                P_tot=out[16:25].reshape(3, 3), W12=out[25], UCD=out[26:30].reshape(2, 2), f_T=out[30],
                C=out[31:35].reshape(2, 2))


def element(prm, xy, TL, EU, ied=0, jac=True):
    """Res[9][6], A[54][54] (A[(inod,ieq)][(jnod,jeq)]), Ac[54] of one element."""
    xy = np.ascontiguousarray(xy, dtype=np.float64)
    TL = np.ascontiguousarray(TL, dtype=np.float64)
    Res = np.zeros((9, 6))
    A = np.zeros((54, 54)) if jac else None
    Ac = np.zeros(54) if jac else None
    lib().vo_element(ctypes.byref(prm), _p(xy), _p(TL), float(EU), ied, 1 if jac else 0, _p(Res), _p(A), _p(Ac))
    return Res, A, Ac


def count_flops(prm, xy, TL, EU, ied=0):
    out = np.zeros(5, dtype=np.uint64)
    xy = np.ascontiguousarray(xy, dtype=np.float64)
    TL = np.ascontiguousarray(TL, dtype=np.float64)
    lib(flops=True).vo_count_flops(ctypes.byref(prm), _p(xy), _p(TL), float(EU), ied, _p(out))
    return dict(add=int(out[0]), mul=int(out[1]), div=int(out[2]), sqrt=int(out[3]), total=int(out[4]))


def count_flops_geometry(prm, xy):
    """Operations of the geometry-only part of one residualBulk evaluation (all
    9 Gauss points: basis2D, Helem**2, BvN/H2, DfN/H2)."""
    out = np.zeros(5, dtype=np.uint64)
    xy = np.ascontiguousarray(xy, dtype=np.float64)
    lib(flops=True).vo_count_flops_geometry(ctypes.byref(prm), _p(xy), _p(out))
    return dict(add=int(out[0]), mul=int(out[1]), div=int(out[2]), sqrt=int(out[3]), total=int(out[4]))


def pattern(mesh):
    conn = np.ascontiguousarray(mesh.conn, dtype=np.int32)
    nnz = lib().vo_pattern_nnz(mesh.n_nodes, mesh.n_elements, _p(conn))
    rowptr = np.zeros(mesh.n_unknowns + 1, dtype=np.int64)
    col = np.zeros(nnz, dtype=np.int32)
    lib().vo_pattern(mesh.n_nodes, mesh.n_elements, _p(conn), _p(rowptr), _p(col))
    return rowptr, col


def assemble(prm, mesh, x, jac=True, rowptr=None, col=None, el_range=None, nthreads=0):
    """Reference-order assembly: returns (b[N], values[nnz])."""
    if rowptr is None:
        rowptr, col = pattern(mesh)
    coords = np.ascontiguousarray(mesh.coords, dtype=np.float64)
    conn = np.ascontiguousarray(mesh.conn, dtype=np.int32)
    edges = np.ascontiguousarray(mesh.outlet_edges, dtype=np.int32).reshape(-1, 2)
    x = np.ascontiguousarray(x, dtype=np.float64)
    N = mesh.n_unknowns
    b = np.zeros(N)
    vals = np.zeros(int(rowptr[-1])) if jac else None
    e0, e1 = (0, mesh.n_elements) if el_range is None else el_range
    rc = lib().vo_assemble(ctypes.byref(prm), mesh.n_nodes, mesh.n_elements, _p(coords), _p(conn), edges.shape[0],
                           _p(edges), _p(x[:-1]), float(x[-1]), 1 if jac else 0, _p(rowptr), _p(col), _p(b), _p(vals),
                           e0, e1, nthreads)
    if rc != 0:
        raise RuntimeError("oracle assembly failed: pattern mismatch")
    return b, vals


def arclength(mode, n_nodes, x, xo, t, increment, LN0, dArN, dSo, rowptr, col, vals, b, jac=True):
    """Applies the arclength row to (vals, b) in place (reference order)."""
    x, xo, t = (np.ascontiguousarray(a, dtype=np.float64) for a in (x, xo, t))
    lib().vo_arclength(mode, n_nodes, _p(x), _p(xo), _p(t), increment, LN0, dArN, dSo, 1 if jac else 0, _p(rowptr),
                       _p(col), _p(vals), _p(b))


def drag(prm, mesh, x, edges):
    """Drag term over the given boundary edges (element, edge id 1..4), in the
    reference's edge order; returns (Fd, per-edge terms)."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    edges = np.ascontiguousarray(edges, dtype=np.int32).reshape(-1, 2)
    Fd = np.zeros(1)
    per = np.zeros(max(len(edges), 1))
    rc = lib().vo_drag(ctypes.byref(prm), mesh.n_nodes, mesh.n_elements, _p(np.ascontiguousarray(mesh.coords)),
                       _p(np.ascontiguousarray(mesh.conn)), _p(x), len(edges), _p(edges), _p(Fd), _p(per))
    if rc != 0:
        raise ValueError("vo_drag failed")
    return float(Fd[0]), per[:len(edges)]


def nodal_stress(BvN, WiN, n_nodes, x):
    """[n_nodes][4] = T(1,1), T(1,2), T(2,2), T(3,3) of the SQRT formulation."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    out = np.zeros((n_nodes, 4))
    lib().vo_nodal_stress(float(BvN), float(WiN), n_nodes, _p(x), _p(out))
    return out


# ---------------------------------------------------------------------------
# Linear stability assembly (Cylinder_Stability_2D, vef_oracle.cpp's f3 part).
# ---------------------------------------------------------------------------
def gauss_tau(prm, xy, TL, EU, gp):
    """Stabilization quantities at bulk Gauss point gp of one element."""
    xy = np.ascontiguousarray(xy, dtype=np.float64)
    TL = np.ascontiguousarray(TL, dtype=np.float64)
    out = np.zeros(6)
    if lib().vo_gauss_tau(ctypes.byref(prm), _p(xy), _p(TL), float(EU), int(gp), _p(out)) != 0:
        raise ValueError(f"gauss_tau: Gauss point {gp} out of range")
    return dict(H2=out[0], tsugn=out[1], tlsme=out[2], tlsic=out[3], tlsce=out[4], tdcs=out[5])


def tau_plant(which: int):
    """Plants a known mistake in the tau coefficients (0 removes it)."""
    if lib().vo_tau_plant(int(which)) != 0:
        raise ValueError(f"tau_plant: no plant {which}")


def node_terms(prm, xy, TL, EU, gp):
    """The summands of the element residual's test-node terms at bulk Gauss
    point gp, as the oracle adds them (vo_node_terms): mom[n][i][t] (t:
    ReN*Convection*phi, P_tot(i,:).grad phi, LSME, LSIC, LSCE), cont[n][t]
    (Continuity*phi, LSME), str[n][s][t] (s = xx, xy, yy; t: Con_EQ*phi, DfN
    term, LSME, LSCE, DCS), WET and tdcs*|Con_EQ|."""
    xy = np.ascontiguousarray(xy, dtype=np.float64)
    TL = np.ascontiguousarray(TL, dtype=np.float64)
    out = np.zeros(245)
    if lib().vo_node_terms(ctypes.byref(prm), _p(xy), _p(TL), float(EU), int(gp), _p(out)) != 0:
        raise ValueError(f"node_terms: Gauss point {gp} out of range")
    t = out[:243].reshape(9, 27)
    return dict(mom=t[:, :10].reshape(9, 2, 5), cont=t[:, 10:12], str=t[:, 12:].reshape(9, 3, 5), WET=out[243],
                dcs_coef=out[244])


def term_plant(which: int):
    """Plants a known mistake in the test-node terms (0 removes it)."""
    if lib().vo_term_plant(int(which)) != 0:
        raise ValueError(f"term_plant: no plant {which}")


def _stab_rc(rc, what):
    if rc == -2:
        raise ValueError(f"{what}: the stability assembly needs DfN > 1e-12 (discontinuity capturing off)")
    if rc != 0:
        raise RuntimeError(f"{what} failed ({rc})")


def stab_plant(which: int):
    """Plants a known mistake in the stability restatement (0 removes it)."""
    lib().vo_stab_plant(int(which))


def stab_element(prm, xy, TLb, EU):
    """A, B [54 columns (inod,ieq)][54 rows (jnod,jeq)] of storeBulkStability."""
    xy = np.ascontiguousarray(xy, dtype=np.float64)
    TLb = np.ascontiguousarray(TLb, dtype=np.float64)
    A = np.zeros((54, 54))
    B = np.zeros((54, 54))
    _stab_rc(lib().vo_stab_element(ctypes.byref(prm), _p(xy), _p(TLb), float(EU), _p(A), _p(B)), "vo_stab_element")
    return A, B


def stab_residual(prm, xy, TLb, EU, TLd):
    """Res_a[9][6], Res_b[9][6] of residualBulkStability for the disturbance TLd."""
    xy, TLb, TLd = (np.ascontiguousarray(a, dtype=np.float64) for a in (xy, TLb, TLd))
    RA = np.zeros((9, 6))
    RB = np.zeros((9, 6))
    _stab_rc(lib().vo_stab_residual(ctypes.byref(prm), _p(xy), _p(TLb), float(EU), _p(TLd), _p(RA), _p(RB)),
             "vo_stab_residual")
    return RA, RB


def stab_linearization(prm, xy, TLb, EU, TLd, gp, h):
    """(stability test-node terms, central difference of the frozen-coefficient
    base terms) at bulk Gauss point gp, each [9][6]."""
    xy, TLb, TLd = (np.ascontiguousarray(a, dtype=np.float64) for a in (xy, TLb, TLd))
    st = np.zeros((9, 6))
    fd = np.zeros((9, 6))
    _stab_rc(lib().vo_stab_linearization(ctypes.byref(prm), _p(xy), _p(TLb), float(EU), _p(TLd), int(gp), float(h),
                                         _p(st), _p(fd)), "vo_stab_linearization")
    return st, fd


def stab_pattern(mesh):
    conn = np.ascontiguousarray(mesh.conn, dtype=np.int32)
    nnz = lib().vo_stab_pattern_nnz(mesh.n_nodes, mesh.n_elements, _p(conn))
    rowptr = np.zeros(mesh.n_nodes * 6 + 1, dtype=np.int64)
    col = np.zeros(nnz, dtype=np.int32)
    lib().vo_stab_pattern(mesh.n_nodes, mesh.n_elements, _p(conn), _p(rowptr), _p(col))
    return rowptr, col


def stab_assemble(prm, mesh, x, rowptr=None, col=None, el_range=None, nthreads=0):
    """setEquationsBulkStability in the reference's order: (A values, B values)."""
    if rowptr is None:
        rowptr, col = stab_pattern(mesh)
    coords = np.ascontiguousarray(mesh.coords, dtype=np.float64)
    conn = np.ascontiguousarray(mesh.conn, dtype=np.int32)
    x = np.ascontiguousarray(x, dtype=np.float64)
    A = np.zeros(int(rowptr[-1]))
    B = np.zeros(int(rowptr[-1]))
    e0, e1 = (0, mesh.n_elements) if el_range is None else el_range
    _stab_rc(lib().vo_stab_assemble(ctypes.byref(prm), mesh.n_nodes, mesh.n_elements, _p(coords), _p(conn), _p(x),
                                    _p(rowptr), _p(col), _p(A), _p(B), e0, e1, nthreads), "vo_stab_assemble")
    return A, B


def stab_dirichlet(rowptr, col, A, B, rows):
    """applyDirichletBCStability (src/Dirichlet_mod.F90:287-316): MatZeroRows
    of A with diagonal 1 and of B with diagonal 0, in place."""
    for r in np.unique(np.asarray(rows, dtype=np.int64)):
        s, e = rowptr[r], rowptr[r + 1]
        A[s:e] = 0.0
        B[s:e] = 0.0
        A[s:e][col[s:e] == r] = 1.0


def count_flops_stability(prm, xy, TLb, EU, TLd):
    out = np.zeros(10, dtype=np.uint64)
    xy, TLb, TLd = (np.ascontiguousarray(a, dtype=np.float64) for a in (xy, TLb, TLd))
    lib(flops=True).vo_count_flops_stability(ctypes.byref(prm), _p(xy), _p(TLb), float(EU), _p(TLd), _p(out))
    keys = ("add", "mul", "div", "sqrt", "total")
    return dict(zip(keys, map(int, out[:5]))), dict(zip(keys, map(int, out[5:])))
