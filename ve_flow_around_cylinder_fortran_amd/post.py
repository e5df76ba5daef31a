"""Post-processing of a converged state and the solution file format.

* Drag: `Assembler.drag` (GPU) integrates dot(-n, P_tot(:,1)) over the
  cylinder edges (calculateEdgeQuantities, src/Post_Process_mod.F90:1796-1890);
  `cylinder_drag` here sets the edges and returns Fd as a float.
* Wall stress: `Assembler.nodal_stress` (GPU) gives the nodal stress tensor
  (nodalStressTensorSQRT, :2176-2236); `txx_profile` selects the nodes the
  reference writes to its Txx file (calculateNodalQuantities, :1992-2168) and
  `peak_wall_stress` the maximum Txx on the cylinder.
* Solution files: `write_solution` / `read_solution` use the reference's
  `.DTA` layout (writeSolution / readSolution, :1271-1424): a Fortran
  unformatted sequential file with one record per value, element by element,
  local node by local node, equation by equation (so shared nodes repeat),
  then the continuation unknowns.
"""
from __future__ import annotations

import math

import numpy as np

from .mesh import MESH_TOLERANCE, Mesh, boundary_edges

# ----------------------------------------------------------------------------
# Drag and wall stress
# ----------------------------------------------------------------------------


def cylinder_drag(asm, x) -> float:
    """Fd on the cylinder (the value the reference prints as 'Fd = ')."""
    if getattr(asm, "_drag_boundary", None) != "Cylinder":
        asm.set_drag_edges(boundary_edges(asm.mesh, "Cylinder"))
        asm._drag_boundary = "Cylinder"
    return float(asm.drag(x).item())


def drag_coefficient(Fd: float, ReN: float) -> float:
    """drag_coef of calculateEdgeQuantities: (2*Fd/(2*ReN + 1e-8))/2."""
    c = 2.0 * Fd / (2.0 * ReN + 1.0e-8)
    return c / 2.0


def _wall_masks(mesh: Mesh):
    R = mesh.geometry["R"]
    x, y = mesh.coords[:, 0], mesh.coords[:, 1]
    eps = MESH_TOLERANCE
    # distanceToExtraBND (src/Post_Process_mod.F90:2322-2349), ibnd = 1 and 2
    on_cyl = (y >= 0.0) & (np.abs(x ** 2 + y ** 2 - R ** 2) < eps)
    on_line = (np.abs(y - 0.0) < eps) & (x >= R) & (x <= 8.0 * R)
    return on_cyl, on_line


def txx_profile(mesh: Mesh, T: np.ndarray) -> np.ndarray:
    """Rows (L_arc, Txx) the reference writes to Results/.../Txx/*.dat, sorted
    by L_arc: the upper cylinder surface with L_arc = atan2(y, x)*R, except
    L_arc = 0 at x = -R and pi*R at x = +R (as the reference sets them), then
    the wake centreline R <= x <= 8R with L_arc = pi*R + (x - R)."""
    R = mesh.geometry["R"]
    eps = MESH_TOLERANCE
    T = np.asarray(T)
    on_cyl, on_line = _wall_masks(mesh)
    rows = []
    for g in range(mesh.n_nodes):
        xx, yy = mesh.coords[g]
        if on_cyl[g]:
            L = math.atan2(yy, xx) * R
            if abs(xx + R) < eps:
                L = 0.0
            if abs(xx - R) < eps:
                L = math.pi * R
            rows.append((L, T[g, 0]))
        if on_line[g]:
            rows.append((math.pi * R + (xx - R), T[g, 0]))
    out = np.array(rows, dtype=np.float64).reshape(-1, 2)
    return out[np.argsort(out[:, 0], kind="stable")]


def peak_wall_stress(mesh: Mesh, T: np.ndarray) -> float:
    """Maximum Txx over the upper cylinder surface nodes."""
    on_cyl, _ = _wall_masks(mesh)
    return float(np.asarray(T)[on_cyl, 0].max())


def write_txx_file(path: str, profile: np.ndarray) -> None:
    """The reference's '(*(es16.8))' lines."""
    with open(path, "w") as f:
        for L, t in profile:
            f.write(f"{_es16_8(L)}{_es16_8(t)}\n")


def _es16_8(v: float) -> str:
    """Fortran ES16.8: one digit before the point, 8 after, exponent E+dd."""
    m, e = f"{v:.8E}".split("E")
    return f"{m}E{int(e):+03d}".rjust(16)


# ----------------------------------------------------------------------------
# Solution files (.DTA, Fortran unformatted sequential, 4-byte record markers)
# ----------------------------------------------------------------------------
_MARK = np.array([8], dtype="<i4")


def write_solution(path: str, mesh: Mesh, x: np.ndarray) -> None:
    """writeSolution: TL(gnod, ieq) for every element's local nodes, then EU."""
    x = np.asarray(x, dtype=np.float64)
    TL = x[:-1].reshape(mesh.n_nodes, 6)
    vals = np.concatenate([TL[mesh.conn.astype(np.int64)].reshape(-1), x[-1:]])
    rec = np.empty((vals.size, 4), dtype="<i4")
    rec[:, 0] = 8
    rec[:, 3] = 8
    rec[:, 1:3] = vals.astype("<f8").view("<i4").reshape(-1, 2)
    with open(path, "wb") as f:
        f.write(rec.tobytes())


def read_solution(path: str, mesh: Mesh) -> np.ndarray:
    """readSolution: the global state (the last element's copy of a shared node
    wins, as the reference's loop overwrites in element order)."""
    raw = np.fromfile(path, dtype="<i4")
    if raw.size % 4 != 0:
        raise ValueError("not a sequence of 8-byte records")
    rec = raw.reshape(-1, 4)
    if not (np.all(rec[:, 0] == 8) and np.all(rec[:, 3] == 8)):
        raise ValueError("record markers are not 8")
    vals = np.ascontiguousarray(rec[:, 1:3]).view("<f8").reshape(-1)
    n_tl = mesh.n_elements * 9 * 6
    if vals.size < n_tl + 1:
        raise ValueError("file shorter than the mesh's element-node values plus EU")
    TL = np.zeros((mesh.n_nodes, 6))
    TL[mesh.conn.astype(np.int64).reshape(-1)] = vals[:n_tl].reshape(-1, 6)
    x = np.empty(mesh.n_unknowns)
    x[:-1] = TL.reshape(-1)
    x[-1] = vals[n_tl]
    return x
