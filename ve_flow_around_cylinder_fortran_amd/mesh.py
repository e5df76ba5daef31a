"""Body-fitted 9-node quadrilateral meshes of a cylinder in a channel.

The reference reads its meshes from ``<name>_MESH.DTA`` files produced offline
by SALOME (``mesh.py``/``mesh2.py`` of the reference, not runnable without
SALOME).  This module generates the same kind of mesh directly:

* a cylinder of radius ``R`` centred at the origin in the channel
  ``x in [-L_up, L_down]``, ``y in [-H, H]`` (blockage ``R/H``);
* an O-grid of four sectors between the circle and the square
  ``[-H, H]^2`` plus an upstream and a downstream block, all structured;
* 9-node (biquadratic) elements whose local node order is the reference's
  (corners (-1,-1) (1,-1) (1,1) (-1,1), mid-edges bottom right top left,
  centre; ``src/GaussIntegration/Gauss_Quadrangle_mod.F90:413-421``), with
  counter-clockwise orientation so that the edge normals of
  ``setEdgeQuantities`` (``src/Element_Calculations_mod.F90:887-905``) point
  outwards;
* reverse Cuthill-McKee node numbering and elements sorted by their lowest
  node, so that neighbouring elements touch nearby memory.

Boundaries are found geometrically with the reference's tolerance
(``Mesh_Tolerance = 1e-6``, ``src/Parameters_mod.F90:226``), and the outlet
edges are listed as (element, edge id 1..4) in ascending element order, the
order ``setBoundaryEdges`` produces (``src/Boundary_mod.F90:476-531``).

``write_dta``/``read_dta`` use the reference's ``.DTA`` layout
(``src/FEM_mod.F90:172-256``): a line ``Nnodes Nel``, node lines
``id x y z``, element lines ``id type n1 .. nk`` with ``type = 100*dim +
nodes`` and the top-dimension elements last; ids are 1-based.
"""
from __future__ import annotations

import dataclasses
import math
from typing import Dict

import numpy as np

MESH_TOLERANCE = 1.0e-6

# Local nodes of each edge id (1: eta=-1, 2: xi=+1, 3: eta=+1, 4: xi=-1), 0-based.
EDGE_NODES = {1: (0, 1, 4), 2: (1, 2, 5), 3: (2, 3, 6), 4: (3, 0, 7)}


@dataclasses.dataclass
class Mesh:
    coords: np.ndarray  # (n_nodes, 2) float64
    conn: np.ndarray  # (n_elements, 9) int32, 0-based
    outlet_edges: np.ndarray  # (n_edges, 2) int32: (element, edge id 1..4)
    boundary_nodes: Dict[str, np.ndarray]  # name -> sorted 0-based node ids
    geometry: Dict[str, float]

    @property
    def n_nodes(self) -> int:
        return int(self.coords.shape[0])

    @property
    def n_elements(self) -> int:
        return int(self.conn.shape[0])

    @property
    def n_unknowns(self) -> int:
        """N = 6*Nnodes + 1 (the arclength unknown last)."""
        return 6 * self.n_nodes + 1


def _geometric(n: int, ratio: float) -> np.ndarray:
    """n+1 points on [0,1]; last cell / first cell = ratio."""
    if n == 1 or abs(ratio - 1.0) < 1e-12:
        return np.linspace(0.0, 1.0, n + 1)
    q = ratio ** (1.0 / (n - 1))
    w = q ** np.arange(n)
    t = np.concatenate([[0.0], np.cumsum(w)])
    return t / t[-1]


def _q2_params(t: np.ndarray) -> np.ndarray:
    """Element-corner parameters -> 2n+1 node parameters (mid nodes halfway)."""
    out = np.empty(2 * len(t) - 1)
    out[0::2] = t
    out[1::2] = 0.5 * (t[:-1] + t[1:])
    return out


def _block_elements(nx: int, ny: int, node_id: np.ndarray) -> np.ndarray:
    """Connectivity of an nx-by-ny block of Q2 elements on a (2nx+1, 2ny+1) node grid."""
    conn = np.empty((nx * ny, 9), dtype=np.int64)
    k = 0
    for j in range(ny):
        for i in range(nx):
            a, b = 2 * i, 2 * j
            conn[k] = [
                node_id[a, b], node_id[a + 2, b], node_id[a + 2, b + 2], node_id[a, b + 2],
                node_id[a + 1, b], node_id[a + 2, b + 1], node_id[a + 1, b + 2], node_id[a, b + 1],
                node_id[a + 1, b + 1],
            ]
            k += 1
    return conn


def cylinder_channel(n_t: int = 16, n_r: int = 12, n_up: int = 20, n_down: int = 32, *, R: float = 1.0,
                     H: float = 2.0, L_up: float = 15.0, L_down: float = 15.0, radial_ratio: float = 6.0,
                     axial_ratio: float = 8.0, reorder: bool = True) -> Mesh:
    """Structured body-fitted Q2 mesh; n_t*(4*n_r + n_up + n_down) elements.

    n_t: elements along each side of the square around the cylinder (a
    quarter of the circle); n_r: radial elements between circle and square;
    n_up/n_down: elements along x upstream/downstream of the square.
    """
    if H <= R:
        raise ValueError("the channel half-width must exceed the radius")
    pts = []
    blocks = []

    def add_block(X, Y):
        blocks.append((X, Y))

    s = _q2_params(np.linspace(-1.0, 1.0, n_t + 1))  # along a square side, counter-clockwise
    t = _q2_params(_geometric(n_r, radial_ratio))  # circle (0) -> square (1)
    corners = [(H, -H), (H, H), (-H, H), (-H, -H), (H, -H)]
    for sec in range(4):
        q0 = np.array(corners[sec])
        q1 = np.array(corners[sec + 1])
        Q = q0[None, :] + 0.5 * (s[:, None] + 1.0) * (q1 - q0)[None, :]  # (2n_t+1, 2)
        C = R * Q / np.linalg.norm(Q, axis=1, keepdims=True)
        # xi: radial (outwards), eta: along the side (counter-clockwise) -> positive orientation
        X = C[None, :, 0] + t[:, None] * (Q[None, :, 0] - C[None, :, 0])
        Y = C[None, :, 1] + t[:, None] * (Q[None, :, 1] - C[None, :, 1])
        X[-1, :] = Q[:, 0]
        Y[-1, :] = Q[:, 1]
        add_block(X, Y)
    y = H * s
    xu = -L_up + (L_up - H) * (1.0 - _q2_params(_geometric(n_up, axial_ratio))[::-1])
    xd = H + (L_down - H) * _q2_params(_geometric(n_down, axial_ratio))
    add_block(np.repeat(xu[:, None], len(y), 1), np.repeat(y[None, :], len(xu), 0))
    add_block(np.repeat(xd[:, None], len(y), 1), np.repeat(y[None, :], len(xd), 0))

    # Merge coincident nodes between blocks.
    key_to_id: Dict[tuple, int] = {}
    conn_all = []
    scale = 1.0e9
    for X, Y in blocks:
        nid = np.empty(X.shape, dtype=np.int64)
        for a in range(X.shape[0]):
            for b in range(X.shape[1]):
                k = (int(round(X[a, b] * scale)), int(round(Y[a, b] * scale)))
                if k not in key_to_id:
                    key_to_id[k] = len(pts)
                    pts.append((X[a, b], Y[a, b]))
                nid[a, b] = key_to_id[k]
        conn_all.append(_block_elements((X.shape[0] - 1) // 2, (X.shape[1] - 1) // 2, nid))
    coords = np.asarray(pts, dtype=np.float64)
    conn = np.concatenate(conn_all).astype(np.int64)

    if reorder:
        coords, conn = rcm_reorder(coords, conn)
    geometry = dict(R=R, H=H, L_up=L_up, L_down=L_down, n_t=n_t, n_r=n_r, n_up=n_up, n_down=n_down)
    return finish_mesh(coords, conn, geometry)


def finish_mesh(coords: np.ndarray, conn: np.ndarray, geometry: Dict[str, float]) -> Mesh:
    """Boundary node sets and outlet edges from the geometry."""
    R, H = geometry["R"], geometry["H"]
    L_up, L_down = geometry["L_up"], geometry["L_down"]
    x, y = coords[:, 0], coords[:, 1]
    tol = MESH_TOLERANCE
    bnd = {
        "Inlet": np.nonzero(np.abs(x + L_up) < tol)[0],
        "Outlet": np.nonzero(np.abs(x - L_down) < tol)[0],
        "Wall": np.nonzero(np.abs(np.abs(y) - H) < tol)[0],
        "Cylinder": np.nonzero(np.abs(np.hypot(x, y) - R) < tol)[0],
    }
    on_outlet = np.abs(x - L_down) < tol
    hit = np.stack([np.all(on_outlet[conn[:, list(EDGE_NODES[ied])]], axis=1) for ied in (1, 2, 3, 4)], axis=1)
    el, ed = np.nonzero(hit)  # row-major: ascending element, then edge id
    outlet = np.stack([el, ed + 1], axis=1).astype(np.int32).reshape(-1, 2)
    return Mesh(coords=np.ascontiguousarray(coords, dtype=np.float64),
                conn=np.ascontiguousarray(conn, dtype=np.int32), outlet_edges=outlet,
                boundary_nodes={k: v.astype(np.int64) for k, v in bnd.items()}, geometry=dict(geometry))


def boundary_edges(mesh: Mesh, name: str) -> np.ndarray:
    """Edges (element, edge id 1..4) whose three nodes lie on boundary `name`,
    ascending in element then edge id — the order of setBoundaryEdges
    (src/Boundary_mod.F90:477-522) — with the distance tests of
    distanceToBNDCylinder (:1383-1414): Inlet |x + L|, Outlet |x - L|,
    Wall |y -+ H|, Cylinder |x^2 + y^2 - R^2|, each below the mesh tolerance."""
    g = mesh.geometry
    x, y = mesh.coords[:, 0], mesh.coords[:, 1]
    tol = MESH_TOLERANCE
    if name == "Inlet":
        on = np.abs(x + g["L_up"]) < tol
    elif name == "Outlet":
        on = np.abs(x - g["L_down"]) < tol
    elif name == "Wall":
        on = (np.abs(y + g["H"]) < tol) | (np.abs(y - g["H"]) < tol)
    elif name == "Cylinder":
        on = np.abs(x * x + y * y - g["R"] * g["R"]) < tol
    else:
        raise ValueError(f"unknown boundary {name!r}")
    if name == "Wall":  # the two walls are separate boundaries; an edge lies on one of them
        lo = np.abs(y + g["H"]) < tol
        hi = np.abs(y - g["H"]) < tol
        hit = np.stack([np.all(lo[mesh.conn[:, list(EDGE_NODES[ied])]], axis=1)
                        | np.all(hi[mesh.conn[:, list(EDGE_NODES[ied])]], axis=1) for ied in (1, 2, 3, 4)], axis=1)
    else:
        hit = np.stack([np.all(on[mesh.conn[:, list(EDGE_NODES[ied])]], axis=1) for ied in (1, 2, 3, 4)], axis=1)
    el, ed = np.nonzero(hit)
    return np.stack([el, ed + 1], axis=1).astype(np.int32).reshape(-1, 2)


def rcm_reorder(coords: np.ndarray, conn: np.ndarray):
    """Reverse Cuthill-McKee node numbering; elements sorted by lowest node."""
    import scipy.sparse as sp
    from scipy.sparse.csgraph import reverse_cuthill_mckee

    n = coords.shape[0]
    rows = np.repeat(conn, 9, axis=1).ravel()
    cols = np.tile(conn, (1, 9)).ravel()
    A = sp.csr_matrix((np.ones_like(rows, dtype=np.int8), (rows, cols)), shape=(n, n))
    perm = reverse_cuthill_mckee(A, symmetric_mode=True)  # new -> old
    inv = np.empty(n, dtype=np.int64)
    inv[perm] = np.arange(n)
    conn2 = inv[conn]
    order = np.lexsort((conn2.max(axis=1), conn2.min(axis=1)))
    return coords[perm], conn2[order]


def mesh_for_elements(target: int, **kw) -> Mesh:
    """The channel mesh family scaled to about `target` elements.

    Proportions n_t : n_r : n_up : n_down = 16 : 12 : 20 : 32 (1,600 elements
    at scale 1); exact sizes used by the workloads are chosen explicitly in
    `workload_mesh`.
    """
    s = math.sqrt(target / 1600.0)
    return cylinder_channel(max(2, round(16 * s)), max(2, round(12 * s)), max(2, round(20 * s)),
                            max(2, round(32 * s)), **kw)


# Block sizes of the named workloads: n_t*(4*n_r + n_up + n_down) elements.
WORKLOAD_DIMS = {
    1600: (16, 12, 20, 32),
    40000: (80, 60, 100, 160),
    100000: (125, 94, 157, 267),
}


def workload_mesh(n_elements: int) -> Mesh:
    if n_elements in WORKLOAD_DIMS:
        return cylinder_channel(*WORKLOAD_DIMS[n_elements])
    return mesh_for_elements(n_elements)


# ----------------------------------------------------------------------------
# .DTA format (src/FEM_mod.F90:172-256)
# ----------------------------------------------------------------------------
def write_dta(mesh: Mesh, path: str) -> None:
    with open(path, "w") as f:
        f.write(f"{mesh.n_nodes} {mesh.n_elements}\n")
        for i, (xx, yy) in enumerate(mesh.coords):
            f.write(f"{i + 1} {xx:.17g} {yy:.17g} 0\n")
        for e, row in enumerate(mesh.conn):
            f.write(f"{e + 1} 209 " + " ".join(str(int(v) + 1) for v in row) + "\n")


def read_dta(path: str, geometry: Dict[str, float] | None = None) -> Mesh:
    """Reads a 2D Q2 mesh; like setConnectivityArrays, keeps the elements of the
    top dimension (type 2xx), which the format lists last."""
    with open(path) as f:
        n_nodes, n_total = (int(v) for v in f.readline().split()[:2])
        coords = np.empty((n_nodes, 2))
        for i in range(n_nodes):
            parts = f.readline().split()
            coords[i] = float(parts[1]), float(parts[2])
        rows = []
        for _ in range(n_total):
            parts = f.readline().split()
            etype = int(parts[1])
            nn = etype % 100
            if etype // 100 == 2:
                if nn != 9:
                    raise ValueError(f"element type {etype}: only 9-node quadrangles (209) are supported")
                rows.append([int(v) - 1 for v in parts[2:2 + nn]])
    conn = np.asarray(rows, dtype=np.int64)
    if geometry is None:
        geometry = dict(R=1.0, H=float(np.abs(coords[:, 1]).max()), L_up=float(-coords[:, 0].min()),
                        L_down=float(coords[:, 0].max()))
    return finish_mesh(coords, conn, geometry)
