"""Mesh generator and the reference's .DTA format (src/FEM_mod.F90:172-256)."""
import os

import numpy as np
import pytest

from ve_flow_around_cylinder_fortran_amd import cylinder_channel, read_dta, workload_mesh, write_dta
from ve_flow_around_cylinder_fortran_amd.mesh import EDGE_NODES, WORKLOAD_DIMS


@pytest.mark.parametrize("n", [1600])
def test_workload_sizes(n):
    m = workload_mesh(n)
    assert m.n_elements == n
    nt, nr, nu, nd = WORKLOAD_DIMS[n]
    assert n == nt * (4 * nr + nu + nd)


def test_mesh_validity(oracle):
    m = cylinder_channel(8, 6, 10, 16)
    # every node used; connectivity in range; all elements positively oriented
    assert np.array_equal(np.unique(m.conn), np.arange(m.n_nodes))
    for e in range(0, m.n_elements, 7):
        xy = m.coords[m.conn[e]]
        for g in range(9):
            assert oracle.basis2d(xy, g)["detJ"] > 0
    # area of the fluid domain (isoparametric quadrature; circle approximated by Q2 arcs)
    area = sum(oracle.basis2d(m.coords[m.conn[e]], g)["WET"] for e in range(m.n_elements) for g in range(9))
    assert area == pytest.approx(30 * 4 - np.pi, rel=2e-5)
    # boundary sets
    x, y = m.coords.T
    assert np.allclose(x[m.boundary_nodes["Outlet"]], 15.0)
    assert np.allclose(np.hypot(x, y)[m.boundary_nodes["Cylinder"]], 1.0)
    # outlet edges: ascending elements, edge nodes on x = 15, one edge per outlet element side
    assert len(m.outlet_edges) == 8
    assert np.all(np.diff(m.outlet_edges[:, 0]) > 0)
    for e, ied in m.outlet_edges:
        assert np.allclose(x[m.conn[e, list(EDGE_NODES[ied])]], 15.0)


def test_dta_round_trip(tmp_path):
    m = cylinder_channel(4, 2, 3, 4)
    path = os.path.join(tmp_path, "Cyl_MESH.DTA")
    write_dta(m, path)
    m2 = read_dta(path)
    assert np.array_equal(m2.conn, m.conn)
    assert np.array_equal(m2.coords, m.coords)  # %.17g round-trips doubles
    assert np.array_equal(m2.outlet_edges, m.outlet_edges)


def test_dta_reader_keeps_top_dimension_elements(tmp_path):
    """Lower-dimension elements (boundary lines, type 103) may precede the
    quadrangles; only type 2xx elements form the connectivity."""
    m = cylinder_channel(2, 1, 1, 2)
    path = os.path.join(tmp_path, "X_MESH.DTA")
    with open(path, "w") as f:
        f.write(f"{m.n_nodes} {m.n_elements + 1}\n")
        for i, (a, b) in enumerate(m.coords):
            f.write(f"{i + 1} {a:.17g} {b:.17g} 0\n")
        f.write("1 103 1 2 3\n")
        for e, row in enumerate(m.conn):
            f.write(f"{e + 2} 209 " + " ".join(str(v + 1) for v in row) + "\n")
    m2 = read_dta(path)
    assert np.array_equal(m2.conn, m.conn)
