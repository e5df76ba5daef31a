"""Symbolic analysis of the multifrontal LU (nested dissection, fronts,
buckets, maps), checked on the CPU: every unknown is the pivot of exactly one
front, every stored entry of the node block lands in exactly one front
position, each child's update rows are rows of its parent, and a numpy/scipy
execution of the plan (scatter, extend-add, partial LU per front, forward and
backward solves: what libvef_mf.so runs) solves Newton matrices of the
cylinder problem to scipy's accuracy."""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as sla

from newton_cpu import apply_dirichlet, dirichlet_rows
from ve_flow_around_cylinder_fortran_amd import cylinder_channel, smooth_state
from ve_flow_around_cylinder_fortran_amd.multifrontal import (build_plan, buckets_by_size, nested_dissection,
                                                                 numpy_factor_solve, value_map)
from ve_flow_around_cylinder_fortran_amd.newton import cylinder_dirichlet


def newton_matrix(oracle, mesh, EU=0.3):
    prm = oracle.params()
    rowptr, col = oracle.pattern(mesh)
    x = smooth_state(mesh, EU=EU)
    b, v = oracle.assemble(prm, mesh, x, rowptr=rowptr, col=col)
    r, vv = dirichlet_rows(*cylinder_dirichlet(mesh, EU), len(b))
    apply_dirichlet(r, vv, x, b, v, rowptr, col)
    return rowptr, col, v


def test_dissection_tree_properties():
    mesh = cylinder_channel(8, 4, 6, 8)
    fronts, order, pos = nested_dissection(mesh.coords, mesh.conn, leaf=4)
    assert np.array_equal(np.sort(order), np.arange(mesh.n_nodes))
    for t, f in enumerate(fronts):
        # pivots consecutive and last in the subtree's range; updates after it
        assert np.array_equal(pos[f.nodes], np.arange(f.last - len(f.nodes) + 1, f.last + 1))
        assert np.all(pos[f.update] > f.last) and np.all(np.diff(pos[f.update]) > 0)
        for c in f.children:
            assert fronts[c].parent == t and fronts[c].height < f.height
            # a child's update nodes are pivots or update nodes of its parent
            assert np.all(np.isin(fronts[c].update, np.concatenate([f.nodes, f.update])))
        # every element touching a pivot node belongs to the front's element set
        touching = np.nonzero(np.isin(mesh.conn, f.nodes).any(axis=1))[0]
        assert np.all(np.isin(touching, f.elements))


def test_buckets_cover_fronts_once():
    rng = np.random.default_rng(1)
    s = rng.integers(0, 300, 200)
    u = rng.integers(0, 900, 200)
    got = buckets_by_size(s, u)
    members = np.concatenate([np.array(m) for m, _, _ in got])
    assert np.array_equal(np.sort(members), np.arange(200))
    for m, s_max, u_max in got:
        assert s_max == s[m].max() and u_max == u[m].max()


@pytest.mark.parametrize("dims", [(8, 4, 6, 8), (16, 12, 20, 32)])
def test_plan_solves_newton_matrix(oracle, dims):
    mesh = cylinder_channel(*dims)
    rowptr, col, v = newton_matrix(oracle, mesh)
    plan = build_plan(mesh.coords, mesh.conn)
    n1 = len(rowptr) - 2
    # every unknown is the pivot of one front
    piv = np.concatenate([plan.rows[plan.rows_off[f]:plan.rows_off[f] + plan.front_s[f]] for f in range(plan.n_fronts)])
    assert np.array_equal(np.sort(piv), np.arange(n1))
    # every node-block entry maps to one distinct front position; the border to none
    dst = value_map(plan, rowptr, col)
    rows = np.repeat(np.arange(n1 + 1), np.diff(rowptr))
    inner = (rows < n1) & (col < n1)
    assert np.all(dst[~inner] == -1) and np.all(dst[inner] >= 0) and np.all(dst[inner] < plan.total)
    assert len(np.unique(dst[inner])) == int(inner.sum())
    # the plan's execution solves the node block
    A = sp.csr_matrix((v, col, rowptr), shape=(n1 + 1, n1 + 1))[:n1, :n1].tocsc()
    b = np.random.default_rng(0).standard_normal(n1)
    x = numpy_factor_solve(plan, dst, v, b)
    ref = sla.spsolve(A, b)
    assert np.abs(x - ref).max() <= 1e-10 * np.abs(ref).max()
    assert plan.flops_padded <= 1.5 * plan.flops
