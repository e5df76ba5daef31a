"""The C-ABI library: it loads, exports every function include/vef.h declares,
and rejects invalid arguments before touching the GPU (src/Equations_mod.F90
stops on invalid enums; the C ABI returns an error code instead)."""
import ctypes
import os
import re

import numpy as np
import pytest

from ve_flow_around_cylinder_fortran_amd import Params, VefError, cylinder_channel
from ve_flow_around_cylinder_fortran_amd import assembly

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions(header="vef.h"):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(vef_[a-z_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = assembly.load_library()
    names = declared_functions()
    assert len(names) >= 12
    for n in names:
        assert hasattr(lib, n), n


def test_mlp_library_exports_every_declared_symbol():
    from ve_flow_around_cylinder_fortran_amd import learned

    lib = learned.load_library()
    names = declared_functions("vef_mlp.h")
    assert len(names) == 9
    for n in names:
        assert hasattr(lib, n), n


def test_band_library_exports_every_declared_symbol():
    from ve_flow_around_cylinder_fortran_amd import linsolve

    lib = linsolve._band_lib()
    names = declared_functions("vef_band.h")
    assert len(names) == 9
    for n in names:
        assert hasattr(lib, n), n


def test_band_library_rejects_bad_arguments_without_a_gpu():
    """Argument checks come before any device call (block size other than 64,
    empty systems, NULL handles)."""
    from ve_flow_around_cylinder_fortran_amd import linsolve

    lib = linsolve._band_lib()
    h = ctypes.c_void_p()
    assert lib.vef_band_create(100, 10, 32, ctypes.byref(h)) != 0 and not h.value
    assert lib.vef_band_create(0, 10, 64, ctypes.byref(h)) != 0 and not h.value
    assert lib.vef_band_create_prec(100, 10, 64, 16, ctypes.byref(h)) != 0 and not h.value
    assert lib.vef_band_factor(None, None) != 0
    assert lib.vef_band_uses_graph(None) == 0
    assert lib.vef_band_solve(None, None, None, None) != 0
    assert lib.vef_band_set(None, 0, None, None, None, None) != 0


def test_mf_library_exports_every_declared_symbol():
    from ve_flow_around_cylinder_fortran_amd import multifrontal

    lib = multifrontal._mf_lib()
    names = declared_functions("vef_mf.h")
    assert len(names) == 14
    for n in names:
        assert hasattr(lib, n), n


def test_mf_profile_calls_reject_null_handles():
    """The profiling entry points check their arguments on the host."""
    from ve_flow_around_cylinder_fortran_amd import multifrontal

    lib = multifrontal._mf_lib()
    assert lib.vef_mf_profile(None, 1) != 0
    assert lib.vef_mf_heights(None) == -1
    ms = np.zeros(4, dtype=np.float32)
    assert lib.vef_mf_profile_read(None, ms.ctypes.data) != 0
    assert lib.vef_mf_set_pivoting(None, 0) != 0


def test_mf_library_rejects_bad_plans_without_a_gpu():
    """The plan is checked on the host before any device allocation: front
    sizes against their bucket, parents above their children, rows in range,
    every unknown the pivot of exactly one front; NULL handles."""
    from ve_flow_around_cylinder_fortran_amd import multifrontal

    lib = multifrontal._mf_lib()
    i32 = lambda *a: np.array(a, dtype=np.int32)  # noqa: E731
    i64 = lambda *a: np.array(a, dtype=np.int64)  # noqa: E731

    def plan(front_s=(6, 6), rows=None, parent=(1, -1), slot=(0, 0), bucket_h=(0, 1), n=12):
        # two fronts in two buckets: a leaf (6 pivots, 6 updates) under a root (6 pivots)
        arrs = dict(bh=i32(*bucket_h), bc=i32(1, 1), bs=i32(6, 6), bu=i32(6, 0), fs=i32(*front_s),
                    fu=i32(6, 0), ro=i64(0, 12, 18), rows=i32(*(rows or list(range(12)) + list(range(6, 12)))),
                    par=i32(*parent), slot=i32(*slot), eo=i64(0, 6, 6), em=i32(*range(6)))
        vp = lambda a: ctypes.c_void_p(a.ctypes.data)  # noqa: E731
        p = multifrontal._Plan(n, 2, vp(arrs["bh"]), vp(arrs["bc"]), vp(arrs["bs"]), vp(arrs["bu"]), 2,
                               vp(arrs["fs"]), vp(arrs["fu"]), vp(arrs["ro"]), vp(arrs["rows"]), vp(arrs["par"]),
                               vp(arrs["slot"]), vp(arrs["eo"]), vp(arrs["em"]))
        return p, arrs

    def create(p):
        h = ctypes.c_void_p()
        rc = lib.vef_mf_create(ctypes.byref(p), 64, ctypes.byref(h))
        msg = lib.vef_mf_last_error(h).decode() if h.value else ""
        lib.vef_mf_destroy(h)
        return rc, msg

    p, keep = plan(bucket_h=(1, 0))
    assert create(p) == (-1, "plan: buckets not ordered by height")
    p, keep = plan(front_s=(12, 6))
    assert create(p) == (-1, "plan: front larger than its bucket")
    p, keep = plan(rows=list(range(11)) + [12] + list(range(6, 12)))
    assert create(p) == (-1, "plan: row out of range")
    p, keep = plan(rows=list(range(12)) + list(range(0, 6)))
    assert create(p) == (-1, "plan: an unknown is a pivot of two fronts")
    p, keep = plan(parent=(0, -1))
    assert create(p) == (-1, "plan: parent not above its child")
    p, keep = plan()
    h = ctypes.c_void_p()
    assert lib.vef_mf_create(ctypes.byref(p), 16, ctypes.byref(h)) != 0 and not h.value
    assert lib.vef_mf_factor(None, None) != 0
    assert lib.vef_mf_solve(None, None, None, None) != 0
    assert lib.vef_mf_set(None, 0, None, None, None) != 0
    assert lib.vef_mf_zero_pivots(None, None) != 0


def test_every_header_is_checked():
    assert sorted(os.listdir(os.path.join(ROOT, "include"))) == ["vef.h", "vef_band.h", "vef_mf.h", "vef_mlp.h"]


def test_fortran_binding_names_match_header():
    """The Fortran interface module binds exactly names the header declares."""
    src = open(os.path.join(ROOT, "ve_flow_around_cylinder_fortran_amd", "fortran", "vef_assembly_mod.f90")).read()
    bound = set(re.findall(r"bind\(C, name='(vef_[a-z_]+)'\)", src))
    assert bound and bound <= set(declared_functions())


def _create(mesh, params, conn=None, edges=None):
    lib = assembly.load_library()
    coords = np.ascontiguousarray(mesh.coords)
    conn = np.ascontiguousarray(mesh.conn if conn is None else conn, dtype=np.int32)
    edges = np.ascontiguousarray(mesh.outlet_edges if edges is None else edges, dtype=np.int32)
    h = ctypes.c_void_p()
    cp = params.to_c()
    rc = lib.vef_create(mesh.n_nodes, conn.shape[0], coords.ctypes.data, conn.ctypes.data, edges.shape[0],
                        edges.ctypes.data, ctypes.byref(cp), ctypes.byref(h))
    return rc, lib.vef_last_error(None).decode()


def test_create_rejects_bad_meshes_and_params():
    m = cylinder_channel(2, 1, 1, 2)
    bad = m.conn.copy()
    bad[0, 0] = m.n_nodes
    assert _create(m, Params(), conn=bad) == (2, "connectivity out of range")
    rep = m.conn.copy()
    rep[1, 3] = rep[1, 2]
    assert _create(m, Params(), conn=rep)[0] == 2
    edges = m.outlet_edges.copy()
    edges[0, 1] = 5
    assert _create(m, Params(), edges=edges)[0] == 2
    rc, msg = _create(m, Params(model="L-PTT", WiN0=0.0))
    assert rc == 1 and "WiN0" in msg
    p = Params().to_c()
    p.model = 7
    lib = assembly.load_library()
    h = ctypes.c_void_p()
    conn = np.ascontiguousarray(m.conn, dtype=np.int32)
    rc = lib.vef_create(m.n_nodes, m.n_elements, np.ascontiguousarray(m.coords).ctypes.data, conn.ctypes.data, 0,
                        None, ctypes.byref(p), ctypes.byref(h))
    assert rc == 4


def test_unknown_model_and_flag_names():
    with pytest.raises(VefError):
        Params(model="e-PTT").to_c()
    with pytest.raises(VefError):
        assembly.Assembler._flag("XYZ")
    assert assembly.Assembler._flag("NRP") == 1 and assembly.Assembler._flag("MNR") == 0
