"""Host-side mirror of the reference's assembly interface over the C ABI.

The reference assembles once per Newton iteration in assemblePetscObjects
(src/Newton_Raphson_mod.F90:281-400):

    initializeLSObjects(Flag_NR, LS)                 ! zero b (and A if NRP)
    setEquationsBulk(..., Sol, basis_p, Flag_NR, Elem, LS)
    setEquationsEdge(..., Boundary(Outlet), ..., Flag_NR, Elem, LS)
    ... arclength row, Dirichlet rows ...

`Assembler.assemble(x, flag_nr)` performs the first three steps on the GPU
through vef_assemble (include/vef.h) and returns the residual b and the CSR
matrix values; `Assembler.rowptr`/`colind` are the preallocated pattern of
preAllocatePetsc (src/Petsc_mod.F90:629-743).  Flag_NR keeps the reference's
spelling: 'NRP' assembles residual and Jacobian, 'MNR' the residual only.

There is no CPU fallback: constructing an Assembler without the built HIP
library raises.
"""
from __future__ import annotations

import ctypes
import dataclasses
import os
from typing import Optional, Tuple

import numpy as np

from .mesh import Mesh

_LIB_NAME = "libvef.so"
_lib = None

VEF_MODEL_LPTT = 0
VEF_MODEL_GIESEKUS = 1
VEF_RES_ONLY = 0
VEF_RES_JAC = 1
VEF_RES_JAC_FP32 = 2
MODELS = {"L-PTT": VEF_MODEL_LPTT, "GIESEKUS": VEF_MODEL_GIESEKUS}


class VefError(RuntimeError):
    pass


class _Params(ctypes.Structure):
    _fields_ = [("ReN", ctypes.c_double), ("BvN", ctypes.c_double), ("VeN", ctypes.c_double),
                ("DfN", ctypes.c_double), ("WiN0", ctypes.c_double), ("model", ctypes.c_int32),
                ("reserved", ctypes.c_int32)]


@dataclasses.dataclass
class Params:
    """Dimensionless numbers (src/Parameters_mod.F90:98-111).

    Model_Name 'L-PTT' with VeN = 0 is Oldroyd-B; 'GIESEKUS' uses VeN as the
    mobility.  DfN defaults to the reference's DfN0 = 1e-3.
    """
    ReN: float = 0.0
    BvN: float = 0.59
    VeN: float = 0.0
    DfN: float = 1.0e-3
    WiN0: float = 1.0
    model: str = "L-PTT"

    def to_c(self) -> _Params:
        if self.model not in MODELS:
            raise VefError(f"Model_Name {self.model!r} is not supported (L-PTT, GIESEKUS)")
        return _Params(self.ReN, self.BvN, self.VeN, self.DfN, self.WiN0, MODELS[self.model], 0)


def library_path() -> str:
    """The in-tree libvef.so, or the file VEF_LIBRARY names."""
    return os.environ.get("VEF_LIBRARY") or os.path.join(os.path.dirname(os.path.abspath(__file__)), _LIB_NAME)


def load_library():
    """Loads the HIP assembler library; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    path = library_path()
    if not os.path.exists(path):
        raise VefError(f"{path} not found: build it with __graft_entry__.build() "
                       "(python -c 'import __graft_entry__ as g; g.build()')")
    lib = ctypes.CDLL(path)
    c_i32, c_i64, c_dp, c_ip, c_vp = ctypes.c_int32, ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p
    lib.vef_create.argtypes = [c_i32, c_i32, c_dp, c_ip, c_i32, c_ip, ctypes.POINTER(_Params),
                               ctypes.POINTER(ctypes.c_void_p)]
    lib.vef_create.restype = ctypes.c_int
    lib.vef_set_params.argtypes = [c_vp, ctypes.POINTER(_Params)]
    lib.vef_set_params.restype = ctypes.c_int
    lib.vef_num_rows.argtypes = [c_vp]
    lib.vef_num_rows.restype = c_i64
    lib.vef_nnz.argtypes = [c_vp]
    lib.vef_nnz.restype = c_i64
    lib.vef_get_pattern.argtypes = [c_vp, c_vp, c_vp]
    lib.vef_get_pattern.restype = ctypes.c_int
    lib.vef_device_rowptr.argtypes = [c_vp]
    lib.vef_device_rowptr.restype = c_vp
    lib.vef_device_colind.argtypes = [c_vp]
    lib.vef_device_colind.restype = c_vp
    lib.vef_assemble.argtypes = [c_vp, c_vp, c_i32, c_vp, c_vp, c_vp]
    lib.vef_assemble.restype = ctypes.c_int
    lib.vef_arclength.argtypes = [c_vp, c_i32, c_vp, c_vp, c_vp, c_i32, ctypes.c_double, ctypes.c_double,
                                  ctypes.c_double, c_i32, c_vp, c_vp, c_vp]
    lib.vef_arclength.restype = ctypes.c_int
    lib.vef_set_dirichlet.argtypes = [c_vp, c_i64, c_vp, c_vp]
    lib.vef_set_dirichlet.restype = ctypes.c_int
    lib.vef_apply_dirichlet.argtypes = [c_vp, c_vp, c_i32, c_vp, c_vp, c_vp]
    lib.vef_apply_dirichlet.restype = ctypes.c_int
    lib.vef_set_drag_edges.argtypes = [c_vp, c_i32, c_ip]
    lib.vef_set_drag_edges.restype = ctypes.c_int
    lib.vef_drag.argtypes = [c_vp, c_vp, c_vp, c_vp, c_vp]
    lib.vef_drag.restype = ctypes.c_int
    lib.vef_nodal_stress.argtypes = [c_vp, c_vp, ctypes.c_double, c_vp, c_vp]
    lib.vef_nodal_stress.restype = ctypes.c_int
    lib.vef_assemble_host.argtypes = [c_vp, c_vp, c_i32, c_vp, c_vp]
    lib.vef_stability_nnz.argtypes = [c_vp]
    lib.vef_stability_nnz.restype = c_i64
    lib.vef_get_stability_pattern.argtypes = [c_vp, c_vp, c_vp]
    lib.vef_get_stability_pattern.restype = ctypes.c_int
    lib.vef_set_stability_dirichlet.argtypes = [c_vp, c_i64, c_vp]
    lib.vef_set_stability_dirichlet.restype = ctypes.c_int
    lib.vef_stability_assemble.argtypes = [c_vp, c_vp, c_vp, c_vp, c_vp]
    lib.vef_stability_assemble.restype = ctypes.c_int
    lib.vef_assemble_host.restype = ctypes.c_int
    lib.vef_profile.argtypes = [c_vp, c_i32]
    lib.vef_profile.restype = ctypes.c_int
    lib.vef_profile_read.argtypes = [c_vp, c_vp, c_vp, c_vp]
    lib.vef_profile_read.restype = ctypes.c_int
    lib.vef_last_error.argtypes = [c_vp]
    lib.vef_last_error.restype = ctypes.c_char_p
    lib.vef_destroy.argtypes = [c_vp]
    lib.vef_destroy.restype = None
    _lib = lib
    return lib


def _ptr(a: np.ndarray):
    return ctypes.c_void_p(a.ctypes.data)


class Assembler:
    """GPU assembler for one mesh (vef_create .. vef_destroy)."""

    def __init__(self, mesh: Mesh, params: Params):
        self.lib = load_library()
        self.mesh = mesh
        self.params = params
        coords = np.ascontiguousarray(mesh.coords, dtype=np.float64)
        conn = np.ascontiguousarray(mesh.conn, dtype=np.int32)
        edges = np.ascontiguousarray(mesh.outlet_edges, dtype=np.int32).reshape(-1, 2)
        h = ctypes.c_void_p()
        cp = params.to_c()
        rc = self.lib.vef_create(mesh.n_nodes, mesh.n_elements, _ptr(coords), _ptr(conn), edges.shape[0],
                                 _ptr(edges) if edges.size else None, ctypes.byref(cp), ctypes.byref(h))
        if rc != 0:
            raise VefError(f"vef_create failed ({rc}): {self.lib.vef_last_error(None).decode()}")
        self._h = h
        self.N = int(self.lib.vef_num_rows(h))
        self.nnz = int(self.lib.vef_nnz(h))
        self._rowptr = None
        self._colind = None

    def _check(self, rc: int, what: str):
        if rc != 0:
            raise VefError(f"{what} failed ({rc}): {self.lib.vef_last_error(self._h).decode()}")

    def close(self):
        if getattr(self, "_h", None):
            self.lib.vef_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_params(self, params: Params):
        cp = params.to_c()
        self._check(self.lib.vef_set_params(self._h, ctypes.byref(cp)), "vef_set_params")
        self.params = params

    def pattern(self) -> Tuple[np.ndarray, np.ndarray]:
        if self._rowptr is None:
            rowptr = np.empty(self.N + 1, dtype=np.int64)
            colind = np.empty(self.nnz, dtype=np.int32)
            self._check(self.lib.vef_get_pattern(self._h, _ptr(rowptr), _ptr(colind)), "vef_get_pattern")
            self._rowptr, self._colind = rowptr, colind
        return self._rowptr, self._colind

    @property
    def rowptr(self) -> np.ndarray:
        return self.pattern()[0]

    @property
    def colind(self) -> np.ndarray:
        return self.pattern()[1]

    # ------------------------------------------------------------------
    def assemble(self, x, flag_nr: str = "NRP", b=None, values=None, stream=None):
        """One assembly on the GPU with torch tensors (float64, cuda).

        x: [N] state; returns (b[N], values[nnz] or None).  Work is queued on
        `stream` (default: torch's current stream).
        """
        import torch

        flag = self._flag(flag_nr)
        if not (x.is_cuda and x.dtype == torch.float64 and x.is_contiguous() and x.numel() == self.N):
            raise VefError("x must be a contiguous float64 CUDA tensor of length N")
        dev = x.device
        if b is None:
            b = torch.empty(self.N, dtype=torch.float64, device=dev)
        if flag != VEF_RES_ONLY and values is None:
            values = torch.empty(self.nnz, dtype=torch.float64, device=dev)
        if stream is None:
            stream = torch.cuda.current_stream(dev)
        rc = self.lib.vef_assemble(self._h, ctypes.c_void_p(x.data_ptr()), flag, ctypes.c_void_p(b.data_ptr()),
                                   ctypes.c_void_p(values.data_ptr()) if values is not None else None,
                                   ctypes.c_void_p(stream.cuda_stream))
        self._check(rc, "vef_assemble")
        return b, (values if flag != VEF_RES_ONLY else None)

    def arclength(self, x, b, values=None, *, increment: int, LN0: float = 0.0, dArN: float = 0.0, x_old=None,
                  tangent=None, dS: float = 0.0, flag_nr: str = "NRP", stream=None):
        """Arclength row N-1 (src/Extra_Equations_mod.F90:698-828): pins lambda
        for increments 1-2, pseudo-arclength row from increment 3 on."""
        import torch

        flag = self._flag(flag_nr)  # the row is the same for either Jacobian
        mode = 0 if increment <= 2 else 1
        if stream is None:
            stream = torch.cuda.current_stream(x.device)
        ptr = lambda t: ctypes.c_void_p(t.data_ptr()) if t is not None else None  # noqa: E731
        rc = self.lib.vef_arclength(self._h, mode, ptr(x), ptr(x_old), ptr(tangent), int(increment), float(LN0),
                                    float(dArN), float(dS), flag, ptr(b), ptr(values),
                                    ctypes.c_void_p(stream.cuda_stream))
        self._check(rc, "vef_arclength")

    def set_dirichlet(self, rows: np.ndarray, values: np.ndarray):
        """Dirichlet rows (global 6*gnod + ieq) and their prescribed values; a
        row listed twice keeps its last value."""
        r = np.ascontiguousarray(rows, dtype=np.int64).reshape(-1)
        v = np.ascontiguousarray(values, dtype=np.float64).reshape(-1)
        if r.shape != v.shape:
            raise VefError("rows and values differ in length")
        self._check(self.lib.vef_set_dirichlet(self._h, len(r), _ptr(r), _ptr(v)), "vef_set_dirichlet")

    def apply_dirichlet(self, x, b, values=None, flag_nr: str = "NRP", stream=None):
        """b[row] = x[row] - value; with 'NRP' (or 'NRP32') the rows of values
        become unit rows."""
        import torch

        flag = self._flag(flag_nr)
        if stream is None:
            stream = torch.cuda.current_stream(x.device)
        rc = self.lib.vef_apply_dirichlet(self._h, ctypes.c_void_p(x.data_ptr()), flag, ctypes.c_void_p(b.data_ptr()),
                                          ctypes.c_void_p(values.data_ptr()) if values is not None else None,
                                          ctypes.c_void_p(stream.cuda_stream))
        self._check(rc, "vef_apply_dirichlet")

    def set_drag_edges(self, edges: np.ndarray):
        """Boundary edges (element, edge id 1..4) of the drag integral, in the
        reference's order (mesh.boundary_edges(mesh, "Cylinder"))."""
        e = np.ascontiguousarray(edges, dtype=np.int32).reshape(-1, 2)
        self._check(self.lib.vef_set_drag_edges(self._h, len(e), _ptr(e)), "vef_set_drag_edges")
        self._n_drag = len(e)

    def drag(self, x, per_edge: bool = False, stream=None):
        """Fd over the drag edges (calculateEdgeQuantities, ieq_ed = 1) as a
        one-element device tensor; with per_edge, also each edge's term."""
        import torch

        if stream is None:
            stream = torch.cuda.current_stream(x.device)
        Fd = torch.empty(1, dtype=torch.float64, device=x.device)
        per = torch.empty(max(getattr(self, "_n_drag", 0), 1), dtype=torch.float64, device=x.device) \
            if per_edge else None
        rc = self.lib.vef_drag(self._h, ctypes.c_void_p(x.data_ptr()), ctypes.c_void_p(Fd.data_ptr()),
                               ctypes.c_void_p(per.data_ptr()) if per is not None else None,
                               ctypes.c_void_p(stream.cuda_stream))
        self._check(rc, "vef_drag")
        if per_edge:
            return Fd, per[:getattr(self, "_n_drag", 0)]
        return Fd

    def nodal_stress(self, x, WiN: float, stream=None):
        """[n_nodes][4] device tensor (Txx, Txy, Tyy, Tzz) of the square-root
        formulation (nodalStressTensorSQRT) at the given WiN."""
        import torch

        if stream is None:
            stream = torch.cuda.current_stream(x.device)
        T = torch.empty((self.mesh.n_nodes, 4), dtype=torch.float64, device=x.device)
        rc = self.lib.vef_nodal_stress(self._h, ctypes.c_void_p(x.data_ptr()), float(WiN),
                                       ctypes.c_void_p(T.data_ptr()), ctypes.c_void_p(stream.cuda_stream))
        self._check(rc, "vef_nodal_stress")
        return T

    # -- linear stability (Cylinder_Stability_2D) ---------------------------
    def stability_pattern(self) -> Tuple[np.ndarray, np.ndarray]:
        """CSR pattern of A and B (the node blocks; rows/cols 6*g + q)."""
        nnz = int(self.lib.vef_stability_nnz(self._h))
        rowptr = np.empty(self.N, dtype=np.int64)  # 6*n_nodes + 1
        col = np.empty(nnz, dtype=np.int32)
        self._check(self.lib.vef_get_stability_pattern(self._h, _ptr(rowptr), _ptr(col)), "vef_get_stability_pattern")
        return rowptr, col

    def set_stability_dirichlet(self, rows: np.ndarray):
        """Rows (6*g + q) of applyDirichletBCStability."""
        rows = np.ascontiguousarray(rows, dtype=np.int64)
        self._check(self.lib.vef_set_stability_dirichlet(self._h, rows.size, _ptr(rows)),
                    "vef_set_stability_dirichlet")

    def stability(self, x, A=None, B=None, stream=None):
        """A and B values (device float64 tensors, stability pattern) for the
        base state x (device tensor, length N: TL node-major, EU last)."""
        import torch

        if stream is None:
            stream = torch.cuda.current_stream(x.device)
        nnz = int(self.lib.vef_stability_nnz(self._h))
        if A is None:
            A = torch.empty(nnz, dtype=torch.float64, device=x.device)
        if B is None:
            B = torch.empty(nnz, dtype=torch.float64, device=x.device)
        if x.numel() != self.N or A.numel() != nnz or B.numel() != nnz:
            raise VefError("x must have length N, A and B the stability nnz")
        rc = self.lib.vef_stability_assemble(self._h, ctypes.c_void_p(x.data_ptr()), ctypes.c_void_p(A.data_ptr()),
                                             ctypes.c_void_p(B.data_ptr()), ctypes.c_void_p(stream.cuda_stream))
        self._check(rc, "vef_stability_assemble")
        return A, B

    def assemble_host(self, x: np.ndarray, flag_nr: str = "NRP"):
        """The same with numpy arrays (copies through the host)."""
        flag = self._flag(flag_nr)
        x = np.ascontiguousarray(x, dtype=np.float64)
        if x.shape != (self.N,):
            raise VefError("x must have length N")
        b = np.empty(self.N)
        values = np.empty(self.nnz) if flag != VEF_RES_ONLY else None
        rc = self.lib.vef_assemble_host(self._h, _ptr(x), flag, _ptr(b), _ptr(values) if values is not None else None)
        self._check(rc, "vef_assemble_host")
        return b, values

    def profile(self, enable: bool = True):
        self._check(self.lib.vef_profile(self._h, 1 if enable else 0), "vef_profile")

    def profile_read(self):
        """(bulk kernel ms, scatter ms, assemblies) summed since the last read."""
        tb, ts, n = ctypes.c_double(), ctypes.c_double(), ctypes.c_int64()
        self._check(self.lib.vef_profile_read(self._h, ctypes.byref(tb), ctypes.byref(ts), ctypes.byref(n)),
                    "vef_profile_read")
        return tb.value, ts.value, n.value

    @staticmethod
    def _flag(flag_nr: str) -> int:
        if flag_nr == "NRP":
            return VEF_RES_JAC
        if flag_nr == "MNR":
            return VEF_RES_ONLY
        if flag_nr == "NRP32":  # FP32 analytic Jacobian (no counterpart in the reference)
            return VEF_RES_JAC_FP32
        raise VefError(f"Flag_NR must be 'NRP', 'MNR' or 'NRP32', got {flag_nr!r}")
