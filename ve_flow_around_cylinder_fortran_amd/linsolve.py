"""Sparse LU solves on the GPU for the Newton counterpart.

The reference factorises each Newton matrix with MUMPS through PETSc
(solveLinearSystem, src/Petsc_mod.F90) and, in modified-Newton ('MNR')
iterations, reuses the last factorisation.  Here the pattern is analysed once
on the host: a symmetric fill-reducing order Q (minimum degree on A + A^T,
SuperLU through scipy) and the structural pattern of T = (L - I) + U of
Q^T A Q with diagonal pivots, obtained by factorising a matrix of that pattern
with random, diagonally dominant values (so no entry of the fill cancels).
Every 'NRP' matrix is then factorised on the GPU in that pattern
(rocSOLVER csrrf_refactlu, static diagonal pivoting) and the solves run on the
GPU (csrrf_solve).  Each factorisation is checked by the residual of its first
solve; if static pivoting has lost accuracy, that system is solved on the host
with threshold partial pivoting instead, and the event is counted.

`precision="f32"` factorises and solves in single precision (the FP64 values
rounded once per factorisation; right-hand sides rounded on the way in, the
solution widened on the way out): the linear solve of an inexact Newton whose
residual stays FP64.  Its check tolerance is correspondingly looser.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_rocblas = None
_rocsolver = None


def _libs():
    global _rocblas, _rocsolver
    if _rocsolver is None:
        rocm = os.environ.get("ROCM_PATH", "/opt/rocm")
        _rocblas = ctypes.CDLL(os.path.join(rocm, "lib", "librocblas.so"))
        _rocsolver = ctypes.CDLL(os.path.join(rocm, "lib", "librocsolver.so"))
        vp, i32 = ctypes.c_void_p, ctypes.c_int32
        _rocblas.rocblas_create_handle.argtypes = [ctypes.POINTER(vp)]
        _rocblas.rocblas_destroy_handle.argtypes = [vp]
        _rocblas.rocblas_set_stream.argtypes = [vp, vp]
        _rocsolver.rocsolver_create_rfinfo.argtypes = [ctypes.POINTER(vp), vp]
        _rocsolver.rocsolver_destroy_rfinfo.argtypes = [vp]
        for t in "sd":
            getattr(_rocsolver, f"rocsolver_{t}csrrf_analysis").argtypes = [vp, i32, i32, i32, vp, vp, vp, i32, vp, vp,
                                                                            vp, vp, vp, vp, i32, vp]
            getattr(_rocsolver, f"rocsolver_{t}csrrf_refactlu").argtypes = [vp, i32, i32, vp, vp, vp, i32, vp, vp, vp,
                                                                            vp, vp, vp]
            getattr(_rocsolver, f"rocsolver_{t}csrrf_solve").argtypes = [vp, i32, i32, i32, vp, vp, vp, vp, vp, vp,
                                                                         i32, vp]
    return _rocblas, _rocsolver


class RfLUError(RuntimeError):
    pass


def _ok(status: int, what: str):
    if status != 0:
        raise RfLUError(f"{what} returned rocblas_status {status}")


def _bundle(L, U, n):
    """T = (L - I) + U in CSR with sorted columns, built from the stored entries
    (the strictly lower entries of L, then the row of U), so no stored position
    is dropped even where its value is 0."""
    import scipy.sparse as sp

    L.sort_indices()
    U.sort_indices()
    rl = np.repeat(np.arange(n), np.diff(L.indptr))
    keep = L.indices < rl
    ru = np.repeat(np.arange(n), np.diff(U.indptr))
    rows = np.concatenate([rl[keep], ru])
    cols = np.concatenate([L.indices[keep], U.indices])
    vals = np.concatenate([L.data[keep], U.data])
    order = np.lexsort((cols, rows))
    indptr = np.zeros(n + 1, dtype=np.int64)
    np.cumsum(np.bincount(rows, minlength=n), out=indptr[1:])
    return sp.csr_matrix((vals[order], cols[order], indptr), shape=(n, n))


class RfLU:
    """LU of one sparsity pattern on the GPU by re-factorisation."""

    def __init__(self, rowptr: np.ndarray, col: np.ndarray, device="cuda:0", check_tol: float = None,
                 precision: str = "f64"):
        import torch

        if precision not in ("f64", "f32"):
            raise RfLUError(f"precision {precision!r}: expected 'f64' or 'f32'")
        self.torch = torch
        self.precision = precision
        self.dtype = torch.float64 if precision == "f64" else torch.float32
        self._t = "d" if precision == "f64" else "s"
        if check_tol is None:
            check_tol = 1e-10 if precision == "f64" else 1e-3
        self.n = len(rowptr) - 1
        if rowptr[-1] >= 2 ** 31:
            raise RfLUError("rocSOLVER's csrrf takes 32-bit indices: pattern too large")
        self.device = device
        self.rowptr_h = np.asarray(rowptr, dtype=np.int64)
        self.col_h = np.asarray(col, dtype=np.int32)
        self.ptrA = torch.from_numpy(self.rowptr_h.astype(np.int32)).to(device)
        self.indA = torch.from_numpy(self.col_h).to(device)
        self.check_tol = check_tol
        rb, rs = _libs()
        self.handle = ctypes.c_void_p()
        _ok(rb.rocblas_create_handle(ctypes.byref(self.handle)), "rocblas_create_handle")
        self.rfinfo = None
        self.analyses = 0
        self.refactorizations = 0
        self.host_fallbacks = 0
        self._A = None
        self._checked = False
        self._host_lu = None

    def close(self):
        rb, rs = _libs()
        if self.rfinfo is not None:
            rs.rocsolver_destroy_rfinfo(self.rfinfo)
            self.rfinfo = None
        if self.handle:
            rb.rocblas_destroy_handle(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _stream(self):
        rb, _ = _libs()
        s = self.torch.cuda.current_stream(self.device)
        _ok(rb.rocblas_set_stream(self.handle, ctypes.c_void_p(s.cuda_stream)), "rocblas_set_stream")

    def _analyse(self):
        """Symmetric order and structural pattern of the factors (host, once)."""
        import scipy.sparse as sp
        import scipy.sparse.linalg as sla

        rb, rs = _libs()
        torch = self.torch
        rng = np.random.default_rng(12345)
        vals = rng.uniform(1.0, 2.0, size=len(self.col_h))
        A = sp.csr_matrix((vals, self.col_h, self.rowptr_h), shape=(self.n, self.n))
        rowsum = np.asarray(abs(A).sum(axis=1)).ravel()
        diag = np.repeat(np.arange(self.n), np.diff(self.rowptr_h)) == self.col_h
        vals[diag] = rowsum[np.repeat(np.arange(self.n), np.diff(self.rowptr_h))[diag]] + 1.0
        A = sp.csr_matrix((vals, self.col_h, self.rowptr_h), shape=(self.n, self.n))
        lu = sla.splu(A.tocsc(), permc_spec="MMD_AT_PLUS_A", options=dict(DiagPivotThresh=0.0, SymmetricMode=True))
        if not np.array_equal(lu.perm_r, lu.perm_c):
            raise RfLUError("the structural factorisation did not keep diagonal pivots")
        T = _bundle(lu.L.tocsr(), lu.U.tocsr(), self.n)
        if T.nnz >= 2 ** 31:
            raise RfLUError("L+U too large for 32-bit indices")
        self.ptrT = torch.from_numpy(T.indptr.astype(np.int32)).to(self.device)
        self.indT = torch.from_numpy(T.indices.astype(np.int32)).to(self.device)
        self.valT = torch.from_numpy(T.data).to(device=self.device, dtype=self.dtype)
        perm = np.argsort(lu.perm_c).astype(np.int32)
        self.pivP = torch.from_numpy(perm).to(self.device)
        self.pivQ = torch.from_numpy(perm.copy()).to(self.device)
        self.nnzT = int(T.nnz)
        M = torch.from_numpy(vals).to(device=self.device, dtype=self.dtype)
        self.rfinfo = ctypes.c_void_p()
        _ok(rs.rocsolver_create_rfinfo(ctypes.byref(self.rfinfo), self.handle), "rocsolver_create_rfinfo")
        self._stream()
        p = lambda t: ctypes.c_void_p(t.data_ptr())  # noqa: E731
        # the solve phase needs an analysis with nrhs >= 1 and a right-hand side buffer
        rhs = torch.zeros(self.n, dtype=self.dtype, device=self.device)
        analysis = getattr(rs, f"rocsolver_{self._t}csrrf_analysis")
        _ok(analysis(self.handle, self.n, 1, int(self.rowptr_h[-1]), p(self.ptrA), p(self.indA), p(M), self.nnzT,
                     p(self.ptrT), p(self.indT), p(self.valT), p(self.pivP), p(self.pivQ), p(rhs), self.n,
                     self.rfinfo), f"rocsolver_{self._t}csrrf_analysis")
        torch.cuda.synchronize(self.device)
        self.analyses += 1

    def factor(self, values):
        """Factorises the matrix with the given device values (same pattern)."""
        _, rs = _libs()
        if self.rfinfo is None:
            self._analyse()
        self._A = values
        self._checked = False
        self._host_lu = None
        self._stream()
        p = lambda t: ctypes.c_void_p(t.data_ptr())  # noqa: E731
        vals = values if values.dtype == self.dtype else values.to(self.dtype)
        refactlu = getattr(rs, f"rocsolver_{self._t}csrrf_refactlu")
        _ok(refactlu(self.handle, self.n, int(self.rowptr_h[-1]), p(self.ptrA), p(self.indA), p(vals), self.nnzT,
                     p(self.ptrT), p(self.indT), p(self.valT), p(self.pivP), p(self.pivQ), self.rfinfo),
            f"rocsolver_{self._t}csrrf_refactlu")
        self.refactorizations += 1

    def _residual(self, values, x, b):
        torch = self.torch
        A = torch.sparse_csr_tensor(self.ptrA.to(torch.int64), self.indA.to(torch.int64), values,
                                    size=(self.n, self.n))
        r = torch.mv(A, x) - b
        return float(torch.linalg.vector_norm(r)), float(torch.linalg.vector_norm(b))

    def solve(self, b):
        """x with A x = b for the last factorised matrix (b is not modified).
        The first solve after a factorisation checks ||A x - b|| <= check_tol
        ||b||; a failed check moves that matrix's solves to the host."""
        _, rs = _libs()
        torch = self.torch
        if self.rfinfo is None or self._A is None:
            raise RfLUError("factor() first")
        if self._host_lu is not None:
            return torch.from_numpy(self._host_lu.solve(b.cpu().numpy())).to(self.device)
        x = b.to(self.dtype, copy=True)
        self._stream()
        p = lambda t: ctypes.c_void_p(t.data_ptr())  # noqa: E731
        solve = getattr(rs, f"rocsolver_{self._t}csrrf_solve")
        _ok(solve(self.handle, self.n, 1, self.nnzT, p(self.ptrT), p(self.indT), p(self.valT), p(self.pivP),
                  p(self.pivQ), p(x), self.n, self.rfinfo), f"rocsolver_{self._t}csrrf_solve")
        x = x.to(b.dtype)
        if not self._checked:
            self._checked = True
            rn, bn = self._residual(self._A, x, b)
            if not (rn <= self.check_tol * max(bn, 1e-300)):
                import scipy.sparse as sp
                import scipy.sparse.linalg as sla

                A = sp.csr_matrix((self._A.cpu().numpy(), self.col_h, self.rowptr_h), shape=(self.n, self.n))
                self._host_lu = sla.splu(A.tocsc(), permc_spec="MMD_AT_PLUS_A", options=dict(DiagPivotThresh=0.01))
                self.host_fallbacks += 1
                return torch.from_numpy(self._host_lu.solve(b.cpu().numpy())).to(self.device)
        return x
