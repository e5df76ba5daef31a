"""Sparse LU solves on the GPU for the Newton counterpart.

Three variants share one interface (factor(values), solve(b)): RfLU (scalar
re-factorisation and scalar triangular solves), BlockLU (6x6 node blocks for
both) and HybridLU (scalar re-factorisation in the node-block order, block
triangular solves; the Newton driver's default, the fastest of the three on
this problem: the scalar factorisation of rocSOLVER and the block solves of
rocSPARSE are each the faster of their pair).

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

    def solve(self, b, check: bool = False):
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


# ---------------------------------------------------------------------------
# Block LU in the 6x6 node blocks (rocSPARSE bsrilu0 + bsrsv)
# ---------------------------------------------------------------------------
_rocsparse = None


def _sparse_lib():
    global _rocsparse
    if _rocsparse is None:
        rocm = os.environ.get("ROCM_PATH", "/opt/rocm")
        L = ctypes.CDLL(os.path.join(rocm, "lib", "librocsparse.so"))
        vp, i32 = ctypes.c_void_p, ctypes.c_int32
        L.rocsparse_create_handle.argtypes = [ctypes.POINTER(vp)]
        L.rocsparse_destroy_handle.argtypes = [vp]
        L.rocsparse_set_stream.argtypes = [vp, vp]
        L.rocsparse_create_mat_descr.argtypes = [ctypes.POINTER(vp)]
        L.rocsparse_destroy_mat_descr.argtypes = [vp]
        L.rocsparse_set_mat_fill_mode.argtypes = [vp, i32]
        L.rocsparse_set_mat_diag_type.argtypes = [vp, i32]
        L.rocsparse_create_mat_info.argtypes = [ctypes.POINTER(vp)]
        L.rocsparse_destroy_mat_info.argtypes = [vp]
        L.rocsparse_bsrilu0_zero_pivot.argtypes = [vp, vp, ctypes.POINTER(i32)]
        for t in "sd":
            getattr(L, f"rocsparse_{t}bsrilu0_buffer_size").argtypes = [vp, i32, i32, i32, vp, vp, vp, vp, i32, vp,
                                                                        ctypes.POINTER(ctypes.c_size_t)]
            getattr(L, f"rocsparse_{t}bsrilu0_analysis").argtypes = [vp, i32, i32, i32, vp, vp, vp, vp, i32, vp, i32,
                                                                     i32, vp]
            getattr(L, f"rocsparse_{t}bsrilu0").argtypes = [vp, i32, i32, i32, vp, vp, vp, vp, i32, vp, i32, vp]
            getattr(L, f"rocsparse_{t}bsrsv_buffer_size").argtypes = [vp, i32, i32, i32, i32, vp, vp, vp, vp, i32, vp,
                                                                      ctypes.POINTER(ctypes.c_size_t)]
            getattr(L, f"rocsparse_{t}bsrsv_analysis").argtypes = [vp, i32, i32, i32, i32, vp, vp, vp, vp, i32, vp,
                                                                   i32, i32, vp]
            getattr(L, f"rocsparse_{t}bsrsv_solve").argtypes = [vp, i32, i32, i32, i32, vp, vp, vp, vp, vp, i32, vp, vp,
                                                                vp, i32, vp]
        _rocsparse = L
    return _rocsparse


_DIR_ROW, _OP_NONE, _ANALYSIS_FORCE, _SOLVE_AUTO = 0, 111, 1, 0
_FILL_LOWER, _FILL_UPPER, _DIAG_NON_UNIT, _DIAG_UNIT = 0, 1, 0, 1


def _sp_ok(status: int, what: str):
    if status != 0:
        raise RfLUError(f"{what} returned rocsparse_status {status}")


class BlockLU:
    """LU of one sparsity pattern on the GPU in 6x6 node blocks.

    The unknowns 6g..6g+5 of node g form one block row and column; the
    arclength unknown is the first unknown of one more block, padded with an
    identity.  The order is minimum degree on the node graph (A + A^T), the
    arclength block last; the block fill is that of a random diagonally
    dominant matrix of the same pattern (host, once).  Each factorisation
    scatters the CSR values into the block storage and runs rocSPARSE's
    bsrilu0, which on a pattern that already holds the fill is the exact LU
    with static diagonal pivots; the solves are two block triangular solves
    (bsrsv).  Row dependencies are resolved per node block rather than per
    unknown, which shortens the sequential chain of the elimination about
    six-fold.  The first solve after a factorisation is checked as in RfLU.
    """

    BS = 6

    def __init__(self, rowptr: np.ndarray, col: np.ndarray, device="cuda:0", check_tol: float = None,
                 precision: str = "f64"):
        import torch

        if precision not in ("f64", "f32"):
            raise RfLUError(f"precision {precision!r}: expected 'f64' or 'f32'")
        self.torch = torch
        self.device = device
        self.precision = precision
        self.dtype = torch.float64 if precision == "f64" else torch.float32
        self._t = "d" if precision == "f64" else "s"
        self.check_tol = (1e-10 if precision == "f64" else 1e-3) if check_tol is None else check_tol
        self.rowptr_h = np.asarray(rowptr, dtype=np.int64)
        self.col_h = np.asarray(col, dtype=np.int32)
        self.n = len(rowptr) - 1
        if (self.n - 1) % self.BS != 0:
            raise RfLUError("BlockLU expects 6 unknowns per node and one arclength unknown")
        self.nn = (self.n - 1) // self.BS
        self.nb = self.nn + 1
        self.ptrA = torch.from_numpy(self.rowptr_h).to(device)
        self.indA = torch.from_numpy(self.col_h.astype(np.int64)).to(device)
        self.handle = None
        self.analyses = 0
        self.refactorizations = 0
        self.host_fallbacks = 0
        self._A = None
        self._checked = False
        self._host_lu = None

    def close(self):
        L = _sparse_lib()
        for name, fn in (("info", L.rocsparse_destroy_mat_info), ("descr", L.rocsparse_destroy_mat_descr),
                         ("descrL", L.rocsparse_destroy_mat_descr), ("descrU", L.rocsparse_destroy_mat_descr),
                         ("handle", L.rocsparse_destroy_handle)):
            h = getattr(self, name, None)
            if h:
                fn(h)
                setattr(self, name, None)

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _stream(self):
        L = _sparse_lib()
        s = self.torch.cuda.current_stream(self.device)
        _sp_ok(L.rocsparse_set_stream(self.handle, ctypes.c_void_p(s.cuda_stream)), "rocsparse_set_stream")

    def _analyse(self):
        import scipy.sparse as sp
        import scipy.sparse.linalg as sla

        torch = self.torch
        n, nn, nb, BS = self.n, self.nn, self.nb, self.BS
        idx = np.arange(n)
        node = np.minimum(idx // BS, nn)
        local = np.where(idx < BS * nn, idx % BS, 0)
        rows = np.repeat(idx, np.diff(self.rowptr_h))
        br, bc = node[rows], node[self.col_h]
        P = sp.csr_matrix((np.ones(len(br)), (br, bc)), shape=(nb, nb))
        P.sum_duplicates()
        rng = np.random.default_rng(12345)
        P.data = rng.uniform(1.0, 2.0, P.nnz)
        P = (P + sp.diags(np.asarray(abs(P).sum(axis=1)).ravel() + 1.0)).tocsr()
        lu = sla.splu(P[:nn, :nn].tocsc(), permc_spec="MMD_AT_PLUS_A",
                      options=dict(DiagPivotThresh=0.0, SymmetricMode=True))
        order = np.empty(nb, dtype=np.int64)   # order[k] = node eliminated k-th
        order[:nn] = np.argsort(lu.perm_c)
        order[nn] = nn
        pos = np.empty(nb, dtype=np.int64)
        pos[order] = np.arange(nb)
        lu2 = sla.splu(P[order][:, order].tocsc(), permc_spec="NATURAL",
                       options=dict(DiagPivotThresh=0.0, SymmetricMode=True))
        if not (np.array_equal(lu2.perm_r, np.arange(nb)) and np.array_equal(lu2.perm_c, np.arange(nb))):
            raise RfLUError("the structural block factorisation did not keep the order")
        T = _bundle(lu2.L.tocsr(), lu2.U.tocsr(), nb)
        if T.nnz * BS * BS >= 2 ** 31:
            raise RfLUError("block factors too large for 32-bit indices")
        self.nnzb = int(T.nnz)
        Trow = np.repeat(np.arange(nb, dtype=np.int64), np.diff(T.indptr))
        keys = Trow * nb + T.indices
        want = pos[br] * nb + pos[bc]
        blk = np.searchsorted(keys, want)
        if not np.array_equal(keys[blk], want):
            raise RfLUError("block pattern misses an entry of the matrix")
        dst = blk * (BS * BS) + local[rows] * BS + local[self.col_h]
        bd = int(np.searchsorted(keys, (nb - 1) * nb + (nb - 1)))
        pad = bd * BS * BS + np.arange(1, BS) * (BS + 1)
        self.ptrT = torch.from_numpy(T.indptr.astype(np.int32)).to(self.device)
        self.indT = torch.from_numpy(T.indices.astype(np.int32)).to(self.device)
        self.dst = torch.from_numpy(dst).to(self.device)
        self.pad = torch.from_numpy(pad).to(self.device)
        self.fwd = torch.from_numpy(pos[node] * BS + local).to(self.device)
        self.vals = torch.zeros(self.nnzb * BS * BS, dtype=self.dtype, device=self.device)

        L = _sparse_lib()
        p = lambda t: ctypes.c_void_p(t.data_ptr())  # noqa: E731
        self.handle = ctypes.c_void_p()
        _sp_ok(L.rocsparse_create_handle(ctypes.byref(self.handle)), "rocsparse_create_handle")
        self.descr, self.descrL, self.descrU = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_void_p()
        for d in (self.descr, self.descrL, self.descrU):
            _sp_ok(L.rocsparse_create_mat_descr(ctypes.byref(d)), "rocsparse_create_mat_descr")
        _sp_ok(L.rocsparse_set_mat_fill_mode(self.descrL, _FILL_LOWER), "set_mat_fill_mode")
        _sp_ok(L.rocsparse_set_mat_diag_type(self.descrL, _DIAG_UNIT), "set_mat_diag_type")
        _sp_ok(L.rocsparse_set_mat_fill_mode(self.descrU, _FILL_UPPER), "set_mat_fill_mode")
        _sp_ok(L.rocsparse_set_mat_diag_type(self.descrU, _DIAG_NON_UNIT), "set_mat_diag_type")
        self.info = ctypes.c_void_p()
        _sp_ok(L.rocsparse_create_mat_info(ctypes.byref(self.info)), "rocsparse_create_mat_info")
        self._stream()
        t = self._t
        sizes = []
        sz = ctypes.c_size_t()
        _sp_ok(getattr(L, f"rocsparse_{t}bsrilu0_buffer_size")(self.handle, _DIR_ROW, nb, self.nnzb, self.descr,
                                                                p(self.vals), p(self.ptrT), p(self.indT), BS,
                                                                self.info, ctypes.byref(sz)), "bsrilu0_buffer_size")
        sizes.append(sz.value)
        for d in (self.descrL, self.descrU):
            _sp_ok(getattr(L, f"rocsparse_{t}bsrsv_buffer_size")(self.handle, _DIR_ROW, _OP_NONE, nb, self.nnzb, d,
                                                                  p(self.vals), p(self.ptrT), p(self.indT), BS,
                                                                  self.info, ctypes.byref(sz)), "bsrsv_buffer_size")
            sizes.append(sz.value)
        self.buffer = torch.empty(max(max(sizes), 8), dtype=torch.uint8, device=self.device)
        _sp_ok(getattr(L, f"rocsparse_{t}bsrilu0_analysis")(self.handle, _DIR_ROW, nb, self.nnzb, self.descr,
                                                             p(self.vals), p(self.ptrT), p(self.indT), BS, self.info,
                                                             _ANALYSIS_FORCE, _SOLVE_AUTO, p(self.buffer)),
               "bsrilu0_analysis")
        for d in (self.descrL, self.descrU):
            _sp_ok(getattr(L, f"rocsparse_{t}bsrsv_analysis")(self.handle, _DIR_ROW, _OP_NONE, nb, self.nnzb, d,
                                                               p(self.vals), p(self.ptrT), p(self.indT), BS,
                                                               self.info, _ANALYSIS_FORCE, _SOLVE_AUTO,
                                                               p(self.buffer)), "bsrsv_analysis")
        self.y = torch.zeros(nb * BS, dtype=self.dtype, device=self.device)
        self.z = torch.zeros(nb * BS, dtype=self.dtype, device=self.device)
        torch.cuda.synchronize(self.device)
        self.analyses += 1

    def factor(self, values):
        """Factorises the matrix with the given device values (same pattern)."""
        L = _sparse_lib()
        if self.handle is None:
            self._analyse()
        self._A = values
        self._checked = False
        self._host_lu = None
        self._stream()
        self.vals.zero_()
        self.vals[self.dst] = values.to(self.dtype)
        self.vals[self.pad] = 1.0
        p = lambda t: ctypes.c_void_p(t.data_ptr())  # noqa: E731
        _sp_ok(getattr(L, f"rocsparse_{self._t}bsrilu0")(self.handle, _DIR_ROW, self.nb, self.nnzb, self.descr,
                                                          p(self.vals), p(self.ptrT), p(self.indT), self.BS,
                                                          self.info, _SOLVE_AUTO, p(self.buffer)), "bsrilu0")
        zp = ctypes.c_int32(-1)
        st = L.rocsparse_bsrilu0_zero_pivot(self.handle, self.info, ctypes.byref(zp))
        if st not in (0, 9):
            _sp_ok(st, "bsrilu0_zero_pivot")
        self._zero_pivot = st == 9
        self.refactorizations += 1

    def _residual(self, values, x, b):
        torch = self.torch
        A = torch.sparse_csr_tensor(self.ptrA, self.indA, values, size=(self.n, self.n))
        r = torch.mv(A, x) - b
        return float(torch.linalg.vector_norm(r)), float(torch.linalg.vector_norm(b))

    def solve(self, b, check: bool = False):
        """x with A x = b for the last factorised matrix (b is not modified)."""
        L = _sparse_lib()
        torch = self.torch
        if self.handle is None or self._A is None:
            raise RfLUError("factor() first")
        if self._host_lu is not None:
            return torch.from_numpy(self._host_lu.solve(b.cpu().numpy())).to(self.device)
        if not self._zero_pivot:
            self._stream()
            self.y.zero_()
            self.y[self.fwd] = b.to(self.dtype)
            one = (ctypes.c_double if self._t == "d" else ctypes.c_float)(1.0)
            p = lambda t: ctypes.c_void_p(t.data_ptr())  # noqa: E731
            solve = getattr(L, f"rocsparse_{self._t}bsrsv_solve")
            _sp_ok(solve(self.handle, _DIR_ROW, _OP_NONE, self.nb, self.nnzb, ctypes.byref(one), self.descrL,
                         p(self.vals), p(self.ptrT), p(self.indT), self.BS, self.info, p(self.y), p(self.z),
                         _SOLVE_AUTO, p(self.buffer)), "bsrsv_solve L")
            _sp_ok(solve(self.handle, _DIR_ROW, _OP_NONE, self.nb, self.nnzb, ctypes.byref(one), self.descrU,
                         p(self.vals), p(self.ptrT), p(self.indT), self.BS, self.info, p(self.z), p(self.y),
                         _SOLVE_AUTO, p(self.buffer)), "bsrsv_solve U")
            x = self.y[self.fwd].to(b.dtype)
        if self._zero_pivot or not self._checked:
            self._checked = True
            ok = False
            if not self._zero_pivot:
                rn, bn = self._residual(self._A, x, b)
                ok = rn <= self.check_tol * max(bn, 1e-300)
            if not ok:
                import scipy.sparse as sp
                import scipy.sparse.linalg as sla

                A = sp.csr_matrix((self._A.cpu().numpy(), self.col_h, self.rowptr_h), shape=(self.n, self.n))
                self._host_lu = sla.splu(A.tocsc(), permc_spec="MMD_AT_PLUS_A", options=dict(DiagPivotThresh=0.01))
                self.host_fallbacks += 1
                return torch.from_numpy(self._host_lu.solve(b.cpu().numpy())).to(self.device)
        return x


class HybridLU(BlockLU):
    """Scalar re-factorisation in the node-block order, block triangular solves.

    The order and block fill are BlockLU's.  The factorisation is rocSOLVER's
    csrrf_refactlu on the scalar expansion of the block fill pattern (without
    the arclength block's padding), which holds every fill entry, so it is the
    exact LU with static diagonal pivots; its L+U values are then gathered into
    the block storage, and the solves are BlockLU's two bsrsv calls.  The
    scalar re-factorisation is several times faster than rocSPARSE's block
    one for 6x6 blocks, the block solves several times faster than the scalar
    ones."""

    def _analyse(self):
        import scipy.sparse as sp

        torch = self.torch
        super()._analyse()  # order, block fill, block storage, rocSPARSE objects
        BS, nn, nb, n = self.BS, self.nn, self.nb, self.n
        ptrT = self.ptrT.cpu().numpy().astype(np.int64)
        indT = self.indT.cpu().numpy().astype(np.int64)
        # scalar expansion of the block pattern in the new order: new index of
        # (block b, local l) is 6b + l for real unknowns; the arclength block has l = 0 only
        rows, cols, dst = [], [], []
        nlocal = lambda b: 1 if b == nb - 1 else BS  # noqa: E731
        for bi in range(nb):
            bcols = indT[ptrT[bi]:ptrT[bi + 1]]
            bidx = np.arange(ptrT[bi], ptrT[bi + 1])
            cl = np.concatenate([bj * BS + np.arange(nlocal(bj)) for bj in bcols])
            cloc = np.concatenate([np.arange(nlocal(bj)) for bj in bcols])
            cblk = np.concatenate([np.full(nlocal(bj), k) for k, bj in zip(bidx, bcols)])
            for lr in range(nlocal(bi)):
                rows.append(np.full(len(cl), bi * BS + lr))
                cols.append(cl)
                dst.append(cblk * BS * BS + lr * BS + cloc)
        rows = np.concatenate(rows)
        cols = np.concatenate(cols)
        self.dst_T = torch.from_numpy(np.concatenate(dst)).to(self.device)
        # new scalar indices are 0..n-1 in this order (arclength last): remap 6b + l -> compact index
        newpos = np.full(nb * BS, -1, dtype=np.int64)
        real = np.concatenate([np.arange(BS * nn), [BS * nn]])
        newpos[real] = np.arange(n)
        r2, c2 = newpos[rows], newpos[cols]
        if (r2 < 0).any() or (c2 < 0).any():
            raise RfLUError("block expansion produced a padding index")
        ptr = np.zeros(n + 1, dtype=np.int64)
        np.cumsum(np.bincount(r2, minlength=n), out=ptr[1:])
        if not np.all(np.diff(r2) >= 0):
            raise RfLUError("block expansion not in row order")
        self.nnzS = len(c2)
        if self.nnzS >= 2 ** 31:
            raise RfLUError("scalar factors too large for 32-bit indices")
        self.ptrS = torch.from_numpy(ptr.astype(np.int32)).to(self.device)
        self.indS = torch.from_numpy(c2.astype(np.int32)).to(self.device)
        self.valS = torch.zeros(self.nnzS, dtype=self.dtype, device=self.device)
        # pivots: position k holds original unknown uorder[k]
        fwd = self.fwd.cpu().numpy()
        uorder = np.empty(n, dtype=np.int64)
        uorder[newpos[fwd]] = np.arange(n)
        piv = uorder.astype(np.int32)
        self.pivP = torch.from_numpy(piv).to(self.device)
        self.pivQ = torch.from_numpy(piv.copy()).to(self.device)
        rb, rs = _libs()
        self.rbhandle = ctypes.c_void_p()
        _ok(rb.rocblas_create_handle(ctypes.byref(self.rbhandle)), "rocblas_create_handle")
        rng = np.random.default_rng(12345)
        M = torch.from_numpy(rng.uniform(1.0, 2.0, size=len(self.col_h))).to(device=self.device, dtype=self.dtype)
        self.ptrA32 = torch.from_numpy(self.rowptr_h.astype(np.int32)).to(self.device)
        self.indA32 = torch.from_numpy(self.col_h).to(self.device)
        self.rfinfo = ctypes.c_void_p()
        _ok(rs.rocsolver_create_rfinfo(ctypes.byref(self.rfinfo), self.rbhandle), "rocsolver_create_rfinfo")
        self._rb_stream()
        p = lambda t: ctypes.c_void_p(t.data_ptr())  # noqa: E731
        rhs = torch.zeros(n, dtype=self.dtype, device=self.device)
        analysis = getattr(rs, f"rocsolver_{self._t}csrrf_analysis")
        _ok(analysis(self.rbhandle, n, 1, int(self.rowptr_h[-1]), p(self.ptrA32), p(self.indA32), p(M), self.nnzS,
                     p(self.ptrS), p(self.indS), p(self.valS), p(self.pivP), p(self.pivQ), p(rhs), n, self.rfinfo),
            f"rocsolver_{self._t}csrrf_analysis")
        torch.cuda.synchronize(self.device)

    def _rb_stream(self):
        rb, _ = _libs()
        s = self.torch.cuda.current_stream(self.device)
        _ok(rb.rocblas_set_stream(self.rbhandle, ctypes.c_void_p(s.cuda_stream)), "rocblas_set_stream")

    def close(self):
        rb, rs = _libs()
        if getattr(self, "rfinfo", None):
            rs.rocsolver_destroy_rfinfo(self.rfinfo)
            self.rfinfo = None
        if getattr(self, "rbhandle", None):
            rb.rocblas_destroy_handle(self.rbhandle)
            self.rbhandle = None
        super().close()

    def factor(self, values):
        _, rs = _libs()
        if self.handle is None:
            self._analyse()
        self._A = values
        self._checked = False
        self._host_lu = None
        self._zero_pivot = False
        self._rb_stream()
        p = lambda t: ctypes.c_void_p(t.data_ptr())  # noqa: E731
        vals = values if values.dtype == self.dtype else values.to(self.dtype)
        refactlu = getattr(rs, f"rocsolver_{self._t}csrrf_refactlu")
        _ok(refactlu(self.rbhandle, self.n, int(self.rowptr_h[-1]), p(self.ptrA32), p(self.indA32), p(vals),
                     self.nnzS, p(self.ptrS), p(self.indS), p(self.valS), p(self.pivP), p(self.pivQ), self.rfinfo),
            f"rocsolver_{self._t}csrrf_refactlu")
        self.vals.zero_()
        self.vals[self.dst_T] = self.valS
        self.vals[self.pad] = 1.0
        self.refactorizations += 1


_band = None


def _band_lib():
    """libvef_band.so (banded LU kernels, include/vef_band.h)."""
    global _band
    if _band is None:
        path = os.environ.get("VEF_BAND_LIBRARY") or os.path.join(os.path.dirname(os.path.abspath(__file__)),
                                                                   "libvef_band.so")
        if not os.path.exists(path):
            raise RfLUError(f"{path} not found: build it with __graft_entry__.build()")
        L = ctypes.CDLL(path)
        vp, i32, i64 = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64
        L.vef_band_create.argtypes = [i64, i32, i32, ctypes.POINTER(vp)]
        L.vef_band_create_prec.argtypes = [i64, i32, i32, i32, ctypes.POINTER(vp)]
        L.vef_band_storage.argtypes = [vp, ctypes.POINTER(vp), ctypes.POINTER(i64), ctypes.POINTER(i32)]
        L.vef_band_set.argtypes = [vp, i64, vp, vp, vp, vp]
        L.vef_band_factor.argtypes = [vp, vp]
        L.vef_band_uses_graph.argtypes = [vp]
        L.vef_band_solve.argtypes = [vp, vp, vp, vp]
        L.vef_band_last_error.argtypes = [vp]
        L.vef_band_last_error.restype = ctypes.c_char_p
        L.vef_band_destroy.argtypes = [vp]
        _band = L
    return _band


class _BorderedLU:
    """Shared part of the node-block factorisations with the dense arclength
    row and column as a border: A = [A11 a12; a21 a22] with A11 the node
    unknowns in the factorisation's own order (`pu`: unknown -> inner index).
    A subclass factorises A11 (`_factor11`) and solves with it (`_solve11`);
    here: z = A11^-1 a12 and the Schur complement s = a22 - a21.z once per
    factorisation, solves x2 = (b2 - a21.A11^-1 b1)/s, x1 = A11^-1 b1 - z x2,
    the residual check of the first solve after a factorisation (a host
    factorisation with threshold pivoting takes over if it fails) and, for
    FP32 storage, FP64 residuals and GMRES refinement with the FP32 factors
    as right preconditioner.  FP64 solves are checked by their normwise
    backward error; one that fails is refined the same way with the FP64
    factors before the host factorisation (only on systems up to
    HOST_FALLBACK_MAX unknowns) is considered."""

    # the host factorisation (threshold pivoting) only takes over on systems
    # up to this size; larger ones raise when refinement cannot pass the check
    HOST_FALLBACK_MAX = 200000
    # GMRES(30) restarts per refined solve: a factorisation that needs more is
    # too far from the matrix (the caller refactorises or restarts)
    gmres_restarts = 3

    def _border_setup(self, rowptr, col, device, precision, check_tol, refine_tol, pu):
        import torch

        if precision not in ("f64", "f32"):
            raise RfLUError(f"{type(self).__name__} precision must be 'f64' or 'f32'")
        self.torch = torch
        self.device = device
        self.precision = precision
        # FP64: a normwise backward error bound; FP32: a relative residual bound
        self.check_tol = (1e-12 if precision == "f64" else 1e-3) if check_tol is None else check_tol
        self.refine_tol = refine_tol
        self.refinements = 0      # solves refined by GMRES
        self.gmres_iterations = 0
        self.rowptr_h = np.asarray(rowptr, dtype=np.int64)
        self.col_h = np.asarray(col, dtype=np.int32)
        n = self.n = len(rowptr) - 1
        n1 = self.n1 = n - 1
        rows = np.repeat(np.arange(n, dtype=np.int64), np.diff(self.rowptr_h))
        cols = self.col_h.astype(np.int64)
        dev = lambda a: torch.from_numpy(np.ascontiguousarray(a, dtype=np.int64)).to(device)  # noqa: E731
        colb = (rows < n1) & (cols == n1)
        rowb = (rows == n1) & (cols < n1)
        corner = np.nonzero((rows == n1) & (cols == n1))[0]
        if len(corner) != 1:
            raise RfLUError("the pattern has no arclength diagonal entry")
        self.src_c, self.dst_c = dev(np.nonzero(colb)[0]), dev(pu[rows[colb]])
        self.src_r, self.dst_r = dev(np.nonzero(rowb)[0]), dev(pu[cols[rowb]])
        self.src_d = int(corner[0])
        self.pu = dev(pu)
        self.a12 = torch.zeros(n1, dtype=torch.float64, device=device)
        self.a21 = torch.zeros(n1, dtype=torch.float64, device=device)
        self.ptrA = torch.from_numpy(self.rowptr_h).to(device)
        self.indA = torch.from_numpy(cols).to(device)
        # rows the sparse product sums with a workgroup each (the arclength row)
        self.long_rows = torch.from_numpy(np.nonzero(np.diff(self.rowptr_h) > 16384)[0].astype(np.int64)).to(device)
        # the CSR product's partial sums of the long rows (256 chunks each)
        self._mv_work = torch.empty(max(1, 256 * len(self.long_rows)), dtype=torch.float64, device=device)
        self.analyses = 1
        self.refactorizations = 0
        self.host_fallbacks = 0
        self._A = None
        self._checked = False
        self._refine_all = False
        self._host_lu = None
        self.last_residual = None
        return rows, cols

    def _stream(self):
        return ctypes.c_void_p(self.torch.cuda.current_stream(self.device).cuda_stream)

    def zero_pivot_note(self) -> str:
        zp = getattr(self, "zero_pivots", None)
        return f"{zp()} fronts with a zero pivot" if zp else "factorisation"

    def factor(self, values):
        """Factorises the matrix with the given device values (same pattern)."""
        torch = self.torch
        self._A = values
        self._checked = False
        self._refine_all = False
        self._host_lu = None
        self._factor11(values)
        self._Anorm = float(torch.linalg.vector_norm(values))  # Frobenius norm, for the backward error
        self.a12.zero_()
        self.a12[self.dst_c] = values[self.src_c]
        self.a21.zero_()
        self.a21[self.dst_r] = values[self.src_r]
        self.z = self._solve11(self.a12.clone())
        self.s = values[self.src_d] - torch.dot(self.a21, self.z)
        self.refactorizations += 1

    def refresh_row(self, values):
        """Takes the border row (the last row: the arclength equation) from
        `values` and recomputes the border's Schur complement s = a22 - a21.z
        with this factorisation's A11 and z = A11^-1 a12, which the row does
        not touch: a new continuation step's arclength row costs one dot
        product, not a factorisation.  `values` becomes the matrix the checks
        and the refinement use."""
        torch = self.torch
        if self._A is None:
            raise RfLUError("factor() first")
        self._A = values
        self.a21.zero_()
        self.a21[self.dst_r] = values[self.src_r]
        self.s = values[self.src_d] - torch.dot(self.a21, self.z)

    def _mv(self, values, x):
        """A x for the matrix with `values` (this pattern): on the GPU the
        library's fixed-order CSR product (vef_mf_csr_mv; torch's sparse
        product has returned NaN once after a product with NaN inputs), on the
        host torch's."""
        torch = self.torch
        if not x.is_cuda:
            return torch.mv(torch.sparse_csr_tensor(self.ptrA, self.indA, values, size=(self.n, self.n)), x)
        from .multifrontal import _mf_lib

        y = torch.empty_like(x)
        p = lambda t: ctypes.c_void_p(t.data_ptr())  # noqa: E731
        rc = _mf_lib().vef_mf_csr_mv(self.n, p(self.ptrA), p(self.indA), p(values), p(x), p(y), len(self.long_rows),
                                     p(self.long_rows) if len(self.long_rows) else None, p(self._mv_work),
                                     self._stream())
        if rc != 0:
            raise RfLUError("vef_mf_csr_mv failed")
        return y

    def _residual(self, values, x, b):
        torch = self.torch
        r = self._mv(values, x) - b
        return float(torch.linalg.vector_norm(r)), float(torch.linalg.vector_norm(b))

    def _gmres(self, b, x, bn, tol, m=30, restarts=10, values=None):
        """Restarted GMRES(m) in FP64 for A x = b from x, right-preconditioned
        by the factorisation's solve; stops at ||b - A x|| <= tol ||b||.  A is
        the factorised matrix, or the one with `values` (same pattern).  The
        basis is orthogonalised by classical Gram-Schmidt applied twice (two
        products with the stored basis per iteration, one transfer to the
        host)."""
        torch = self.torch
        vals = self._A if values is None else values
        V = torch.empty((m + 1, self.n), dtype=torch.float64, device=self.device)
        for _ in range(restarts):
            r = b - self._mv(vals, x)
            beta = float(torch.linalg.vector_norm(r))
            if beta <= tol * bn or not np.isfinite(beta):
                return x, beta
            V[0] = r / beta
            Z = []
            H = np.zeros((m + 1, m))
            g = np.zeros(m + 1)
            g[0] = beta
            y = None
            for j in range(m):
                z = self._precondition(V[j])
                Z.append(z)
                w = self._mv(vals, z)
                Vj = V[:j + 1]
                h1 = torch.mv(Vj, w)
                w = w - torch.mv(Vj.t(), h1)
                h2 = torch.mv(Vj, w)
                w = w - torch.mv(Vj.t(), h2)
                hw = torch.cat([h1 + h2, torch.linalg.vector_norm(w).reshape(1)]).cpu().numpy()
                H[:j + 2, j] = hw
                self.gmres_iterations += 1
                if not np.all(np.isfinite(H[:j + 2, :j + 1])):
                    return x, beta
                y = np.linalg.lstsq(H[:j + 2, :j + 1], g[:j + 2], rcond=None)[0]
                est = float(np.linalg.norm(g[:j + 2] - H[:j + 2, :j + 1] @ y))
                if est <= tol * bn or H[j + 1, j] == 0.0:
                    break
                V[j + 1] = w / H[j + 1, j]
            xn = x
            for i, yi in enumerate(y):
                xn = xn + float(yi) * Z[i]
            rn = float(torch.linalg.vector_norm(b - self._mv(vals, xn)))
            if rn >= beta:  # no progress (the residual is at rounding level): keep the better iterate
                return x, beta
            x = xn
        r = b - self._mv(vals, x)
        return x, float(torch.linalg.vector_norm(r))

    def solve_krylov(self, values, b, tol: float, max_its: int):
        """x with A x = b for the matrix with `values` (this factorisation's
        pattern; typically a later Jacobian than the one factorised) by FP64
        GMRES(max_its) from 0, right-preconditioned by this factorisation.
        Returns (x, converged: ||b - A x|| <= tol ||b||, iterations)."""
        torch = self.torch
        if self._A is None:
            raise RfLUError("factor() first")
        bn = float(torch.linalg.vector_norm(b))
        it0 = self.gmres_iterations
        x, rn = self._gmres(b, torch.zeros_like(b), bn, tol, m=max_its, restarts=1, values=values)
        return x, bool(np.isfinite(rn) and rn <= tol * max(bn, 1e-300)), self.gmres_iterations - it0

    def _precondition(self, b):
        """The node-block solve with the border (FP64 vectors)."""
        torch = self.torch
        n1 = self.n1
        b1 = torch.empty(n1, dtype=torch.float64, device=self.device)
        b1[self.pu] = b[:n1]
        y = self._solve11(b1)
        x2 = (b[n1] - torch.dot(self.a21, y)) / self.s
        x = torch.empty_like(b)
        x[:n1] = (y - self.z * x2)[self.pu]
        x[n1] = x2
        return x

    # A subclass whose factorisation can run without row interchanges sets
    # this and implements _refactor_pivoted(): a solve that fails its check
    # (non-finite, or above the tolerance after refinement) on such a
    # factorisation refactorises the same matrix with interchanges and
    # solves again.
    _unpivoted = False

    def _refactor_pivoted(self):
        raise NotImplementedError

    def solve(self, b, check: bool = False):
        """x with A x = b for the last factorised matrix (b is not modified).
        check=True forms this solve's FP64 residual whatever the policy
        (refined or handed to the host factorisation like a first solve)."""
        try:
            return self._solve_checked(b, check)
        except RfLUError:
            if not self._unpivoted:
                raise
        self._refactor_pivoted()
        return self._solve_checked(b, check)

    def _solve_checked(self, b, check: bool):
        torch = self.torch
        if self._A is None:
            raise RfLUError("factor() first")
        if self._host_lu is not None:
            return torch.from_numpy(self._host_lu.solve(b.cpu().numpy())).to(self.device)
        if self.precision == "f32":
            x = self._precondition(b)
            if self._checked and not self._refine_all and not check:
                return x
            rn, bn = self._residual(self._A, x, b)
            self.last_residual = rn / max(bn, 1e-300)
            if not np.isfinite(rn):
                raise RfLUError(f"non-finite solve ({self.zero_pivot_note()})")
            if rn > self.refine_tol * bn:
                self.refinements += 1
                self._refine_all = True  # this factorisation's solves are checked from now on
                x, rn = self._gmres(b, x, bn, self.refine_tol, restarts=self.gmres_restarts)
            if rn <= self.check_tol * max(bn, 1e-300) or self.n > self.HOST_FALLBACK_MAX:
                # above the bound on a large system: the best iterate (the
                # inexact Newton refactorises when its corrections grow)
                self._checked = True
                return x
            return self._host_solve(b)
        x = self._precondition(b)
        if not self._checked or check or self._refine_all:
            # FP64: the normwise backward error ||r|| / (||A||_F ||x|| + ||b||)
            # (a backward-stable solve reaches eps times the factors' growth,
            # whatever the conditioning); above check_tol the solve is
            # refined by FP64 GMRES with this factorisation as preconditioner
            # and the factorisation's later solves are checked too
            self._checked = True
            rn, bn = self._residual(self._A, x, b)
            scale = self._Anorm * float(torch.linalg.vector_norm(x)) + bn
            self.last_residual = rn / max(scale, 1e-300)
            if not np.isfinite(rn):
                raise RfLUError(f"non-finite solve ({self.zero_pivot_note()})")
            if not (rn <= self.check_tol * scale):
                self.refinements += 1
                self._refine_all = True
                xr, rr = self._gmres(b, x, bn, self.check_tol * scale / max(bn, 1e-300), restarts=self.gmres_restarts)
                if rr < rn:
                    x, rn = xr, rr
                if rn <= self.check_tol * scale:
                    return x
                if self.n > self.HOST_FALLBACK_MAX:
                    raise RfLUError(f"solve backward error {rn / max(scale, 1e-300):.3e} above {self.check_tol:.1e} "
                                    f"after refinement (first solve: {self.last_residual:.3e})")
                return self._host_solve(b)
        return x

    def _host_solve(self, b):
        """Threshold-pivoting LU on the host (SuperLU) for this factorisation's
        matrix, used for the rest of its solves."""
        import scipy.sparse as sp
        import scipy.sparse.linalg as sla

        A = sp.csr_matrix((self._A.cpu().numpy(), self.col_h, self.rowptr_h), shape=(self.n, self.n))
        self._host_lu = sla.splu(A.tocsc(), permc_spec="MMD_AT_PLUS_A", options=dict(DiagPivotThresh=0.01))
        self.host_fallbacks += 1
        return self.torch.from_numpy(self._host_lu.solve(b.cpu().numpy())).to(self.device)


class BandLU(_BorderedLU):
    """Banded LU without pivoting on the GPU (libvef_band.so).

    The node unknowns are ordered by reverse Cuthill-McKee on the node graph
    (6 unknowns per node, consecutive), which on the cylinder meshes gives a
    half-bandwidth of about 6 x 254 unknowns at 1,600 elements; the dense
    arclength row and column form a border.  Each factorisation scatters the
    node block of the matrix into the band and factorises it in 64-column
    steps (a one-wavefront diagonal LU, thread-parallel triangular solves for
    the panel and the diagonal blocks' inverses, a rocBLAS dgemm for the
    trailing update — a dense, dependency-light schedule where the sparse
    re-factorisation is bound by its 3,097-level elimination chain), then
    solves A11 z = a12 and keeps the Schur complement s = a22 - a21.z.  A
    solve is x2 = (b2 - a21.A11^-1 b1)/s, x1 = A11^-1 b1 - z x2.  The first
    solve after a factorisation is checked as in RfLU (host fallback with
    threshold pivoting).  precision="f32" stores and factorises the band in
    FP32 (sgemm panels and updates) with FP64 solve vectors and border: the
    single-precision factorisation of an inexact Newton iteration.  With
    FP32 storage the first solve after a factorisation also forms its FP64
    residual (one sparse product), and so does every later one when that
    first solve needed refining; above refine_tol (relative) a solve is
    refined by restarted GMRES in FP64 with the FP32 factors as right
    preconditioner — what keeps the FP32
    factorisation usable where its plain solves lose too much accuracy (large
    meshes, whose condition number grows)."""

    NB = 64

    def __init__(self, rowptr: np.ndarray, col: np.ndarray, device="cuda:0", check_tol: float = None,
                 precision: str = "f64", refine_tol: float = 1.0e-3):
        import scipy.sparse as sp
        import torch
        from scipy.sparse.csgraph import reverse_cuthill_mckee

        n1 = len(rowptr) - 2
        if n1 % 6 != 0:
            raise RfLUError("BandLU expects 6 unknowns per node and one arclength unknown")
        nn = n1 // 6
        rowptr = np.asarray(rowptr, dtype=np.int64)
        rows = np.repeat(np.arange(n1 + 1, dtype=np.int64), np.diff(rowptr))
        cols = np.asarray(col, dtype=np.int64)
        inner = (rows < n1) & (cols < n1)
        G = sp.csr_matrix((np.ones(int(inner.sum())), (rows[inner] // 6, cols[inner] // 6)), shape=(nn, nn))
        order = reverse_cuthill_mckee(G, symmetric_mode=True).astype(np.int64)
        pos = np.empty(nn, dtype=np.int64)
        pos[order] = np.arange(nn)
        pu = pos[np.arange(n1) // 6] * 6 + np.arange(n1) % 6   # new index of each node unknown
        self._border_setup(rowptr, col, device, precision, check_tol, refine_tol, pu)
        pi, pj = pu[rows[inner]], pu[cols[inner]]
        self.w = int(np.abs(pi - pj).max()) if len(pi) else 0
        L = _band_lib()
        self.h = ctypes.c_void_p()
        bits = 64 if precision == "f64" else 32
        if L.vef_band_create_prec(n1, self.w, self.NB, bits, ctypes.byref(self.h)) != 0:
            raise RfLUError("vef_band_create_prec failed (device memory?)")
        ld, K = ctypes.c_int64(), ctypes.c_int32()
        self._err(L.vef_band_storage(self.h, None, ctypes.byref(ld), ctypes.byref(K)), "vef_band_storage")
        self.ld, self.K = ld.value, K.value
        dev = lambda a: torch.from_numpy(np.ascontiguousarray(a, dtype=np.int64)).to(device)  # noqa: E731
        self.src_in = dev(np.nonzero(inner)[0])
        self.dst_in = dev((self.K + pi - pj) + pj * self.ld)

    def close(self):
        if getattr(self, "h", None):
            _band_lib().vef_band_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _err(self, rc, what):
        if rc != 0:
            raise RfLUError(f"{what}: {_band_lib().vef_band_last_error(self.h).decode()}")

    @property
    def uses_graph(self) -> bool:
        """True once the factorisation replays its captured HIP graph."""
        return bool(_band_lib().vef_band_uses_graph(self.h))

    def _solve11(self, y):
        """A11^-1 y in the band order (y is consumed)."""
        x = self.torch.empty_like(y)
        self._err(_band_lib().vef_band_solve(self.h, ctypes.c_void_p(y.data_ptr()), ctypes.c_void_p(x.data_ptr()),
                                             self._stream()), "vef_band_solve")
        return x

    def _factor11(self, values):
        L = _band_lib()
        st = self._stream()
        p = lambda t: ctypes.c_void_p(t.data_ptr())  # noqa: E731
        self._err(L.vef_band_set(self.h, int(self.src_in.numel()), p(self.src_in), p(self.dst_in), p(values), st),
                  "vef_band_set")
        self._err(L.vef_band_factor(self.h, st), "vef_band_factor")
