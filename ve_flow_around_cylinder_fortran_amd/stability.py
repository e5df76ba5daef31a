"""Linear stability of the cylinder flow (Cylinder_Stability_2D) on the GPU
assembler: A and B of the generalised eigenproblem A q = lambda B q around a
converged base state (include/vef.h, vef_stability_assemble), as
assembleSlepcObjects (src/Stability/Stability_mod.F90:97-199) builds them for
the unchanged SLEPc shift-invert solve (solveEVP, :271-425).

The eigen solve itself is not part of this engine (SLEPc's Krylov-Schur with
shift-invert in the reference); `leading_eigenvalues` (SuperLU, small
meshes) and `leading_eigenvalues_gpu` (the operator (A - sigma B)^-1 B with
the GPU multifrontal LU, for the full-size meshes) drive ARPACK's
shift-invert Arnoldi from scipy as stand-ins for callers without SLEPc, used
by the tests and the bench to look at the spectrum.
"""
from __future__ import annotations

import numpy as np

from .assembly import Assembler
from .mesh import MESH_TOLERANCE, Mesh


def stability_dirichlet_rows(mesh: Mesh) -> np.ndarray:
    """Rows 6*g + q of applyDirichletBCStability for the boundary names of
    Cylinder_Stability_2D (setBoundaryNames, src/Boundary_mod.F90:106-108, and
    setBoundaryEQsStability, :319-439): Inlet and Outlet all six unknowns,
    Wall and Cylinder the velocity; nodes found by the distance tests of
    setBoundaryNodes (Mesh_Tolerance)."""
    g = mesh.geometry
    x, y = mesh.coords[:, 0], mesh.coords[:, 1]
    tol = MESH_TOLERANCE
    rows = []
    for on in (np.abs(x + g["L_up"]) < tol, np.abs(x - g["L_down"]) < tol):
        nodes = np.nonzero(on)[0]
        rows += [6 * nodes + q for q in range(6)]
    for on in (np.abs(y + g["H"]) < tol, np.abs(y - g["H"]) < tol, np.abs(x * x + y * y - g["R"] * g["R"]) < tol):
        nodes = np.nonzero(on)[0]
        rows += [6 * nodes + q for q in (0, 1)]
    return np.unique(np.concatenate(rows)).astype(np.int64)


class StabilityAssembler:
    """A and B on the GPU for base states on one mesh."""

    def __init__(self, mesh: Mesh, params, assembler: Assembler = None):
        self.asm = assembler if assembler is not None else Assembler(mesh, params)
        self.mesh = mesh
        self.rows = stability_dirichlet_rows(mesh)
        self.asm.set_stability_dirichlet(self.rows)
        self.rowptr, self.col = self.asm.stability_pattern()

    def assemble(self, x, A=None, B=None):
        """x: device tensor (base state, length 6*n_nodes + 1).  Returns the
        device value tensors (A, B) in the pattern (rowptr, col)."""
        return self.asm.stability(x, A, B)

    def to_scipy(self, A, B):
        import scipy.sparse as sp

        n = len(self.rowptr) - 1
        mA = sp.csr_matrix((A.cpu().numpy() if hasattr(A, "cpu") else A, self.col, self.rowptr), shape=(n, n))
        mB = sp.csr_matrix((B.cpu().numpy() if hasattr(B, "cpu") else B, self.col, self.rowptr), shape=(n, n))
        return mA, mB


def leading_eigenvalues(A, B, k: int = 10, sigma: float = 0.0, seed: int = 0):
    """The k eigenvalues of A q = lambda B q nearest sigma (shift-invert,
    ARPACK), sorted by decreasing real part.  A, B: scipy sparse matrices."""
    import scipy.sparse.linalg as sla

    n = A.shape[0]
    v0 = np.random.default_rng(seed).standard_normal(n)
    # (A - sigma B)^-1 B q = nu q with nu = 1/(lambda - sigma)
    lu = sla.splu((A - sigma * B).tocsc())
    op = sla.LinearOperator((n, n), matvec=lambda v: lu.solve(B @ v), dtype=np.float64)
    nu = sla.eigs(op, k=k, which="LM", v0=v0, return_eigenvectors=False, tol=1e-12)
    lam = sigma + 1.0 / nu
    return lam[np.argsort(-lam.real)]


def leading_eigenvalues_gpu(A, B, rowptr, col, mesh, k: int = 10, sigma: float = 0.0, seed: int = 0,
                            lu=None, vectors: bool = False, tol: float = 1e-12, ncv: int = 64, maxiter: int = None,
                            progress=None):
    """The k eigenvalues of A q = lambda B q nearest sigma, sorted by
    decreasing real part: ARPACK's shift-invert Arnoldi (scipy) on
    (A - sigma B)^-1 B, the factorisation and solves of A - sigma B on the GPU
    (multifrontal LU in nested-dissection order) and B q as a GPU sparse
    product.  A, B: device value tensors in the pattern (rowptr, col).
    progress: optional callable(applications) called every 25 operator
    applications; maxiter: ARPACK's restart limit.
    Returns (eigenvalues, eigenvectors or None, the LU, operator applications)."""
    import scipy.sparse.linalg as sla
    import torch

    from .multifrontal import NodeBlockLU

    n = len(rowptr) - 1
    dev = A.device
    if lu is None:
        lu = NodeBlockLU(rowptr, col, mesh.coords, mesh.conn, device=str(dev))
    lu.factor(A - sigma * B)
    Bt = torch.sparse_csr_tensor(torch.from_numpy(np.asarray(rowptr, dtype=np.int64)).to(dev),
                                 torch.from_numpy(np.asarray(col, dtype=np.int64)).to(dev), B, size=(n, n))
    count = [0]

    def matvec(v):
        count[0] += 1
        if progress is not None and count[0] % 25 == 0:
            progress(count[0])
        vd = torch.from_numpy(np.ascontiguousarray(v, dtype=np.float64)).to(dev)
        return lu.solve(torch.mv(Bt, vd)).cpu().numpy()

    op = sla.LinearOperator((n, n), matvec=matvec, dtype=np.float64)
    v0 = np.random.default_rng(seed).standard_normal(n)
    out = sla.eigs(op, k=k, which="LM", v0=v0, return_eigenvectors=vectors, tol=tol, ncv=ncv, maxiter=maxiter)
    nu, V = (out if vectors else (out, None))
    lam = sigma + 1.0 / nu
    o = np.lexsort((lam.imag, -lam.real))  # decreasing real part, conjugates by imaginary part
    return lam[o], (V[:, o] if V is not None else None), lu, count[0]
