"""Host check of the bordered factorisations' GMRES paths
(linsolve._BorderedLU): solve_krylov solves a matrix other than the one
factorised, preconditioned by the factorisation, to the requested relative
residual; a preconditioner far from the matrix misses it within the
iteration budget and says so.  A dense stand-in for the node-block
factorisation runs on the CPU."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from ve_flow_around_cylinder_fortran_amd.linsolve import _BorderedLU  # noqa: E402


class DenseLU(_BorderedLU):
    """The node block factorised densely on the host (torch, CPU)."""

    def __init__(self, rowptr, col):
        n1 = len(rowptr) - 2
        self._border_setup(rowptr, col, "cpu", "f32", None, 1e-3, np.arange(n1))

    def _factor11(self, values):
        A = torch.sparse_csr_tensor(self.ptrA, self.indA, values, size=(self.n, self.n)).to_dense()
        self._inv11 = torch.linalg.inv(A[:self.n1, :self.n1])

    def _solve11(self, y):
        return self._inv11 @ y


def bordered(n, rng, shift):
    """A dense bordered matrix (last row and column full) in CSR form."""
    A = rng.standard_normal((n, n)) / np.sqrt(n) + shift * np.eye(n)
    rowptr = np.arange(n + 1, dtype=np.int64) * n
    col = np.tile(np.arange(n, dtype=np.int32), n)
    return A, rowptr, col


def test_krylov_solves_a_later_matrix():
    rng = np.random.default_rng(3)
    n = 60
    A0, rowptr, col = bordered(n, rng, 4.0)
    A1 = A0 + 0.05 * rng.standard_normal((n, n)) / np.sqrt(n)  # a later Jacobian, close to A0
    lu = DenseLU(rowptr, col)
    lu.factor(torch.from_numpy(A0.ravel()))
    b = torch.from_numpy(rng.standard_normal(n))
    x, ok, its = lu.solve_krylov(torch.from_numpy(A1.ravel()), b, 1e-8, 20)
    assert ok and 1 <= its <= 20
    r = np.linalg.norm(A1 @ x.numpy() - b.numpy()) / np.linalg.norm(b.numpy())
    assert r <= 1e-8
    assert np.allclose(x.numpy(), np.linalg.solve(A1, b.numpy()), rtol=1e-6, atol=1e-8)
    # the plain solve still solves the factorised matrix
    y = lu.solve(b)
    assert np.linalg.norm(A0 @ y.numpy() - b.numpy()) <= 1e-3 * np.linalg.norm(b.numpy())


def test_krylov_reports_a_missed_tolerance():
    rng = np.random.default_rng(4)
    n = 60
    A0, rowptr, col = bordered(n, rng, 4.0)
    A1, _, _ = bordered(n, rng, 0.3)  # unrelated to the factorised matrix
    lu = DenseLU(rowptr, col)
    lu.factor(torch.from_numpy(A0.ravel()))
    b = torch.from_numpy(rng.standard_normal(n))
    x, ok, its = lu.solve_krylov(torch.from_numpy(A1.ravel()), b, 1e-10, 3)
    assert not ok and its == 3
    assert np.all(np.isfinite(x.numpy()))


def test_refresh_row_solves_the_new_border_without_refactorising():
    """A new arclength row (the last row) on a carried factorisation: after
    refresh_row the solve is exact for the new matrix (the node block and the
    border column unchanged), with no second factorisation."""
    rng = np.random.default_rng(5)
    n = 50
    A0, rowptr, col = bordered(n, rng, 4.0)
    A1 = A0.copy()
    A1[-1, :] = rng.standard_normal(n)  # a new tangent row
    A1[-1, -1] += 3.0
    lu = DenseLU(rowptr, col)
    lu.factor(torch.from_numpy(A0.ravel()))
    b = torch.from_numpy(rng.standard_normal(n))
    stale = lu.solve(b).numpy()
    assert np.linalg.norm(A1 @ stale - b.numpy()) > 1e-2 * np.linalg.norm(b.numpy())
    lu.refresh_row(torch.from_numpy(A1.ravel()))
    x = lu.solve(b).numpy()
    assert np.linalg.norm(A1 @ x - b.numpy()) <= 1e-10 * np.linalg.norm(b.numpy())
    assert lu.refactorizations == 1


class DenseUnpivotedLU(DenseLU):
    """The node block factorised by a Doolittle LU without row interchanges
    (an exact zero pivot makes its solves non-finite), with the pivoted dense
    inverse as the fallback."""

    def __init__(self, rowptr, col):
        super().__init__(rowptr, col)
        self.fallbacks = 0

    def factor(self, values):
        self._pivot = False
        self._unpivoted = True
        super().factor(values)

    def _factor11(self, values):
        if self._pivot:
            return super()._factor11(values)
        A = torch.sparse_csr_tensor(self.ptrA, self.indA, values, size=(self.n, self.n)).to_dense()
        LU = A[:self.n1, :self.n1].clone()
        with np.errstate(all="ignore"):
            for k in range(self.n1):
                LU[k + 1:, k] = LU[k + 1:, k] / LU[k, k]
                LU[k + 1:, k + 1:] -= torch.outer(LU[k + 1:, k], LU[k, k + 1:])
        self._lu = LU

    def _solve11(self, y):
        if self._pivot:
            return super()._solve11(y)
        L = torch.tril(self._lu, -1) + torch.eye(self.n1, dtype=torch.float64)
        U = torch.triu(self._lu)
        z = torch.linalg.solve_triangular(L, y.reshape(-1, 1), upper=False)
        return torch.linalg.solve_triangular(U, z, upper=True).reshape(-1)

    def _refactor_pivoted(self):
        self._pivot = True
        self._unpivoted = False
        self.fallbacks += 1
        _BorderedLU.factor(self, self._A)


@pytest.mark.parametrize("precision", ["f32", "f64"])
def test_failed_unpivoted_solve_refactorises_with_pivoting(precision):
    """A zero first pivot (the matrix stays nonsingular): the unpivoted
    factorisation's solve is not finite, the same matrix is factorised again
    with interchanges and the solve is exact; the next factorisation goes
    without interchanges again."""
    rng = np.random.default_rng(6)
    n = 40
    A, rowptr, col = bordered(n, rng, 3.0)
    A[0, 0] = 0.0
    lu = DenseUnpivotedLU(rowptr, col)
    lu.precision = precision
    lu.check_tol = 1e-12 if precision == "f64" else 1e-3
    lu.factor(torch.from_numpy(A.ravel()))
    b = torch.from_numpy(rng.standard_normal(n))
    x = lu.solve(b).numpy()
    assert lu.fallbacks == 1
    assert np.linalg.norm(A @ x - b.numpy()) <= 1e-10 * np.linalg.norm(b.numpy())
    A2 = A.copy()
    A2[0, 0] = 5.0
    lu.factor(torch.from_numpy(A2.ravel()))
    x2 = lu.solve(b).numpy()
    assert lu.fallbacks == 1 and not lu._pivot
    assert np.linalg.norm(A2 @ x2 - b.numpy()) <= 1e-10 * np.linalg.norm(b.numpy())
