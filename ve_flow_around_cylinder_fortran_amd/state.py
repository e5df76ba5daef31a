"""Synthetic solution vectors for the assembly workloads.

A state is the global solution vector x[N] in the reference's unknown
ordering: x[6*g + q] = TL(g, q) with q = Ux, Uy, P, Sxx, Sxy, Syy and
x[N-1] = EU(Nex), the arclength parameter (WiN = WiN0*EU, set per Gauss
point by setContinuationParameter, src/Element_Calculations_mod.F90:577-601).

`smooth_state` builds a smooth, divergence-free velocity field that vanishes
on the cylinder (stream function psi = psi0(y) * (1 - R^2/r^2)^2 with a
Poiseuille profile psi0), a linear pressure and the Newtonian conformation
S = I, then adds to every field a sum of 8 random-phase Fourier modes,
sum_m (amplitude/8) sin(k_m . x + phi_m), k_m in +-U[0.2, 2]^2, phi_m in
U[0, 2 pi), drawn from numpy.random.default_rng(seed).  It is a stand-in for
a converged Stokes state: the assembly's cost and data flow do not depend on
the values, and S stays symmetric positive definite.
"""
from __future__ import annotations

import numpy as np

from .mesh import Mesh

FIELDS = ("Ux", "Uy", "P", "Sxx", "Sxy", "Syy")


def base_flow(mesh: Mesh) -> np.ndarray:
    """(n_nodes, 6) smooth base fields."""
    R = mesh.geometry["R"]
    H = mesh.geometry["H"]
    x, y = mesh.coords[:, 0], mesh.coords[:, 1]
    r = np.maximum(np.hypot(x, y), R)
    psi0 = 1.5 * (y - y ** 3 / (3.0 * H * H))
    dpsi0 = 1.5 * (1.0 - (y / H) ** 2)
    q = 1.0 - (R / r) ** 2
    g = q * q
    dg = 2.0 * q * (2.0 * R * R / r ** 3)
    u = np.zeros((mesh.n_nodes, 6))
    u[:, 0] = dpsi0 * g + psi0 * dg * y / r
    u[:, 1] = -psi0 * dg * x / r
    u[:, 2] = -0.75 * x
    u[:, 3] = 1.0
    u[:, 4] = 0.0
    u[:, 5] = 1.0
    return u


def fourier_perturbation(mesh: Mesh, seed: int = 20261203, amplitude: float = 1.0e-2, modes: int = 8) -> np.ndarray:
    rng = np.random.default_rng(seed)
    x, y = mesh.coords[:, 0], mesh.coords[:, 1]
    out = np.zeros((mesh.n_nodes, 6))
    for f in range(6):
        kx = rng.uniform(0.2, 2.0, modes) * rng.choice([-1.0, 1.0], modes)
        ky = rng.uniform(0.2, 2.0, modes) * rng.choice([-1.0, 1.0], modes)
        ph = rng.uniform(0.0, 2.0 * np.pi, modes)
        for m in range(modes):
            out[:, f] += (amplitude / modes) * np.sin(kx[m] * x + ky[m] * y + ph[m])
    return out


def smooth_state(mesh: Mesh, EU: float = 0.5, seed: int = 20261203, amplitude: float = 1.0e-2) -> np.ndarray:
    """x[N]: base flow + Fourier perturbation, EU last."""
    u = base_flow(mesh) + fourier_perturbation(mesh, seed=seed, amplitude=amplitude)
    x = np.empty(mesh.n_unknowns)
    x[:-1] = u.reshape(-1)
    x[-1] = EU
    return x


def rest_state(mesh: Mesh, EU: float = 0.5) -> np.ndarray:
    """U = 0, P = 0, S = I: the equilibrium at rest."""
    u = np.zeros((mesh.n_nodes, 6))
    u[:, 3] = 1.0
    u[:, 5] = 1.0
    x = np.empty(mesh.n_unknowns)
    x[:-1] = u.reshape(-1)
    x[-1] = EU
    return x
