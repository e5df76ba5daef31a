"""Newton counterpart for one continuation step with the arclength unknown
pinned (the reference's increments 1-2), built on the GPU assembler.

Per iteration, in the order of assemblePetscObjects
(src/Newton_Raphson_mod.F90:281-400): residual (and, for 'NRP', the
forward-difference Jacobian) on the GPU; the arclength row pinning EU
(src/Extra_Equations_mod.F90:698-760); the Dirichlet rows — the fully
developed inflow first, then Outlet P = 0, Wall and Cylinder U = 0
(src/Dirichlet_mod.F90:104-190, src/Boundary_mod.F90:218-245); ||b||; the
sparse solve (GPU re-factorisation, `linsolve.RfLU`; 'MNR' iterations reuse the
last factors as the reference reuses its factored matrix); ||s||; the flag
logic of checkNewtonConvergence (:470-528); x <- x - xF*s (updateNewton,
:618-680).  Convergence is the reference's test at the top of the loop:
||s|| < 1e-8 and ||b|| < 1e-8 (Cor_Norm_eps, src/Parameters_mod.F90:214).

The fully developed inflow: the reference solves a separate 1D problem for it
(loopNewtonRaphsonInflow); here it is the closed-form Oldroyd-B channel
solution with mean velocity 1 (the flow rate 2*Height of Inflow_1D,
src/Physical_mod.F90:48-49), so this driver takes Oldroyd-B (VeN = 0) only.
"""
from __future__ import annotations

import dataclasses
import math
import time
from typing import List

import numpy as np

from .assembly import Assembler, Params
from .inflow import InflowError
from .linsolve import RfLUError
from .mesh import MESH_TOLERANCE, Mesh

COR_NORM_EPS = 1.0e-8
MNR_EPS = 1.0e-1
COR_NORM_DIV = 1.0e8
NITER = 50


def oldroyd_b_inflow(y: np.ndarray, H: float, WiN: float):
    """Fully developed Oldroyd-B channel flow, mean velocity 1:
    U = 1.5 (1 - (y/H)^2), gamma = dU/dy, C = [[1 + 2 (Wi gamma)^2, Wi gamma],
    [Wi gamma, 1]] and S = C^(1/2) (the symmetric positive root).  Returns
    (Ux, Uy, Sxx, Sxy, Syy)."""
    y = np.asarray(y, dtype=np.float64)
    U = 1.5 * (1.0 - (y / H) ** 2)
    g = -3.0 * y / H ** 2
    c11 = 1.0 + 2.0 * (WiN * g) ** 2
    c12 = WiN * g
    c22 = np.ones_like(y)
    s = np.sqrt(c11 * c22 - c12 * c12)
    t = np.sqrt(c11 + c22 + 2.0 * s)
    return U, np.zeros_like(y), (c11 + s) / t, c12 / t, (c22 + s) / t


def _fixed_dirichlet_rows(mesh: Mesh):
    """The inlet nodes (closed-form order) and the rows of the conditions that
    do not depend on Wi (outlet P, walls and cylinder U), in the reference's
    insertion order; kept on the mesh, keyed by its coordinate array and a
    sample of its values (a mesh is not changed in place after its first
    solve)."""
    c = mesh.coords
    key = (id(c), c.shape, c.__array_interface__["data"][0], float(np.sum(c[::97])))
    cached = mesh.__dict__.get("_vef_fixed_dirichlet")
    if cached is not None and cached[0] == key:
        return cached[1], cached[2]
    g = mesh.geometry
    x, y = mesh.coords[:, 0], mesh.coords[:, 1]
    tol = MESH_TOLERANCE
    inlet = np.nonzero(np.abs(x + g["L_up"]) < tol)[0]
    rows = [6 * np.nonzero(np.abs(x - g["L_down"]) < tol)[0] + 2]
    for on in (np.abs(y + g["H"]) < tol, np.abs(y - g["H"]) < tol,
               np.abs(x * x + y * y - g["R"] * g["R"]) < tol):
        nodes = np.nonzero(on)[0]
        for q in (0, 1):
            rows.append(6 * nodes + q)
    fixed = np.concatenate(rows).astype(np.int64)
    mesh.__dict__["_vef_fixed_dirichlet"] = (key, inlet, fixed)
    return inlet, fixed


def cylinder_dirichlet(mesh: Mesh, WiN: float, inflow=None):
    """Rows and values of the Cylinder_2D Dirichlet conditions, in the order
    the reference inserts them (a row listed twice keeps the later value).
    The inlet takes the fully developed profile: the Oldroyd-B closed form, or
    with `inflow` (an inflow.FullyDevelopedInflow of this mesh) the solution
    of the reference's Inflow_1D problem at WiN, for any model."""
    g = mesh.geometry
    inlet, fixed = _fixed_dirichlet_rows(mesh)
    if inflow is None:
        prof = oldroyd_b_inflow(mesh.coords[inlet, 1], g["H"], WiN)
    else:
        inlet = inflow.nodes
        prof, _ = inflow.solve(WiN)
    rows = [6 * inlet + q for q in (0, 1, 3, 4, 5)]  # Ux, Uy, Sxx, Sxy, Syy (P is skipped)
    rows.append(fixed)
    vals = list(prof) + [np.zeros(len(fixed))]
    return np.concatenate(rows).astype(np.int64), np.concatenate(vals)


def inlet_sensitivity(mesh: Mesh, WiN: float, inflow=None):
    """Rows of the inlet's Dirichlet conditions (Ux, Uy, Sxx, Sxy, Syy, the
    rows cylinder_dirichlet lists first) and d(value)/dWiN of each: from the
    fully developed inflow problem's sensitivity, or for the Oldroyd-B closed
    form a central difference of it."""
    g = mesh.geometry
    if inflow is None:
        inlet = np.nonzero(np.abs(mesh.coords[:, 0] + g["L_up"]) < MESH_TOLERANCE)[0]
        h = 1.0e-6 * max(1.0, abs(WiN))
        yp = oldroyd_b_inflow(mesh.coords[inlet, 1], g["H"], WiN + h)
        ym = oldroyd_b_inflow(mesh.coords[inlet, 1], g["H"], WiN - h)
        sens = [(a - b) / (2.0 * h) for a, b in zip(yp, ym)]
    else:
        inlet = inflow.nodes
        sens = inflow.sensitivity(WiN)
    rows = np.concatenate([6 * inlet + q for q in (0, 1, 3, 4, 5)]).astype(np.int64)
    return rows, np.concatenate(sens)


def initial_state(mesh: Mesh, EU: float) -> np.ndarray:
    """The rest state U = 0, P = 0, S = I with the arclength unknown EU: the
    first Newton iteration from it is the Stokes-like solve with the boundary
    values, the start the reference takes from its Newtonian state."""
    x = np.zeros(mesh.n_unknowns)
    u = x[:-1].reshape(-1, 6)
    u[:, 3] = 1.0
    u[:, 5] = 1.0
    x[-1] = EU
    return x


@dataclasses.dataclass
class NewtonRecord:
    iteration: int
    flag: str
    res_norm: float
    cor_norm: float
    seconds: float


class NewtonError(RuntimeError):
    pass


class NewtonSolver:
    """Fixed-Wi Newton solve (arclength unknown pinned at LN0) on the GPU.

    reuse_factors=False follows the reference: every solve starts with an
    'NRP' iteration (a fresh Jacobian and factorisation).  reuse_factors=True
    is an inexact Newton that carries the last factorisation into the next
    solve: it starts with 'MNR' iterations on the previous Jacobian and
    refactorises ('NRP') when a modified-Newton correction grows (the
    reference's rule), or when, at the contraction r = ||s_k||/||s_k-1|| of
    the last two corrections, more than reuse_max_iters further iterations
    would be needed to bring ||s|| below the tolerance (a factorisation costs
    about as much as that many modified-Newton iterations), and every 10th
    iteration.  reuse_max_iters="auto" takes that count from the solver's
    own measured times: the extra wall time of a factorising iteration over
    a modified-Newton one, in modified-Newton iterations (running means; at
    least 1) — about 8 on the 1,600-element mesh's banded LU, about 5 for the
    multifrontal LU at 40,000 elements.  The residual, the update and the convergence test are
    unchanged (FP64, ||s|| and ||b|| below 1e-8), so the converged states
    agree with the reference policy's within that tolerance; the first solve
    after an FP64 factorisation is then checked against 1e-6 instead of 1e-10
    (the Newton iteration absorbs a less accurate solve).

    inlet_coupling=True (not the reference's): in an arclength step the
    inlet's Dirichlet values follow EU (the fully developed inflow at WiN =
    WiN0*EU), but the reference's Jacobian keeps those rows as unit rows with
    no EU column, so its Newton iteration (and the tangent J t = e_N) treat
    the inlet values as fixed and converge only linearly in that coupling.
    With the option each inlet row carries -d(value)/dEU in the arclength
    column of every factorised Jacobian (the inflow problem's own
    sensitivity), making the Jacobian of the same equations exact; the
    residual, update and convergence test are unchanged.

    krylov=True (with reuse_factors; not the reference's): an iteration on a
    carried factorisation assembles the current Jacobian (the solver's own
    kind) and solves with it by FP64 GMRES(krylov_max_its) right-
    preconditioned with the carried factorisation, to krylov_tol relative to
    ||b|| (record flag 'KRY'): an inexact Newton iteration on the current
    Jacobian in place of a modified-Newton one on the old, so the carried
    factorisation lasts as long as it preconditions well.  A GMRES run that
    misses the tolerance is not applied; the next iteration refactorises at
    the same iterate.  Everything else (residual, update, convergence test,
    the growth and contraction rules) is as with reuse_factors.

    refine_tol (FP32 factorisations, 'mf' or 'band'): the relative residual to
    which GMRES refines a solve that misses it (1e-3 when None); on the
    40,000-element continuation 1e-2 and 1e-1 saved no time, the stale
    factorisation's corrections then costing Newton iterations instead.

    Escalation (reuse_factors with an FP32 factorisation or the FP32
    Jacobian): after an emergency restart, the rest of that solve uses the
    reference's ingredients, the FP64 forward-difference Jacobian and an FP64
    factorisation (a second factorisation object, built on first use), as
    every iteration does in the reference policy; the next solve returns to
    the configured ones.  Near a singular point of the branch (linear PTT at
    Wi = 0.324 on the 40,000-element mesh) the FP32 Jacobian and factorisation
    do not carry the Newton iteration through, the FP64 ones do."""

    def __init__(self, mesh: Mesh, params: Params, device: str = "cuda:0", linear_precision: str = "f64",
                 reuse_factors: bool = False, reuse_max_iters: float = 8.0, factorization: str = None,
                 jacobian: str = "fd64", inflow: str = None, inlet_coupling: bool = False, krylov: bool = False,
                 krylov_tol: float = 1.0e-3, krylov_max_its: int = 10, refine_tol: float = None):
        import torch

        from .linsolve import BandLU, BlockLU, HybridLU, RfLU

        if not (reuse_max_iters == "auto" or (isinstance(reuse_max_iters, (int, float)) and reuse_max_iters > 0)):
            raise NewtonError("reuse_max_iters must be a positive number or 'auto'")
        if krylov and not reuse_factors:
            raise NewtonError("krylov=True needs reuse_factors=True (it preconditions with a carried factorisation)")
        if not (0.0 < krylov_tol < 1.0) or int(krylov_max_its) < 1:
            raise NewtonError("krylov_tol must lie in (0, 1) and krylov_max_its be at least 1")
        if refine_tol is not None and not (linear_precision == "f32" and 0.0 < refine_tol < 1.0):
            raise NewtonError("refine_tol applies to an FP32 factorisation and must lie in (0, 1)")
        oldroyd_b = params.VeN == 0.0 and params.model == "L-PTT"
        if inflow is None:
            inflow = "closed_form" if oldroyd_b else "fe1d"
        if inflow == "closed_form" and not oldroyd_b:
            raise NewtonError("the closed-form inflow covers Oldroyd-B (L-PTT with VeN = 0) only; use inflow='fe1d'")
        if inflow not in ("closed_form", "fe1d"):
            raise NewtonError("inflow must be 'closed_form' (Oldroyd-B) or 'fe1d' (the reference's Inflow_1D problem)")
        from .inflow import FullyDevelopedInflow

        self.inflow = FullyDevelopedInflow.from_mesh(mesh, params) if inflow == "fe1d" else None
        self.torch = torch
        self.mesh = mesh
        self.params = params
        self.device = device
        self.asm = Assembler(mesh, params)
        rowptr, col = self.asm.pattern()
        from .multifrontal import MultifrontalLU

        classes = {"rf": RfLU, "block": BlockLU, "hybrid": HybridLU, "band": BandLU, "mf": MultifrontalLU}
        if factorization is None:  # the fastest for both precisions
            factorization = "band"
        if factorization not in classes:
            raise NewtonError("factorization must be 'rf' (scalar re-factorisation), 'block' (6x6 node blocks), "
                              "'hybrid' (scalar re-factorisation in node-block order, block solves), 'band' "
                              "(banded LU in reverse Cuthill-McKee node order) or 'mf' (multifrontal LU in "
                              "nested-dissection order)")
        check_tol = 1e-6 if (reuse_factors and linear_precision == "f64") else None

        def make_lu(precision, tol, rtol=None):
            kw = {} if rtol is None else {"refine_tol": rtol}
            if factorization == "mf":
                return MultifrontalLU(rowptr, col, mesh.coords, mesh.conn, device=device, precision=precision,
                                      check_tol=tol, plan=getattr(getattr(self, "lu", None), "plan", None), **kw)
            return classes[factorization](rowptr, col, device=device, precision=precision, check_tol=tol, **kw)

        if refine_tol is not None:
            if factorization not in ("mf", "band"):
                raise NewtonError("refine_tol applies to the 'mf' and 'band' factorisations (GMRES refinement)")
            # an FP32 solve above refine_tol (relative residual) is refined to
            # it by GMRES, and one below it is accepted
            check_tol = max(1e-3, refine_tol)

        self.lu = make_lu(linear_precision, check_tol, refine_tol)
        # the inexact Newton's escalation after an emergency restart (see the
        # class docstring): an FP64 factorisation built on first use
        self._make_lu64 = lambda: make_lu("f64", 1e-6)
        self._lu_primary = self.lu
        self._lu64 = None
        self._escalated = False
        self.escalations = 0
        self.linear_precision = linear_precision
        self.N = self.asm.N
        self.b = torch.empty(self.N, dtype=torch.float64, device=device)
        self.values = torch.empty(self.asm.nnz, dtype=torch.float64, device=device)
        if jacobian not in ("fd64", "ad32"):
            raise NewtonError("jacobian must be 'fd64' (the reference's forward differences) or 'ad32' (analytic, FP32)")
        self.jac_flag = "NRP" if jacobian == "fd64" else "NRP32"
        self._jac_flag_primary = self.jac_flag
        self.reuse_factors = reuse_factors
        self.reuse_max_iters = reuse_max_iters
        self._t_nrp = self._t_mnr = None  # running mean wall times of factorising / modified-Newton iterations
        self.inlet_coupling = inlet_coupling
        self.krylov = krylov
        self.krylov_tol = float(krylov_tol)
        self.krylov_max_its = int(krylov_max_its)
        self.krylov_solves = 0       # Newton-Krylov corrections computed
        self.krylov_failures = 0     # of them, GMRES runs that missed krylov_tol (refactorised instead)
        self._rowptr = torch.from_numpy(np.asarray(rowptr, dtype=np.int64)).to(device)
        self.factored = False
        self.emergencies = []
        self.on_iteration = None  # optional callback(record, records in this solve, xF) per iteration

    def solve(self, x0, LN0: float, xF: float = 1.0, max_iter: int = NITER, increment: int = 1, dArN: float = 0.0,
              x_old=None, tangent=None, dS: float = 0.0, cor_norm_div: float = COR_NORM_DIV):
        """Newton from x0 (numpy or device tensor); returns (x as a device
        tensor, list of NewtonRecord).  Increments 1-2 pin EU at
        LN0 + (increment-1)*dArN; from increment 3 the pseudo-arclength row
        tangent.(x - x_old) = dS closes the system (x_old, tangent: device
        tensors).  The inflow values follow the current EU, as the reference
        re-solves its inflow problem inside every arclength assembly.

        Recovery (checkNewtonConvergence and emergencyNewtonRaphson,
        src/Newton_Raphson_mod.F90:453-584): when ||s|| exceeds cor_norm_div
        or the iteration count passes its cap, the solve restarts from x0 (the
        step's prediction) with the relaxation xF halved, the cap raised to
        50*max_iter, dS halved and a fresh Jacobian; below xF = 1e-2 it stops
        with NewtonError.  A factorisation that cannot solve the system (zero
        pivots) or a non-finite correction counts as divergence; a non-finite
        residual stops at once, as the reference's general stop on NaN does,
        and an iterate whose Wi has no fully developed inflow counts as
        divergence.  With reuse_factors, a modified-Newton correction that
        grows is not applied (the next iteration refactorises at the same
        iterate).  `self.dS_used` and `self.emergencies` report the step
        length and the restarts of the last solve."""
        torch = self.torch
        x0 = x0.clone() if torch.is_tensor(x0) else torch.from_numpy(np.asarray(x0, dtype=np.float64)).to(self.device)
        x = x0.clone()
        if self._escalated:  # back to the configured Jacobian and factorisation after an escalated solve
            self.lu, self.jac_flag = self._lu_primary, self._jac_flag_primary
            self._escalated = False
            self.factored = False
        flag = "MNR" if (self.reuse_factors and self.factored) else "NRP"
        res_norm = cor_new = cor_old = 1.0
        applied = 0  # corrections applied since the solve's start (or last restart)
        it = 0
        cap = max_iter
        hist: List[NewtonRecord] = []
        self.emergencies = []
        last_EU = None
        while True:
            if cor_new < COR_NORM_EPS and res_norm < COR_NORM_EPS:
                break
            t0 = time.perf_counter()
            it += 1
            if it % 10 == 0:
                flag = "NRP"
            jac = flag == "NRP"
            EU = float(x[-1])
            inflow_failed = False
            reject = False
            if EU != last_EU:
                try:
                    rows, vals = cylinder_dirichlet(self.mesh, self.params.WiN0 * EU, self.inflow)
                except InflowError:
                    # no fully developed inflow at this iterate's Wi (a wild
                    # iterate): divergence of this solve
                    inflow_failed = True
                else:
                    self.asm.set_dirichlet(rows, vals)
                    last_EU = EU
            if inflow_failed:
                hist.append(NewtonRecord(it, flag, math.nan, math.inf, 0.0))
                cor_new = math.inf
                s = None
            else:
                # Newton-Krylov (krylov=True): an iteration on a carried
                # factorisation takes the current Jacobian and solves it by
                # GMRES preconditioned with that factorisation
                kry = not jac and self.krylov and self.factored
                mat = jac or kry
                aflag = self.jac_flag if mat else flag
                # on a carried factorisation, the border row follows the
                # current arclength equation (a new step's tangent): written
                # into the values the factorisation checks against, its Schur
                # complement refreshed (no solve)
                carried = (not jac and self.reuse_factors and self.factored and increment >= 3
                           and hasattr(self.lu, "refresh_row"))  # the bordered factorisations
                self.asm.assemble(x, aflag, b=self.b, values=self.values if mat else None)
                self.asm.arclength(x, self.b, self.values if (mat or carried) else None, increment=increment, LN0=LN0,
                                   dArN=dArN, x_old=x_old, tangent=tangent, dS=dS, flag_nr=aflag)
                if carried:
                    self.lu.refresh_row(self.values)
                self.asm.apply_dirichlet(x, self.b, self.values if mat else None, aflag)
                coupled = True
                if mat and self.inlet_coupling and increment >= 3:
                    try:
                        self._couple_inlet(EU)
                    except InflowError:
                        coupled = False  # no inflow sensitivity at this iterate's Wi
                res_norm = float(torch.linalg.vector_norm(self.b))
                if not math.isfinite(res_norm):
                    raise NewtonError("non-finite residual")  # the reference's general stop on NaN
                s, cor_new = None, math.inf  # divergence unless a correction is found
                kry_failed = False
                if coupled:
                    try:
                        if jac:
                            self.lu.factor(self.values)
                            self.factored = True
                        if kry:
                            s, ok, _ = self.lu.solve_krylov(self.values, self.b, self.krylov_tol, self.krylov_max_its)
                            self.krylov_solves += 1
                            if not ok:
                                self.krylov_failures += 1
                                kry_failed = True
                        else:
                            s = self.lu.solve(self.b)
                        cor_new = float(torch.linalg.vector_norm(s))
                    except RfLUError:
                        # a Jacobian this factorisation cannot solve (zero pivots at a
                        # wandering iterate): under MUMPS's static pivoting the
                        # reference sees a huge correction, i.e. divergence
                        s, cor_new = None, math.inf
                if not math.isfinite(cor_new):
                    cor_new = math.inf
                hist.append(NewtonRecord(it, "KRY" if kry else flag, res_norm, cor_new, 0.0))
                self.last_history = hist
                if self.on_iteration is not None:
                    self.on_iteration(hist[-1], len(hist), xF)
                if flag == "NRP":
                    flag = "MNR" if cor_new < MNR_EPS else "NRP"
                elif self.reuse_factors:
                    flag = "MNR"
                    if kry_failed:
                        # the carried factorisation no longer preconditions the
                        # current Jacobian well enough: refactorise at this iterate
                        flag = "NRP"
                        reject = math.isfinite(cor_new)
                    elif cor_new > cor_old:
                        flag = "NRP"
                        # inexact Newton only: a growing correction from a carried
                        # factorisation is not applied; the next iteration
                        # refactorises at the same iterate
                        reject = math.isfinite(cor_new)
                    elif cor_new > COR_NORM_EPS and cor_old > 0.0 and applied > 0:
                        # the contraction of two corrections of this solve
                        # (before the first, cor_old is a placeholder)
                        r = cor_new / cor_old
                        if 0.0 < r < 1.0 and math.log(COR_NORM_EPS / cor_new) / math.log(r) > self._reuse_limit():
                            flag = "NRP"
                else:
                    flag = "NRP" if cor_new > cor_old else "MNR"
            if reject and not (cor_new > cor_norm_div or it > cap):
                torch.cuda.synchronize()
                hist[-1].seconds = time.perf_counter() - t0
                continue
            if cor_new > cor_norm_div or it > cap:
                if s is None or not math.isfinite(cor_new):
                    self.factored = False  # no usable factorisation carried over
                # emergencyNewtonRaphson: restart from the prediction, relaxed
                why = "divergence" if cor_new > cor_norm_div else "iterations"
                xF = 0.5 * xF
                self.emergencies.append((len(hist), why, xF))
                if xF < 1.0e-2:
                    raise NewtonError(f"no convergence: relaxation factor {xF:g} below 1e-2 after "
                                      f"{len(self.emergencies)} restarts (last: {why}, ||s|| = {cor_new:.3e})")
                cap = 50 * max_iter
                dS = 0.5 * dS
                if self.reuse_factors and (self.linear_precision == "f32" or self.jac_flag != "NRP") \
                        and not self._escalated:
                    # inexact Newton: the rest of this solve with the reference's
                    # FP64 forward-difference Jacobian and an FP64 factorisation
                    if self._lu64 is None:
                        self._lu64 = self._make_lu64()
                    self.lu, self.jac_flag = self._lu64, "NRP"
                    self._escalated = True
                    self.escalations += 1
                    self.factored = False
                x = x0.clone()
                it = 0
                res_norm = cor_new = cor_old = 1.0
                applied = 0
                flag = "NRP"
                last_EU = None
                torch.cuda.synchronize()
                hist[-1].seconds = time.perf_counter() - t0
                continue
            cor_old = cor_new
            applied += 1
            x = x - xF * s
            torch.cuda.synchronize()
            hist[-1].seconds = time.perf_counter() - t0
            self._time_iteration(hist[-1])
        self.dS_used = dS
        return x, hist

    def _time_iteration(self, rec: NewtonRecord):
        a = 0.3  # weight of the newest iteration in the running means
        if rec.flag in ("MNR", "KRY"):
            self._t_mnr = rec.seconds if self._t_mnr is None else (1 - a) * self._t_mnr + a * rec.seconds
        else:
            self._t_nrp = rec.seconds if self._t_nrp is None else (1 - a) * self._t_nrp + a * rec.seconds

    def _reuse_limit(self) -> float:
        """Modified-Newton iterations a factorisation is worth (reuse_max_iters)."""
        if self.reuse_max_iters != "auto":
            return float(self.reuse_max_iters)
        if self._t_nrp is None or not self._t_mnr:
            return 8.0
        return max(1.0, (self._t_nrp - self._t_mnr) / self._t_mnr)

    def _couple_inlet(self, EU: float):
        """Arclength-column entries -d(value)/dEU of the inlet's Dirichlet rows
        (the last stored entry of each row: the column N-1)."""
        torch = self.torch
        rows, sens = inlet_sensitivity(self.mesh, self.params.WiN0 * EU, self.inflow)
        r = torch.from_numpy(rows).to(self.device)
        pos = self._rowptr[r + 1] - 1
        self.values[pos] = torch.from_numpy(-self.params.WiN0 * sens).to(self.device)

    def residual_norm(self, x, LN0: float, increment: int, dArN: float = 0.0, x_old=None, tangent=None,
                      dS: float = 0.0) -> float:
        """||b|| of the step's system at x (residual, arclength and Dirichlet
        rows; no Jacobian): how far a predicted state is from the solution."""
        torch = self.torch
        rows, vals = cylinder_dirichlet(self.mesh, self.params.WiN0 * float(x[-1]), self.inflow)
        self.asm.set_dirichlet(rows, vals)
        self.asm.assemble(x, "MNR", b=self.b)
        self.asm.arclength(x, self.b, None, increment=increment, LN0=LN0, dArN=dArN, x_old=x_old, tangent=tangent,
                           dS=dS, flag_nr="MNR")
        self.asm.apply_dirichlet(x, self.b, None, "MNR")
        return float(torch.linalg.vector_norm(self.b))


class ArclengthContinuation:
    """Arclength continuation in WiN = WiN0*EU (continuationStepArclength,
    src/Continuation_Arclength_mod.F90:77-377): increments 1-2 pin EU at
    LN0 + (increment-1)*dArN; then the direction is the secant (increment 2)
    or the normalised solution of J t = e_N with the last factorised Jacobian
    (increment >= 3, setDirectionVectors); the step dS is scaled by
    sqrt(eps_pr/err_pr) (at most doubled, at most dS_max) from the relative
    prediction error (adjustContinuationStep); the predictor is x + t*dS
    (updateSolutionArclength)."""

    def __init__(self, solver: NewtonSolver, LN0: float, dArN: float, dS0: float = 1.0, dS_max: float = 50.0,
                 eps_pr: float = 1.0e-3, predictor=None):
        self.s = solver
        self.LN0, self.dArN, self.dS, self.dS_max, self.eps_pr = LN0, dArN, dS0, dS_max, eps_pr
        self.increment = 0
        self.x = None       # current predictor / state (device)
        self.x_old = None   # last converged state
        self.tangent = None
        self.predictor = predictor  # optional learned predictor (learned.LearnedPredictor)
        self.max_iter = NITER
        self.cor_norm_div = COR_NORM_DIV
        self.choices = []   # per step: "tangent" or "learned" for the next prediction
        self.residuals = []  # per learned proposal: (||b|| at the tangent prediction, at the proposal)
        self._pending = None  # (features, tangent prediction) of the last step

    def start(self, x0):
        torch = self.s.torch
        self.x = torch.from_numpy(np.asarray(x0, dtype=np.float64)).to(self.s.device)

    def _field_errors(self, x, xp):
        """max over the 6 fields and EU of ||x - xp|| / ||x|| (per field)."""
        torch = self.s.torch
        N1 = self.s.N - 1
        d = (x[:N1] - xp[:N1]).view(-1, 6)
        v = x[:N1].view(-1, 6)
        e = torch.sqrt((d * d).sum(0) / (v * v).sum(0))
        eu = abs(float(x[-1] - xp[-1])) / abs(float(x[-1]))
        return max(float(e.max()), eu)

    def step(self):
        """One continuation step; returns (converged state, Newton records)."""
        torch = self.s.torch
        self.increment += 1
        k = self.increment
        predicted = self.x.clone()
        x, hist = self.s.solve(self.x, self.LN0, increment=k, dArN=self.dArN, x_old=self.x_old,
                               tangent=self.tangent, dS=self.dS, max_iter=self.max_iter,
                               cor_norm_div=self.cor_norm_div)
        self.dS = self.s.dS_used  # halved by each emergency restart (dSo is global in the reference)
        if k > 1:
            if k == 2:
                d = x - self.x_old
                self.tangent = d / torch.linalg.vector_norm(d)
            else:
                e = torch.zeros(self.s.N, dtype=torch.float64, device=self.s.device)
                e[-1] = 1.0
                # checked (and refined) whatever the factorisation's policy: a
                # carried FP32 factorisation is not otherwise checked here
                t = self.s.lu.solve(e, check=True)
                self.tangent = t / torch.linalg.vector_norm(t)
            err = self._field_errors(x, predicted)
            trns = math.sqrt(self.eps_pr / err) if err > 0 else 2.0
            trns = min(trns, 2.0)
            self.dS = min(self.dS * trns, self.dS_max)
        self.x_old = x
        self.x = x + self.tangent * self.dS if k > 1 else x.clone()
        choice = "tangent"
        if self.predictor is not None and k >= 2:
            if self._pending is not None:
                # the error the previous tangent prediction made
                self.predictor.observe(self._pending[0], x - self._pending[1])
            X = self.predictor.features(x, self.x)
            self._pending = (X, self.x.clone())
            if self.predictor.ready():
                cand = self.predictor.propose(X, self.x, self.tangent)
                nxt = k + 1
                r_t = self.s.residual_norm(self.x, self.LN0, nxt, self.dArN, self.x_old, self.tangent, self.dS)
                try:
                    r_l = self.s.residual_norm(cand, self.LN0, nxt, self.dArN, self.x_old, self.tangent, self.dS)
                except InflowError:  # the proposal's Wi has no fully developed inflow
                    r_l = math.inf
                self.residuals.append((r_t, r_l))
                if r_l < r_t:
                    self.x = cand
                    choice = "learned"
            self.predictor.train_pending()
        self.choices.append(choice)
        return x, hist
