#!/usr/bin/env python
"""Assembly benchmark: residual + forward-difference Jacobian assembly of the
100,000-element cylinder mesh on one MI355X.

A step is one full assembly (vef_assemble, Flag_NR = 'NRP'): element
residuals and their 56-evaluation forward-difference Jacobians, outlet-edge
traction, and the deterministic scatter into the fixed CSR pattern.  Inputs
(mesh, pattern, state) are resident in HBM before the timed region.

Workload: channel x in [-15, 15], y in [-2, 2], cylinder R = 1, 100,000 9-node
elements (401,848 nodes, N = 2,411,089 unknowns); Oldroyd-B (L-PTT with
VeN = 0), Wi = WiN0*EU = 0.5, beta = 0.59, Re = 0, DfN = 1e-3; state = smooth
divergence-free flow + 8-mode Fourier perturbation (amplitude 1e-2, seed
20261203) on every field.

Prints one JSON line (rank 0).  `--gpus N` with N > 1 exits: this build runs
on one GPU.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# Algorithmic operations of one element-residual evaluation (adds, muls,
# divisions, square roots; FMA-free) for this workload's configuration, from
# the operation-counting build of the CPU restatement; pinned by
# tests/test_flop_count.py.
FLOPS_PER_RESIDUAL_EVAL = 24426
# Of those, the geometry (basis2D, Helem**2, BvN/H2 and DfN/H2 at the 9 Gauss
# points) does not depend on the state: the residual pass computes it once per
# element and the Jacobian pass reuses it.
FLOPS_GEOMETRY_PER_EVAL = 2979
EVALS_PER_ELEMENT = 56            # base + 54 nodal + 1 arclength perturbation
QUOTIENT_FLOPS = 54 * 55 * 3      # ((dR - R)*h)/(h*h): sub, mul, div per entry
SCATTER_FLOPS = 2916 + 2 * 54     # one add per ADD_VALUES contribution
# The Jacobian kernel's algorithmic work per element: the 55 perturbed
# evaluations without their geometry, and the quotients.
JACOBIAN_KERNEL_FLOPS = 55 * (FLOPS_PER_RESIDUAL_EVAL - FLOPS_GEOMETRY_PER_EVAL) + QUOTIENT_FLOPS
# The whole assembly's: geometry once, 56 evaluations without it, quotients,
# ADD_VALUES additions.
STEP_FLOPS = (FLOPS_GEOMETRY_PER_EVAL + EVALS_PER_ELEMENT * (FLOPS_PER_RESIDUAL_EVAL - FLOPS_GEOMETRY_PER_EVAL)
              + QUOTIENT_FLOPS + SCATTER_FLOPS)
# Linear stability assembly (A and B): one residualBulkStability evaluation
# counts STAB_FLOPS_PER_EVAL operations, STAB_FLOPS_BASE of them for the base
# state's Gauss-point quantities, which the GPU computes once per element;
# each of the 54 forward differences repeats the rest except the test-node
# coefficients that depend on the base state only (STAB_COEF_FLOPS per Gauss
# point and test node, computed once per element).  Quotients: sub, mul, div
# per entry of A and B.  Pinned by tests/test_flop_count.py and
# tests/test_stability_host_math.py.
STAB_FLOPS_PER_EVAL = 35289
STAB_FLOPS_BASE = 9261
STAB_COEF_FLOPS = 79
STAB_QUOTIENT_FLOPS = 2 * 2916 * 3
# The FP32 analytic Jacobian's work per element (VEF_RES_JAC_FP32): 171 dual-
# number evaluations of the Gauss-point terms (9 Gauss points x 19
# directions, AD32_DUAL_FLOPS each), the field interpolations, the WET
# products, forming E (37 operations per (k, q, m) and Gauss point) and the
# outer contraction with the test functions.  Pinned by
# tests/test_ad_jacobian_host.py.
AD32_DUAL_FLOPS = 1064
AD32_FLOPS_PER_ELEMENT = 171 * AD32_DUAL_FLOPS + 9 * 22 * 9 * 2 + 171 * 18 + 9 * (162 * 37 + (2916 + 54) * 6)
PEAK_FP64_TFLOPS = 78.6           # MI355X FP64 (vector = matrix), FMA counted as 2
PEAK_FP32_TFLOPS = 157.3          # MI355X FP32 (vector = matrix), FMA counted as 2
PEAK_HBM_GBS = 8000.0
N_ELEMENTS = 100000


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--elements", type=int, default=N_ELEMENTS)
    ap.add_argument("--cpu-seconds", type=float, default=15.0, help="target CPU work of the baseline sample")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--newton", action="store_true",
                    help="measure seconds per converged arclength step (fixed-pattern GPU Newton + continuation on "
                         "the 1,600-element mesh) instead of the assembly throughput")
    ap.add_argument("--arc-steps", type=int, default=10, help="continuation steps for --newton (2 pinned + the rest)")
    ap.add_argument("--newton-model", default="L-PTT", choices=["L-PTT", "GIESEKUS"],
                    help="constitutive model of --newton (L-PTT with --newton-ven 0 is Oldroyd-B)")
    ap.add_argument("--newton-ven", type=float, default=0.0,
                    help="L-PTT epsilon or Giesekus alpha of --newton; other than Oldroyd-B, the inlet values come "
                         "from the fully developed inflow problem at each Wi")
    ap.add_argument("--continuation", action="store_true",
                    help="arclength continuation in Wi at full size (40,000 elements unless --elements) up to "
                         "--to-wi, for the modes in --modes")
    ap.add_argument("--to-wi", type=float, default=0.9)
    ap.add_argument("--max-steps", type=int, default=80)
    ap.add_argument("--modes", default="reference,fp32",
                    help="comma-separated: reference (FP64 reference policy), fp32 (inexact Newton, FP32 "
                         "Jacobian and factorisation), fp32_learned (and the learned predictor), fp32_krylov "
                         "(Newton-Krylov: the current FP32 Jacobian solved by GMRES preconditioned with the "
                         "carried FP32 factorisation)")
    ap.add_argument("--ds-max", type=float, default=50.0,
                    help="largest arclength step (the reference's dS_max, src/Parameters_mod.F90:27: an absolute "
                         "length in the state's 2-norm, which grows as the square root of the unknowns)")
    ap.add_argument("--factorization", default="mf", choices=["mf", "band"],
                    help="linear solver of --continuation: multifrontal (nested dissection) or banded LU")
    ap.add_argument("--mode-seconds", type=float, default=600.0,
                    help="wall-time budget of each --continuation mode (the mode stops at the next step)")
    ap.add_argument("--refine-tol", type=float, default=None,
                    help="continuation: relative residual to which the FP32 modes' solves are refined by GMRES "
                         "(default 1e-3)")
    ap.add_argument("--check-every", type=int, default=10,
                    help="--continuation: every this many steps, an FP64 Newton solve at the step's Wi checks the "
                         "drag of the non-reference modes")
    ap.add_argument("--inlet-coupling", action="store_true",
                    help="--continuation: the inlet rows' dependence on EU in the Jacobian (NewtonSolver "
                         "inlet_coupling; not the reference's Newton)")
    ap.add_argument("--learned-epochs", type=int, default=50)
    ap.add_argument("--learned-nodes", type=int, default=0,
                    help="nodes sampled per training step of the learned predictor (0: all)")
    ap.add_argument("--stability", action="store_true",
                    help="linear stability assembly (A and B) of the 40,000-element mesh instead of the Newton "
                         "assembly")
    ap.add_argument("--base", default="converged", choices=["converged", "synthetic"],
                    help="--stability: the base state — a converged steady state at --base-wi (FP32 inexact "
                         "continuation, then an FP64 Newton solve pinned at that Wi) with the 10 leading eigenvalues "
                         "of the GPU and of the oracle-assembled matrices, or the smooth synthetic state")
    ap.add_argument("--base-wi", type=float, default=0.6)
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain 1 GiB device-to-device copy three times (the reference launch for "
                         "scaling the FETCH_SIZE / WRITE_SIZE counters) and exit")
    return ap.parse_args()


def copy_calibration():
    """A plain copy kernel of known size: 1 GiB read + 1 GiB written per launch."""
    import torch

    n = 128 * 1024 * 1024
    src = torch.ones(n, dtype=torch.float64, device="cuda:0")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_bytes_read": n * 8, "copy_calibration_bytes_written": n * 8}), flush=True)


def newton_bench(args):
    """Arclength continuation in Wi on the 1,600-element mesh: two pinned steps
    (Wi = 0.1, 0.12), then pseudo-arclength steps; value = mean wall time of a
    converged pseudo-arclength step (assembly, arclength and Dirichlet rows,
    GPU factorisation and solves, tangent, predictor) of the reference's
    Newton policy in FP64 (a fresh Jacobian at the start of every step) with
    the banded LU, the state resident on the GPU.  Reported beside it: the same
    policy with the two sparse factorisations, and the inexact Newton (the
    last factorisation carried across steps) alone, with the learned
    predictor, with the FP32 analytic Jacobian, with both, and with both and
    an FP32 (sparse) factorisation; each with its Wi gained per
    second and the relative drag difference to an FP64 Newton solve at its
    final Wi.  The CPU baseline runs the reference policy with the oracle's
    assembly and SuperLU on the host cores."""
    import numpy as np
    import torch

    from ve_flow_around_cylinder_fortran_amd import Params, workload_mesh
    from ve_flow_around_cylinder_fortran_amd.learned import LearnedPredictor
    from ve_flow_around_cylinder_fortran_amd.mesh import boundary_edges
    from ve_flow_around_cylinder_fortran_amd.newton import ArclengthContinuation, NewtonSolver, initial_state
    from ve_flow_around_cylinder_fortran_amd.post import peak_wall_stress

    pset = dict(ReN=0.0, BvN=0.59, VeN=args.newton_ven, DfN=1.0e-3, WiN0=1.0, model=args.newton_model)
    oldroyd_b = pset["VeN"] == 0.0 and pset["model"] == "L-PTT"
    model_name = "Oldroyd-B" if oldroyd_b else (f"linear PTT eps={pset['VeN']}" if pset["model"] == "L-PTT"
                                                else f"Giesekus alpha={pset['VeN']}")
    LN0, dArN = 0.1, 0.02
    mesh = workload_mesh(1600)
    x0 = initial_state(mesh, LN0)
    edges = boundary_edges(mesh, "Cylinder")

    def run(factorization="band", reuse=False, learned=False, precision="f64", jacobian="fd64", overlap=False):
        solver = NewtonSolver(mesh, Params(**pset), factorization=factorization, reuse_factors=reuse,
                              linear_precision=precision, jacobian=jacobian)
        cont = ArclengthContinuation(solver, LN0=LN0, dArN=dArN,
                                     predictor=LearnedPredictor(mesh, seed=0, epochs=50, overlap=overlap)
                                     if learned else None)
        cont.start(x0)
        times, flags, EU = [], [], []
        for _ in range(args.arc_steps):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            x, hist = cont.step()
            torch.cuda.synchronize()
            times.append(time.perf_counter() - t0)
            flags.append("".join(h.flag[0] for h in hist))
            EU.append(float(x[-1]))
        solver.asm.set_drag_edges(edges)
        Fd = float(solver.asm.drag(x).item())
        Pw = peak_wall_stress(mesh, solver.asm.nodal_stress(x, pset["WiN0"] * float(x[-1])).cpu().numpy())
        # FP64 Newton (reference policy) at the same Wi, started from this state
        ref = NewtonSolver(mesh, Params(**pset))
        xr, _ = ref.solve(x, float(x[-1]))
        ref.asm.set_drag_edges(edges)
        Fr = float(ref.asm.drag(xr).item())
        Pr = peak_wall_stress(mesh, ref.asm.nodal_stress(xr, pset["WiN0"] * float(xr[-1])).cpu().numpy())
        arc = times[2:]
        return {"s_per_step": float(np.mean(arc)), "step_seconds": times, "newton_flags_per_step": flags,
                "EU_per_step": EU, "Wi_gain_per_second": (EU[-1] - EU[1]) * pset["WiN0"] / sum(arc),
                "factorizations": solver.lu.refactorizations, "host_fallbacks": solver.lu.host_fallbacks,
                "solver": type(solver.lu).__name__,
                "refined_solves": getattr(solver.lu, "refinements", 0),
                "gmres_iterations": getattr(solver.lu, "gmres_iterations", 0),
                "Fd_last_step": Fd, "drag_rel_diff_vs_fp64_newton": abs(Fd - Fr) / abs(Fr),
                "peak_wall_Txx_last_step": Pw, "peak_wall_Txx_rel_diff_vs_fp64_newton": abs(Pw - Pr) / abs(Pr),
                "choices": cont.choices if learned else None, "factor_pattern_nnz":
                    getattr(solver.lu, "nnzS", None) or getattr(solver.lu, "nnzT", None)}

    plain = run()
    modes = {
        "reference_policy_hybrid_lu": run(factorization="hybrid"),
        "reference_policy_scalar_order_lu": run(factorization="rf"),
        "inexact_newton": run(reuse=True),
        "inexact_newton_learned": run(reuse=True, learned=True),
        "inexact_newton_fp32_jacobian": run(reuse=True, jacobian="ad32"),
        "inexact_newton_learned_fp32_jacobian": run(reuse=True, learned=True, jacobian="ad32"),
        "inexact_newton_fp32_jacobian_fp32_factorization": run(reuse=True, precision="f32", jacobian="ad32"),
        "inexact_newton_learned_fp32_jacobian_fp32_factorization": run(
            reuse=True, learned=True, precision="f32", jacobian="ad32"),
        "inexact_newton_learned_overlapped_training_fp32_jacobian_fp32_factorization": run(
            reuse=True, learned=True, precision="f32", jacobian="ad32", overlap=True),
        "inexact_newton_learned_fp32_jacobian_fp32_factorization_hybrid_lu": run(
            factorization="hybrid", reuse=True, learned=True, precision="f32", jacobian="ad32"),
    }
    for m in modes.values():
        m["speedup_per_step_vs_reference_policy"] = plain["s_per_step"] / m["s_per_step"]
        m["Wi_per_second_vs_reference_policy"] = m["Wi_gain_per_second"] / plain["Wi_gain_per_second"]
    out = {
        "metric": "s per converged arclength step",
        "value": plain["s_per_step"],
        "unit": "s",
        "n_gpus": 1,
        "steps": args.arc_steps - 2,
        "warmup": 2,
        "ms_per_step": plain["s_per_step"] * 1e3,
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": f"cylinder-in-channel Q2 mesh, 1600 elements, {model_name} beta=0.59 Re=0 DfN=1e-3, "
                        + ("" if oldroyd_b else "fully developed inflow solved at each Wi, ") + "arclength "
                        f"continuation in Wi from {LN0} (pinned steps {LN0}, {LN0 + dArN}), Newton tolerance 1e-8 on "
                        "||b|| and ||s||; reference Newton policy, FP64, banded LU in reverse Cuthill-McKee node "
                        "order (libvef_band.so)",
            "elements": mesh.n_elements, "unknowns": mesh.n_unknowns,
            "reference_policy": plain,
            "modes": modes,
        },
    }
    if not args.no_cpu_baseline:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle
        from newton_cpu import continuation_cpu

        threads = max(1, min(int(os.environ.get("OMP_NUM_THREADS", "0")) or min(16, os.cpu_count() or 1), 16))
        n_cpu = min(args.arc_steps, 6)
        cpu_times = []
        from ve_flow_around_cylinder_fortran_amd.inflow import FullyDevelopedInflow

        states, _ = continuation_cpu(oracle, mesh, pset, x0, LN0, dArN, n_cpu, nthreads=threads, times=cpu_times,
                                     inflow=None if oldroyd_b else FullyDevelopedInflow.from_mesh(mesh, Params(**pset)))
        cpu_arc = float(np.mean(cpu_times[2:]))
        out["cpu_baseline"] = {"value": cpu_arc, "unit": "s", "cores": threads, "kind": "port",
                               "sample": f"the same continuation, steps 3..{n_cpu}: oracle assembly on {threads} "
                                         "threads, SuperLU (minimum degree on A+A^T) on the host"}
        out["config"]["EU_cpu_last"] = float(states[-1][-1])
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


def continuation_bench(args):
    """Arclength continuation in Wi at the configuration's full size (the
    40,000-element mesh by default), as loopArclength runs it
    (src/Continuation_Arclength_mod.F90:62-128): two pinned steps (Wi = 0.1,
    0.12), then pseudo-arclength steps until Wi reaches --to-wi (or
    --max-steps).  Modes: "reference" = the reference's Newton policy in FP64
    (a fresh forward-difference Jacobian at the start of every step, MNR
    iterations on its factorisation), "fp32" = the inexact Newton with the
    FP32 analytic Jacobian and FP32 factorisation carried across steps (FP64
    residual, update and convergence test), "fp32_learned" = the same with the
    learned predictor, "fp32_krylov" = the inexact Newton whose iterations on
    a carried factorisation solve the current FP32 Jacobian by GMRES
    preconditioned with it.  Linear solves: the multifrontal LU in nested-
    dissection order.  Per step: Wi, Newton flags, seconds, drag Fd and peak
    wall Txx; per mode, an FP64 reference-policy Newton solve at the mode's
    final Wi from its final state gives the drag and wall-stress errors.
    value = mean seconds per pseudo-arclength step of the first mode."""
    import numpy as np
    import torch

    from ve_flow_around_cylinder_fortran_amd import Params, workload_mesh
    from ve_flow_around_cylinder_fortran_amd.mesh import boundary_edges
    from ve_flow_around_cylinder_fortran_amd.newton import ArclengthContinuation, NewtonSolver, initial_state
    from ve_flow_around_cylinder_fortran_amd.post import peak_wall_stress

    pset = dict(ReN=0.0, BvN=0.59, VeN=args.newton_ven, DfN=1.0e-3, WiN0=1.0, model=args.newton_model)
    oldroyd_b = pset["VeN"] == 0.0 and pset["model"] == "L-PTT"
    model_name = "Oldroyd-B" if oldroyd_b else (f"linear PTT eps={pset['VeN']}" if pset["model"] == "L-PTT"
                                                else f"Giesekus alpha={pset['VeN']}")
    n_el = 40000 if args.elements == N_ELEMENTS else args.elements
    LN0, dArN = 0.1, 0.02
    t_mesh = time.perf_counter()
    mesh = workload_mesh(n_el)
    x0 = initial_state(mesh, LN0)
    edges = boundary_edges(mesh, "Cylinder")
    print(f"[continuation] {n_el} elements, {model_name}: mesh {time.perf_counter() - t_mesh:.1f} s", flush=True)

    def drag_and_wall(solver, x):
        solver.asm.set_drag_edges(edges)
        Fd = float(solver.asm.drag(x).item())
        Pw = peak_wall_stress(mesh, solver.asm.nodal_stress(x, pset["WiN0"] * float(x[-1])).cpu().numpy())
        return Fd, Pw

    def run(mode):
        learned = mode.endswith("learned")
        fp32 = mode.startswith("fp32")
        krylov = "krylov" in mode
        t_setup = time.perf_counter()
        solver = NewtonSolver(mesh, Params(**pset), factorization=args.factorization, reuse_factors=fp32,
                              linear_precision="f32" if fp32 else "f64", jacobian="ad32" if fp32 else "fd64",
                              inlet_coupling=args.inlet_coupling, reuse_max_iters="auto", krylov=krylov,
                              refine_tol=args.refine_tol if fp32 else None)
        predictor = None
        if learned:
            from ve_flow_around_cylinder_fortran_amd.learned import LearnedPredictor

            predictor = LearnedPredictor(mesh, seed=0, epochs=args.learned_epochs, overlap=True,
                                         sample_nodes=args.learned_nodes)
        def progress(rec, n, xF):  # long solves (restarts) report as they go
            if n > 12 and n % 5 == 0:
                print(f"[continuation] {mode}   iteration {n}: {rec.flag} ||b|| {rec.res_norm:.2e} "
                      f"||s|| {rec.cor_norm:.2e} xF {xF:g}", flush=True)
            if time.perf_counter() - t_mode > args.mode_seconds:
                raise TimeoutError(f"time budget of {args.mode_seconds:g} s spent inside a step")

        solver.on_iteration = progress
        cont = ArclengthContinuation(solver, LN0=LN0, dArN=dArN, predictor=predictor, dS_max=args.ds_max)
        cont.start(x0)
        setup = time.perf_counter() - t_setup
        steps = []
        failure = None
        x = None
        checks = []
        t_mode = time.perf_counter()
        for k in range(args.max_steps):
            if time.perf_counter() - t_mode > args.mode_seconds:
                failure = {"step": k + 1, "error": f"time budget of {args.mode_seconds:g} s spent"}
                print(f"[continuation] {mode}: time budget spent at step {k + 1}", flush=True)
                break
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            try:
                x, hist = cont.step()
            except Exception as e:  # recorded with what the solver saw; the mode ends here
                h = getattr(solver, "last_history", [])
                vals = solver.values
                failure = {"step": k + 1, "error": f"{type(e).__name__}: {e}",
                           "newton": [(r.flag, r.res_norm, r.cor_norm) for r in h],
                           "values_finite": bool(torch.isfinite(vals).all()),
                           "b_finite": bool(torch.isfinite(solver.b).all()),
                           "predictor_Wi": pset["WiN0"] * float(cont.x[-1]), "dS": cont.dS}
                print(f"[continuation] {mode} step {k + 1} failed: {failure}", flush=True)
                break
            torch.cuda.synchronize()
            dt = time.perf_counter() - t0
            Fd, Pw = drag_and_wall(solver, x)
            Wi = pset["WiN0"] * float(x[-1])
            steps.append({"Wi": Wi, "seconds": dt, "flags": "".join(h.flag[0] for h in hist),
                          "cor_norms": [h.cor_norm for h in hist],
                          "newton_iterations": len(hist), "Fd": Fd, "peak_wall_Txx": Pw, "dS": cont.dS,
                          "restarts": len(solver.emergencies), "lu_check_residual": solver.lu.last_residual,
                          "refined_solves": solver.lu.refinements})
            print(f"[continuation] {mode} step {k + 1}: Wi {Wi:.5f} {dt:.3f} s flags {steps[-1]['flags']} "
                  f"Fd {Fd:.10f} lu residual {solver.lu.last_residual} refined {solver.lu.refinements}", flush=True)
            if mode != "reference" and (k + 1) % args.check_every == 0:
                # FP64 reference-policy Newton at this step's Wi from its state
                if check_solver[0] is None:
                    check_solver[0] = NewtonSolver(mesh, Params(**pset), factorization=args.factorization,
                                                   inlet_coupling=args.inlet_coupling)
                xr, hr = check_solver[0].solve(x, float(x[-1]))
                Fr, Pr = drag_and_wall(check_solver[0], xr)
                checks.append({"step": k + 1, "Wi": Wi, "Fd": Fd, "Fd_fp64": Fr,
                               "drag_rel_diff": abs(Fd - Fr) / abs(Fr),
                               "peak_wall_Txx_rel_diff": abs(Pw - Pr) / abs(Pr), "fp64_iterations": len(hr)})
                print(f"[continuation] {mode} step {k + 1}: FP64 check |dFd|/Fd {checks[-1]['drag_rel_diff']:.2e} "
                      f"({len(hr)} iterations)", flush=True)
            if Wi >= args.to_wi:
                break
        # FP64 reference policy at the final Wi, from the final state
        x = cont.x_old
        ref = NewtonSolver(mesh, Params(**pset), factorization="mf")
        xr, hr = ref.solve(x, float(x[-1]))
        Fr, Pr = drag_and_wall(ref, xr)
        arc = [s["seconds"] for s in steps[2:]]
        out = {"s_per_step": float(np.mean(arc)) if arc else None, "steps": len(steps), "setup_seconds": setup,
               "Wi_final": steps[-1]["Wi"], "total_seconds": float(sum(s["seconds"] for s in steps)),
               "newton_iterations": int(sum(s["newton_iterations"] for s in steps)),
               "s_per_step_median": float(np.median(arc)) if arc else None,
               "factorizations": solver.lu.refactorizations, "refined_solves": solver.lu.refinements,
               "gmres_iterations": solver.lu.gmres_iterations, "krylov_solves": solver.krylov_solves,
               "krylov_failures": solver.krylov_failures, "escalations": solver.escalations,
               "host_fallbacks": solver.lu.host_fallbacks,
               "Fd_final": steps[-1]["Fd"], "fp64_newton_iterations_at_final_Wi": len(hr),
               "drag_rel_diff_vs_fp64_newton": abs(steps[-1]["Fd"] - Fr) / abs(Fr),
               "peak_wall_Txx_rel_diff_vs_fp64_newton": abs(steps[-1]["peak_wall_Txx"] - Pr) / abs(Pr),
               "per_step": steps, "failure": failure, "fp64_checks": checks,
               "max_drag_rel_diff_fp64_checks": max([c["drag_rel_diff"] for c in checks], default=None)}
        if learned:
            out["choices"] = cont.choices
            # per proposal: ||b|| at the tangent prediction and at the learned proposal
            out["proposal_residuals"] = [[float(a), float(b)] for a, b in cont.residuals]
            out["training_losses"] = [float(v) for v in predictor.losses]
        del solver, ref, cont
        torch.cuda.empty_cache()
        return out

    check_solver = [None]  # FP64 Newton for the periodic checks (built on first use, shared by the modes)
    modes = {m: run(m) for m in args.modes.split(",")}
    first = next(iter(modes.values()))
    for m in modes.values():
        m["speedup_per_step_vs_first"] = first["s_per_step"] / m["s_per_step"] if m["s_per_step"] else None
        m["speedup_total_vs_first"] = first["total_seconds"] / m["total_seconds"]
    out = {
        "metric": "s per converged arclength step",
        "value": first["s_per_step"],
        "unit": "s",
        "n_gpus": 1,
        "steps": first["steps"] - 2,
        "warmup": 2,
        "ms_per_step": first["s_per_step"] * 1e3,
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": f"cylinder-in-channel Q2 mesh, {mesh.n_elements} elements, {model_name} beta=0.59 Re=0 "
                        "DfN=1e-3, " + ("" if oldroyd_b else "fully developed inflow solved at each Wi, ")
                        + f"arclength continuation in Wi from {LN0} to {args.to_wi} (pinned steps {LN0}, "
                        f"{LN0 + dArN}; dS0 = 1, dS_max = {args.ds_max:g}, eps_pr = 1e-3), Newton tolerance 1e-8 "
                        f"on ||b|| and ||s||, {'multifrontal' if args.factorization == 'mf' else 'banded'} LU",
            "elements": mesh.n_elements, "unknowns": mesh.n_unknowns,
            "inlet_coupling": bool(args.inlet_coupling),
            "modes": modes,
        },
        "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)


def stability_flops_per_element():
    coef = 81 * STAB_COEF_FLOPS
    return STAB_FLOPS_BASE + coef + 54 * (STAB_FLOPS_PER_EVAL - STAB_FLOPS_BASE - coef) + STAB_QUOTIENT_FLOPS


def stability_bench(args):
    """A and B of the linear stability problem (Cylinder_Stability_2D) around a
    base state at Wi = 0.6 on the 40,000-element mesh; a step is one
    vef_stability_assemble (element kernel, ordered sums of shared node pairs,
    Dirichlet rows).  The reference's A and B do not depend on the spanwise
    wavenumber (its spanwise terms are commented out,
    src/Stability/Element_Calculations_Stability_mod.F90:51-81), so one
    assembly serves every wavenumber of a batch."""
    import numpy as np  # noqa: F401
    import torch

    from ve_flow_around_cylinder_fortran_amd import Params, smooth_state, workload_mesh
    from ve_flow_around_cylinder_fortran_amd.stability import StabilityAssembler

    n_el = 40000 if args.elements == N_ELEMENTS else args.elements
    ven = args.newton_ven if args.newton_ven > 0.0 else 0.25
    pset = dict(ReN=0.0, BvN=0.59, VeN=ven, DfN=1.0e-3, WiN0=1.0, model=args.newton_model)
    model_name = (f"linear PTT epsilon={ven}" if args.newton_model == "L-PTT" else f"Giesekus alpha={ven}")
    mesh = workload_mesh(n_el)
    base_info = None
    if args.base == "converged":
        xd, base_info = converged_base(mesh, pset, args.base_wi)
        x = xd.cpu().numpy()
    else:
        x = smooth_state(mesh, EU=args.base_wi)
        xd = torch.from_numpy(x).to("cuda:0")
    sa = StabilityAssembler(mesh, Params(**pset))
    nnz = len(sa.col)
    A = torch.empty(nnz, dtype=torch.float64, device="cuda:0")
    B = torch.empty(nnz, dtype=torch.float64, device="cuda:0")
    print("[stability] assembling A and B", flush=True)
    for _ in range(args.warmup):
        sa.assemble(xd, A, B)
    torch.cuda.synchronize()
    sa.asm.profile(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        sa.assemble(xd, A, B)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    kern_ms, rest_ms, launches = sa.asm.profile_read()
    sa.asm.profile(False)
    ms_step = (t1 - t0) * 1e3 / args.steps
    ne = mesh.n_elements
    kern_ms_launch = kern_ms / max(launches, 1)
    flops = ne * stability_flops_per_element()
    achieved = flops / (kern_ms_launch * 1e-3) / 1e12
    out = {
        "metric": "elements/s stability A+B assembly",
        "value": ne / (ms_step * 1e-3),
        "unit": "elements/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic mesh; converged base state" if base_info else "synthetic",
        "config": {
            "workload": f"cylinder-in-channel Q2 mesh, {ne} elements, linear stability A and B "
                        f"(Cylinder_Stability_2D) around a {'converged' if base_info else 'smooth synthetic'} base "
                        f"state at Wi={args.base_wi}, {model_name} beta=0.59 Re=0 DfN=1e-3, Dirichlet rows applied",
            "base_state": base_info,
            "elements": ne, "unknowns": 6 * mesh.n_nodes, "nnz": nnz,
            "wavenumbers": "A and B do not depend on the spanwise wavenumber in the reference (spanwise terms "
                           "commented out); one assembly serves a batch of 16",
            "element_kernel_ms": kern_ms_launch,
            "element_kernels": "vef_stab_prep_kernel + vef_stab_kernel",
            "scatter_ms": rest_ms / max(launches, 1),
            "flops_per_element": stability_flops_per_element(),
        },
        "roofline": {
            "kernel": "vef_stab_prep_kernel + vef_stab_kernel (base state and test-node coefficients once per "
                      "element, then 54 forward differences per element, A and B)",
            "bound": "fp64 valu",
            "ceiling": "FP64 compute (MI355X FP64 vector peak = FP64 matrix peak = 78.6 TFLOP/s; no FMA "
                       "contraction, so 0.5 of peak is its ceiling)",
            "achieved": achieved,
            "peak": PEAK_FP64_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / PEAK_FP64_TFLOPS,
            "traffic": stability_traffic_from_profiles(ne),
            "flops_per_launch": flops,
        },
    }
    if base_info is not None:
        out["config"]["eigenvalues"] = stability_eigenvalues(mesh, pset, x, sa, A, B)
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = stability_cpu_baseline(mesh, x, pset, sa.rowptr, sa.col, args.cpu_seconds)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


def converged_base(mesh, pset, Wi):
    """A converged steady state at WiN = Wi: arclength continuation from Wi =
    0.1 with the FP32 inexact Newton (exact inlet coupling) until Wi is
    passed, then an FP64 Newton solve with the reference policy pinned at Wi
    from the nearest converged state."""
    import torch

    from ve_flow_around_cylinder_fortran_amd import Params
    from ve_flow_around_cylinder_fortran_amd.newton import ArclengthContinuation, NewtonSolver, initial_state

    t0 = time.perf_counter()
    solver = NewtonSolver(mesh, Params(**pset), factorization="mf", reuse_factors=True, linear_precision="f32",
                          jacobian="ad32", inlet_coupling=True, reuse_max_iters="auto")
    def progress(rec, n, xF):  # long solves (restarts) report as they go
        if n % 20 == 0:
            print(f"[stability]   iteration {n}: {rec.flag} ||b|| {rec.res_norm:.2e} ||s|| {rec.cor_norm:.2e} "
                  f"xF {xF:g}", flush=True)

    solver.on_iteration = progress
    cont = ArclengthContinuation(solver, LN0=0.1, dArN=0.02)
    cont.start(initial_state(mesh, 0.1))
    near, steps = None, 0
    while True:
        x, _ = cont.step()
        steps += 1
        w = pset["WiN0"] * float(x[-1])
        print(f"[stability] continuation step {steps}: Wi {w:.5f} ({time.perf_counter() - t0:.1f} s)", flush=True)
        if near is None or abs(w - Wi) < abs(pset["WiN0"] * float(near[-1]) - Wi):
            near = x
        if w >= Wi or steps >= 200:
            break
    print(f"[stability] continuation to Wi {w:.4f} in {steps} steps, {time.perf_counter() - t0:.1f} s", flush=True)
    del solver, cont
    ref = NewtonSolver(mesh, Params(**pset), factorization="mf")
    xb, hist = ref.solve(near, Wi / pset["WiN0"])
    torch.cuda.synchronize()
    info = {"continuation_steps": steps, "continuation_Wi": w, "fp64_newton_iterations": len(hist),
            "final_residual_norm": hist[-1].res_norm, "final_correction_norm": hist[-1].cor_norm,
            "seconds": time.perf_counter() - t0}
    print(f"[stability] FP64 Newton at Wi {Wi}: {len(hist)} iterations ({info['seconds']:.1f} s)", flush=True)
    del ref
    torch.cuda.empty_cache()
    return xb, info


def mirror_map(mesh):
    """Node of each node's mirror image in y = 0 (the channel's axis)."""
    import numpy as np

    key = np.round(mesh.coords * 1e9).astype(np.int64)
    order = np.lexsort((key[:, 1], key[:, 0]))
    mk = key.copy()
    mk[:, 1] = -mk[:, 1]
    morder = np.lexsort((mk[:, 1], mk[:, 0]))
    m = np.empty(len(key), dtype=np.int64)
    m[morder] = order
    if not np.array_equal(key[m][:, 0], key[:, 0]) or not np.array_equal(key[m][:, 1], -key[:, 1]):
        return None
    return m


def stability_eigenvalues(mesh, pset, x, sa, A, B, k=10):
    """The k eigenvalues nearest 0 of A q = lambda B q (shift-invert Arnoldi
    on the GPU multifrontal LU) for the GPU matrices and for the oracle's,
    their largest relative difference, and the mirror symmetry of the leading
    mode (Ux, P, Sxx, Syy even and Uy, Sxy odd in y: symmetric)."""
    import numpy as np
    import torch

    from ve_flow_around_cylinder_fortran_amd.stability import leading_eigenvalues_gpu, stability_dirichlet_rows

    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle

    t0 = time.perf_counter()
    say = lambda n: print(f"[stability] eigen solve: {n} operator applications, {time.perf_counter() - t0:.0f} s",  # noqa: E731
                          flush=True)
    lam, V, lu, napply = leading_eigenvalues_gpu(A, B, sa.rowptr, sa.col, mesh, k=k, vectors=True, tol=1e-10,
                                                 progress=say, maxiter=2000)
    t_gpu = time.perf_counter() - t0
    print(f"[stability] GPU matrices: {napply} applications, {t_gpu:.1f} s; the oracle's matrices", flush=True)
    oA, oB = oracle.stab_assemble(oracle.params(**pset), mesh, x, sa.rowptr, sa.col, nthreads=16)
    oracle.stab_dirichlet(sa.rowptr, sa.col, oA, oB, stability_dirichlet_rows(mesh))
    bitwise = bool(np.array_equal(oA, A.cpu().numpy()) and np.array_equal(oB, B.cpu().numpy()))
    lam_o, _, _, _ = leading_eigenvalues_gpu(torch.from_numpy(oA).cuda(), torch.from_numpy(oB).cuda(), sa.rowptr,
                                             sa.col, mesh, k=k, lu=lu, tol=1e-10, progress=say, maxiter=2000)
    rel = max(float(np.abs(lam - lo).min() / abs(lo)) for lo in lam_o)
    sym = None
    m = mirror_map(mesh)
    if m is not None:
        q = V[:, 0].reshape(-1, 6)
        D = np.array([1, -1, 1, 1, -1, 1], dtype=float)
        qm = q[m] * D
        nq = np.linalg.norm(q)
        sym = {"symmetric_part": float(np.linalg.norm(q + qm) / (2 * nq)),
               "antisymmetric_part": float(np.linalg.norm(q - qm) / (2 * nq))}
    out = {"k": k, "gpu": [[float(v.real), float(v.imag)] for v in lam],
           "oracle_matrices": [[float(v.real), float(v.imag)] for v in lam_o],
           "max_rel_diff": rel, "matrices_bitwise_equal_oracle": bitwise, "operator_applications": napply,
           "seconds_gpu_eigensolve": t_gpu, "leading_mode_mirror": sym}
    print(f"[stability] eigenvalues {lam[:4]} ... max relative difference to the oracle matrices' {rel:.1e}; "
          f"leading mode {sym}", flush=True)
    return out


def stability_cpu_baseline(mesh, x, pset, rowptr, col, target_s):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle

    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or min(16, os.cpu_count() or 1)
    threads = max(1, min(threads, 16))
    prm = oracle.params(**pset)
    probe = min(1000, mesh.n_elements)
    t0 = time.perf_counter()
    oracle.stab_assemble(prm, mesh, x, rowptr, col, el_range=(0, probe), nthreads=threads)
    rate = probe / (time.perf_counter() - t0)
    n = int(min(mesh.n_elements, max(probe, rate * target_s)))
    t0 = time.perf_counter()
    oracle.stab_assemble(prm, mesh, x, rowptr, col, el_range=(0, n), nthreads=threads)
    dt = time.perf_counter() - t0
    return {"value": n / dt, "unit": "elements/s", "cores": threads, "kind": "port",
            "sample": f"elements 0..{n} of the same mesh and base state ({n} of {mesh.n_elements}), A and B by "
                      f"forward differences + serial ADD_VALUES scatter, {dt:.1f} s"}


def stability_traffic_from_profiles(n_elements):
    path = os.path.join(ROOT, "profiles", "vef_stab_kernel_pmc_40k.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        d = json.load(f)
    if d.get("elements") != n_elements:
        return None
    return d.get("hbm_bytes_per_launch")


def cpu_baseline(mesh, x, target_s):
    """The CPU restatement (oracle) on the host cores, reference order: element
    work on OpenMP threads, the ADD_VALUES scatter serial."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle

    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or min(16, os.cpu_count() or 1)
    threads = max(1, min(threads, 16))
    prm = oracle.params(ReN=0.0, BvN=0.59, VeN=0.0, DfN=1e-3, WiN0=1.0)
    rowptr, col = oracle.pattern(mesh)
    probe = min(2000, mesh.n_elements)
    t0 = time.perf_counter()
    oracle.assemble(prm, mesh, x, rowptr=rowptr, col=col, el_range=(0, probe), nthreads=threads)
    rate = probe / (time.perf_counter() - t0)
    n = int(min(mesh.n_elements, max(probe, rate * target_s)))
    t0 = time.perf_counter()
    oracle.assemble(prm, mesh, x, rowptr=rowptr, col=col, el_range=(0, n), nthreads=threads)
    dt = time.perf_counter() - t0
    return {"value": n / dt, "unit": "elements/s", "cores": threads, "kind": "port",
            "sample": f"elements 0..{n} of the same mesh and state ({n} of {mesh.n_elements}), residual + "
                      f"FD Jacobian + serial ADD_VALUES scatter, {dt:.1f} s"}


def traffic_from_profiles(n_elements, key="hbm_bytes_per_launch"):
    """A figure of the committed counter profile of this workload (FETCH_SIZE
    and WRITE_SIZE, scaled by a plain copy's): the Jacobian kernel's HBM bytes
    per launch, with key="step_hbm_bytes" a whole NRP assembly's, with
    key="launches_per_assembly" the Jacobian pass's launches (colours)."""
    path = os.path.join(ROOT, "profiles", "vef_bulk_kernel_pmc_100k.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        d = json.load(f)
    if d.get("elements") != n_elements:
        return None
    return d.get(key)


def main():
    args = parse()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus > 1 or world > 1:
        print("bench.py: this assembler runs on one GPU; multi-GPU runs are not part of this build "
              f"(asked for --gpus {args.gpus}, WORLD_SIZE {world})", file=sys.stderr)
        sys.exit(2)
    if args.copy_calibration:
        copy_calibration()
        return
    if args.continuation:
        continuation_bench(args)
        return
    if args.newton:
        newton_bench(args)
        return
    if args.stability:
        stability_bench(args)
        return

    import numpy as np  # noqa: F401
    import torch

    from ve_flow_around_cylinder_fortran_amd import Assembler, Params, smooth_state, workload_mesh

    mesh = workload_mesh(args.elements)
    x = smooth_state(mesh, EU=0.5)
    params = Params(ReN=0.0, BvN=0.59, VeN=0.0, DfN=1.0e-3, WiN0=1.0, model="L-PTT")
    asm = Assembler(mesh, params)
    xd = torch.from_numpy(x).to("cuda:0")
    b = torch.empty(asm.N, dtype=torch.float64, device="cuda:0")
    v = torch.empty(asm.nnz, dtype=torch.float64, device="cuda:0")
    for _ in range(args.warmup):
        asm.assemble(xd, "NRP", b=b, values=v)
    torch.cuda.synchronize()

    stream = torch.cuda.current_stream()
    marks = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps + 1)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    marks[0].record(stream)
    for i in range(args.steps):
        asm.assemble(xd, "NRP", b=b, values=v)
        marks[i + 1].record(stream)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    step_ms = sorted(marks[i].elapsed_time(marks[i + 1]) for i in range(args.steps))
    # kernel durations from a separate, untimed pass (the library records HIP
    # events around its kernels on the launch stream while profiling is on)
    asm.profile(True)
    for _ in range(args.steps):
        asm.assemble(xd, "NRP", b=b, values=v)
    bulk_ms, scatter_ms, launches = asm.profile_read()
    asm.profile(False)
    pct = lambda q: step_ms[min(len(step_ms) - 1, int(round(q * (len(step_ms) - 1))))]  # noqa: E731

    # residual-only assemblies (Flag_NR = 'MNR'), reported beside
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    for _ in range(args.steps):
        asm.assemble(xd, "MNR", b=b)
    torch.cuda.synchronize()
    t3 = time.perf_counter()

    # the same assembly with the FP32 analytic Jacobian (VEF_RES_JAC_FP32), reported beside
    v32 = torch.empty_like(v)
    for _ in range(args.warmup):
        asm.assemble(xd, "NRP32", b=b, values=v32)
    torch.cuda.synchronize()
    t4 = time.perf_counter()
    for _ in range(args.steps):
        asm.assemble(xd, "NRP32", b=b, values=v32)
    torch.cuda.synchronize()
    t5 = time.perf_counter()
    asm.profile(True)
    for _ in range(args.steps):
        asm.assemble(xd, "NRP32", b=b, values=v32)
    jac32_ms, scatter32_ms, launches32 = asm.profile_read()
    asm.profile(False)
    rel32 = float(torch.linalg.vector_norm(v32 - v) / torch.linalg.vector_norm(v))
    del v32

    ms_step = (t1 - t0) * 1e3 / args.steps
    ne = mesh.n_elements
    # The Jacobian pass is one launch per element colour (two on this mesh);
    # the library's events bracket the whole pass, so a launch's average is
    # the pass's time over the launches, and its work the pass's over them.
    colours = traffic_from_profiles(ne, "launches_per_assembly") or 1
    bulk_ms_pass = bulk_ms / max(launches, 1)
    bulk_ms_launch = bulk_ms_pass / colours
    bulk_flops = ne * JACOBIAN_KERNEL_FLOPS / colours
    achieved = bulk_flops / (bulk_ms_launch * 1e-3) / 1e12
    total_flops = ne * STEP_FLOPS
    out = {
        "metric": "elements/s residual+Jacobian assembly",
        "value": ne / (ms_step * 1e-3),
        "unit": "elements/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": f"cylinder-in-channel Q2 mesh, {ne} elements, residual + FD Jacobian (Flag_NR=NRP), "
                        "Oldroyd-B Wi=0.5 beta=0.59 Re=0 DfN=1e-3, fixed CSR pattern",
            "elements": ne, "unknowns": asm.N, "nnz": asm.nnz,
            "step_ms_p10_p50_p90": [pct(0.1), pct(0.5), pct(0.9)],
            "residual_only_ms": (t3 - t2) * 1e3 / args.steps,
            "fp32_jacobian": {
                "ms_per_step": (t5 - t4) * 1e3 / args.steps,
                "elements_per_s": ne / ((t5 - t4) / args.steps),
                "jacobian_kernel_ms": jac32_ms / max(launches32, 1),
                "scatter_ms": scatter32_ms / max(launches32, 1),
                "relative_difference_vs_fp64_fd": rel32,
                "roofline": {
                    "kernel": "vef_jac32_kernel", "bound": "fp32 compute",
                    "achieved": ne * AD32_FLOPS_PER_ELEMENT / (jac32_ms / max(launches32, 1) * 1e-3) / 1e12,
                    "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                    "frac": ne * AD32_FLOPS_PER_ELEMENT / (jac32_ms / max(launches32, 1) * 1e-3) / 1e12
                            / PEAK_FP32_TFLOPS,
                    "flops_per_element": AD32_FLOPS_PER_ELEMENT,
                },
                "note": "VEF_RES_JAC_FP32: the same FP64 residual, the Jacobian from forward-mode "
                        "differentiation of the Gauss-point terms in FP32 contracted with the basis functions",
            },
            "element_kernel_ms": bulk_ms_pass,
            "scatter_ms": scatter_ms / max(launches, 1),
            "total_tflops": total_flops / (ms_step * 1e-3) / 1e12,
        },
        "roofline": {
            "kernel": "vef_bulk_kernel<1,7,56> (the 55 perturbed element residuals and their quotients; "
                      "geometry from the residual pass)",
            "bound": "fp64 valu",
            "ceiling": "FP64 compute (MI355X FP64 vector peak = FP64 matrix peak = 78.6 TFLOP/s; "
                       "the kernel issues FP64 VALU ops without FMA contraction, so 0.5 of peak is its ceiling)",
            "achieved": achieved,
            "peak": PEAK_FP64_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / PEAK_FP64_TFLOPS,
            "traffic": traffic_from_profiles(ne),
            "flops_per_launch": bulk_flops,
            "launches_per_assembly": colours,
            "avg_launch_ms": bulk_ms_launch,
        },
    }
    step_traffic = traffic_from_profiles(ne, "step_hbm_bytes")
    if step_traffic:
        out["config"]["step_hbm_bytes"] = step_traffic
        out["config"]["step_hbm_tb_per_s"] = step_traffic / (ms_step * 1e-3) / 1e12
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(mesh, x, args.cpu_seconds)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
