"""Newton counterpart on the GPU: Dirichlet rows against their numpy
restatement (bit for bit), the re-factorisation LU against scipy, and a full
fixed-Wi Newton solve against the same iteration on the CPU with the oracle's
assembly and SuperLU (same flags and iteration count; states within 1e-8)."""
import numpy as np
import pytest

from ve_flow_around_cylinder_fortran_amd import Assembler, Params, cylinder_channel, smooth_state
from ve_flow_around_cylinder_fortran_amd.mesh import boundary_edges
from ve_flow_around_cylinder_fortran_amd.newton import NewtonSolver, cylinder_dirichlet, initial_state

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

OLDROYD_B = dict(ReN=0.0, BvN=0.59, VeN=0.0, DfN=1e-3, WiN0=1.0, model="L-PTT")


def test_dirichlet_rows_bitwise(oracle):
    from newton_cpu import apply_dirichlet, dirichlet_rows

    mesh = cylinder_channel(8, 4, 6, 8)
    x = smooth_state(mesh, EU=0.3)
    rows, vals = cylinder_dirichlet(mesh, 0.3)
    asm = Assembler(mesh, Params(**OLDROYD_B))
    asm.set_dirichlet(rows, vals)
    xd = torch.from_numpy(x).cuda()
    b, v = asm.assemble(xd, "NRP")
    asm.apply_dirichlet(xd, b, v, "NRP")
    torch.cuda.synchronize()
    rowptr, col = asm.pattern()
    ob, ov = oracle.assemble(oracle.params(**OLDROYD_B), mesh, x, rowptr=rowptr, col=col)
    r, vv = dirichlet_rows(rows, vals, asm.N)
    apply_dirichlet(r, vv, x, ob, ov, rowptr, col)
    assert np.array_equal(b.cpu().numpy(), ob) and np.array_equal(v.cpu().numpy(), ov)


def test_refactorization_lu_matches_scipy(oracle):
    import scipy.sparse as sp
    import scipy.sparse.linalg as sla

    from newton_cpu import apply_dirichlet, dirichlet_rows
    from ve_flow_around_cylinder_fortran_amd.linsolve import RfLU

    mesh = cylinder_channel(8, 4, 6, 8)
    asm = Assembler(mesh, Params(**OLDROYD_B))
    rows, vals = cylinder_dirichlet(mesh, 0.1)
    asm.set_dirichlet(rows, vals)
    rowptr, col = asm.pattern()
    lu = RfLU(rowptr, col)
    for k, x in enumerate([initial_state(mesh, 0.1), smooth_state(mesh, EU=0.1)]):
        print(f"[gpu newton] LU matrix {k}", flush=True)
        xd = torch.from_numpy(x).cuda()
        b, v = asm.assemble(xd, "NRP")
        asm.arclength(xd, b, v, increment=1, LN0=0.1)
        asm.apply_dirichlet(xd, b, v, "NRP")
        lu.factor(v)
        s = lu.solve(b)
        torch.cuda.synchronize()
        A = sp.csr_matrix((v.cpu().numpy(), col, rowptr), shape=(asm.N, asm.N))
        ref = sla.spsolve(A.tocsc(), b.cpu().numpy())
        err = np.abs(s.cpu().numpy() - ref).max() / np.abs(ref).max()
        assert err < 1e-9, err
    assert lu.analyses == 1 and lu.refactorizations == 2 and lu.host_fallbacks == 0


@pytest.mark.parametrize("kind", ["block", "hybrid", "band", "band-f32"])
def test_block_factorisations_match_scipy(kind):
    """BlockLU (6x6 node blocks), HybridLU (scalar re-factorisation in the
    node-block order, block solves) and BandLU (banded LU in reverse
    Cuthill-McKee node order, arclength border) on Newton matrices with the
    arclength and Dirichlet rows: solutions within 1e-9 of scipy's (1e-4 for
    BandLU's FP32 factorisation), for two matrices in a row (the second
    factorisation reuses the analysis)."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as sla

    from ve_flow_around_cylinder_fortran_amd.linsolve import BandLU, BlockLU, HybridLU

    mesh = cylinder_channel(8, 4, 6, 8)
    asm = Assembler(mesh, Params(**OLDROYD_B))
    rows, vals = cylinder_dirichlet(mesh, 0.1)
    asm.set_dirichlet(rows, vals)
    rowptr, col = asm.pattern()
    if kind == "band-f32":
        lu, tol = BandLU(rowptr, col, precision="f32"), 1e-4
    else:
        lu, tol = {"block": BlockLU, "hybrid": HybridLU, "band": BandLU}[kind](rowptr, col), 1e-9
    for k, x in enumerate([initial_state(mesh, 0.1), smooth_state(mesh, EU=0.1)]):
        print(f"[gpu newton] {kind} LU matrix {k}", flush=True)
        xd = torch.from_numpy(x).cuda()
        b, v = asm.assemble(xd, "NRP")
        asm.arclength(xd, b, v, increment=1, LN0=0.1)
        asm.apply_dirichlet(xd, b, v, "NRP")
        lu.factor(v)
        s = lu.solve(b)
        torch.cuda.synchronize()
        A = sp.csr_matrix((v.cpu().numpy(), col, rowptr), shape=(asm.N, asm.N))
        ref = sla.spsolve(A.tocsc(), b.cpu().numpy())
        err = np.abs(s.cpu().numpy() - ref).max() / np.abs(ref).max()
        print(f"[gpu newton] {kind} LU matrix {k}: max error {err:.2e} relative", flush=True)
        assert err < tol, (kind, k, err)
    assert lu.analyses == 1 and lu.refactorizations == 2 and lu.host_fallbacks == 0


@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_band_factor_graph_replays_match_direct_launches(precision):
    """BandLU's first factorisation launches its kernels directly; the second
    captures them into a HIP graph and every later one replays it.  Four
    factorisations in a row, with solves, another matrix and host
    synchronisations between them: every replay gives the direct
    factorisation's solution bit for bit."""
    from ve_flow_around_cylinder_fortran_amd.linsolve import BandLU

    mesh = cylinder_channel(8, 4, 6, 8)
    asm = Assembler(mesh, Params(**OLDROYD_B))
    rows, vals = cylinder_dirichlet(mesh, 0.1)
    asm.set_dirichlet(rows, vals)
    rowptr, col = asm.pattern()
    mats = []
    for x in (initial_state(mesh, 0.1), smooth_state(mesh, EU=0.1)):
        xd = torch.from_numpy(x).cuda()
        b, v = asm.assemble(xd, "NRP")
        asm.arclength(xd, b, v, increment=1, LN0=0.1)
        asm.apply_dirichlet(xd, b, v, "NRP")
        mats.append((b.clone(), v.clone()))
    lu = BandLU(rowptr, col, precision=precision)
    print(f"[gpu newton] band {precision} graph replays", flush=True)
    import ctypes

    from ve_flow_around_cylinder_fortran_amd.linsolve import _band_lib

    ab, ld, K = ctypes.c_void_p(), ctypes.c_int64(), ctypes.c_int32()
    rc = _band_lib().vef_band_storage(lu.h, ctypes.byref(ab), ctypes.byref(ld), ctypes.byref(K))
    assert (rc == 0) == (precision == "f64")  # an FP32 band is not double: ab must be NULL
    assert _band_lib().vef_band_storage(lu.h, None, ctypes.byref(ld), ctypes.byref(K)) == 0
    assert (ld.value, K.value) == (lu.ld, lu.K)
    ref = {}
    for k in range(5):
        m = 1 if k == 2 else 0  # the other matrix in between
        b, v = mats[m]
        lu.factor(v)
        s = lu.solve(b)
        torch.cuda.synchronize()
        if m not in ref:
            ref[m] = s.clone()
            assert k < 2 or lu.uses_graph
        else:
            assert torch.equal(s, ref[m]), (precision, k)
        assert lu.uses_graph == (k >= 1)
    assert lu.host_fallbacks == 0


def test_band_f32_gmres_refinement_matches_scipy():
    """FP32 band factors as the right preconditioner of restarted FP64 GMRES
    (BandLU's refinement, forced here with a 1e-12 target): every solve is
    refined and lands within 1e-10 of scipy's FP64 solution."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as sla

    from ve_flow_around_cylinder_fortran_amd.linsolve import BandLU

    mesh = cylinder_channel(8, 4, 6, 8)
    asm = Assembler(mesh, Params(**OLDROYD_B))
    rows, vals = cylinder_dirichlet(mesh, 0.1)
    asm.set_dirichlet(rows, vals)
    rowptr, col = asm.pattern()
    lu = BandLU(rowptr, col, precision="f32", refine_tol=1e-12)
    xd = torch.from_numpy(smooth_state(mesh, EU=0.1)).cuda()
    b, v = asm.assemble(xd, "NRP")
    asm.arclength(xd, b, v, increment=1, LN0=0.1)
    asm.apply_dirichlet(xd, b, v, "NRP")
    print("[gpu newton] band f32 + GMRES refinement", flush=True)
    lu.factor(v)
    A = sp.csr_matrix((v.cpu().numpy(), col, rowptr), shape=(asm.N, asm.N))
    for k in range(2):
        rhs = b if k == 0 else torch.from_numpy(np.random.default_rng(3).standard_normal(asm.N)).cuda()
        s = lu.solve(rhs).cpu().numpy()
        ref = sla.spsolve(A.tocsc(), rhs.cpu().numpy())
        err = np.abs(s - ref).max() / np.abs(ref).max()
        print(f"[gpu newton] refined solve {k}: error {err:.2e}", flush=True)
        assert err < 1e-10
    assert lu.refinements == 2 and lu.gmres_iterations >= 2 and lu.host_fallbacks == 0


def test_newton_matches_cpu_iteration(oracle):
    from newton_cpu import newton_cpu

    mesh = cylinder_channel(8, 4, 6, 8)
    LN0 = 0.1
    x0 = initial_state(mesh, LN0)
    print("[gpu newton] full solve, small mesh", flush=True)
    solver = NewtonSolver(mesh, Params(**OLDROYD_B))
    xg, hist = solver.solve(x0, LN0)
    xc, hc = newton_cpu(oracle, mesh, OLDROYD_B, x0, LN0)
    assert [h.flag for h in hist] == [h[1] for h in hc]
    xg = xg.cpu().numpy()
    assert np.abs(xg - xc).max() <= 1e-8 * np.abs(xc).max()
    assert hist[-1].cor_norm < 1e-8 and hist[-1].res_norm < 1e-8
    # quadratic convergence at the end: each NRP correction below the square of the previous (scaled)
    cors = [h.cor_norm for h in hist]
    assert cors[-1] < 1e-8 and cors[-2] < 1e-4
    # the drag of the converged state, GPU against the oracle on the CPU state
    edges = boundary_edges(mesh, "Cylinder")
    solver.asm.set_drag_edges(edges)
    Fd = solver.asm.drag(torch.from_numpy(xg).cuda()).item()
    oFd, _ = oracle.drag(oracle.params(**OLDROYD_B), mesh, xc, edges)
    assert abs(Fd - oFd) <= 1e-8 * abs(oFd)


@pytest.mark.parametrize("model,VeN", [("L-PTT", 0.25), ("GIESEKUS", 0.1)])
def test_newton_other_models_match_cpu_iteration(oracle, model, VeN):
    """Linear PTT and Giesekus, whose inlet values come from the fully
    developed inflow problem (inflow.FullyDevelopedInflow, solved at the
    state's Wi): the GPU Newton solve takes the CPU restatement's iteration
    flags (the CPU side with its own inflow solver), the states agree within
    the Newton tolerance, and so does the drag."""
    from newton_cpu import newton_cpu

    from ve_flow_around_cylinder_fortran_amd.inflow import FullyDevelopedInflow

    pset = dict(OLDROYD_B, VeN=VeN, model=model)
    mesh = cylinder_channel(8, 4, 6, 8)
    LN0 = 0.1  # from rest, as the continuation starts
    x0 = initial_state(mesh, LN0)
    print(f"[gpu newton] {model} full solve, small mesh", flush=True)
    solver = NewtonSolver(mesh, Params(**pset))
    assert solver.inflow is not None
    xg, hist = solver.solve(x0, LN0)
    xc, hc = newton_cpu(oracle, mesh, pset, x0, LN0, inflow=FullyDevelopedInflow.from_mesh(mesh, Params(**pset)))
    assert [h.flag for h in hist] == [h[1] for h in hc]
    xg = xg.cpu().numpy()
    assert np.abs(xg - xc).max() <= 1e-8 * np.abs(xc).max()
    assert hist[-1].cor_norm < 1e-8 and hist[-1].res_norm < 1e-8
    # the inlet rows hold the inflow solution
    (U, _, Sxx, Sxy, Syy), _ = solver.inflow.solve(LN0)
    nodes = solver.inflow.nodes
    assert np.abs(xg[6 * nodes] - U).max() < 1e-12 and np.abs(xg[6 * nodes + 4] - Sxy).max() < 1e-12
    edges = boundary_edges(mesh, "Cylinder")
    solver.asm.set_drag_edges(edges)
    Fd = solver.asm.drag(torch.from_numpy(xg).cuda()).item()
    oFd, _ = oracle.drag(oracle.params(**pset), mesh, xc, edges)
    assert abs(Fd - oFd) <= 1e-8 * abs(oFd)


@pytest.mark.parametrize("factorization", ["band", "hybrid"])
def test_single_precision_factorisation_newton(factorization):
    """Inexact Newton: the FP64 Jacobian factorised in FP32 (BandLU: FP32 band
    with FP64 solve vectors; HybridLU: FP32 re-factorisation and solves), the
    residual, update and convergence test in FP64.  It converges with the same
    iteration flags, and the drag differs from the FP64 Newton's by at most
    1e-8 relative."""
    mesh = cylinder_channel(8, 4, 6, 8)
    LN0 = 0.1
    x0 = initial_state(mesh, LN0)
    print(f"[gpu newton] fp32 factorisation ({factorization}), small mesh", flush=True)
    edges = boundary_edges(mesh, "Cylinder")
    out = {}
    for prec in ("f64", "f32"):
        solver = NewtonSolver(mesh, Params(**OLDROYD_B), linear_precision=prec, factorization=factorization)
        x, hist = solver.solve(x0, LN0)
        assert hist[-1].cor_norm < 1e-8 and hist[-1].res_norm < 1e-8
        assert solver.lu.host_fallbacks == 0
        solver.asm.set_drag_edges(edges)
        out[prec] = ([h.flag for h in hist], solver.asm.drag(x).item())
    assert out["f32"][0] == out["f64"][0]
    assert abs(out["f32"][1] - out["f64"][1]) <= 1e-8 * abs(out["f64"][1])


def test_arclength_continuation_matches_cpu(oracle):
    """Six continuation steps (two pinned, four pseudo-arclength) on the GPU
    against the same continuation on the CPU: the same step sizes (to 1e-4),
    and EU and states equal to within the Newton tolerance (corrections below
    1e-8)."""
    from newton_cpu import continuation_cpu
    from ve_flow_around_cylinder_fortran_amd.newton import ArclengthContinuation

    mesh = cylinder_channel(8, 4, 6, 8)
    x0 = initial_state(mesh, 0.1)
    print("[gpu newton] continuation, small mesh", flush=True)
    cont = ArclengthContinuation(NewtonSolver(mesh, Params(**OLDROYD_B)), LN0=0.1, dArN=0.02)
    cont.start(x0)
    gpu_states, gpu_dS = [], []
    for _ in range(6):
        gpu_dS.append(cont.dS)
        x, _ = cont.step()
        gpu_states.append(x.cpu().numpy())
    cpu_states, cpu_dS = continuation_cpu(oracle, mesh, OLDROYD_B, x0, 0.1, 0.02, 6)
    # dS follows sqrt(eps_pr/err) of the relative prediction error, which sees
    # the converged states' differences (within the Newton tolerance) amplified
    assert np.allclose(gpu_dS, cpu_dS, rtol=1e-4, atol=0)
    for g, c in zip(gpu_states, cpu_states):
        assert abs(g[-1] - c[-1]) <= 2e-8
        assert np.abs(g - c).max() <= 1e-7 * np.abs(c).max()
    assert gpu_states[-1][-1] > gpu_states[2][-1] > 0.12


def test_emergency_restarts_match_cpu(oracle):
    """emergencyNewtonRaphson forced from the host: with the iteration cap at
    2 every step of a four-step continuation (two pinned, two
    pseudo-arclength) restarts from its prediction with xF = 1/2, the cap at
    100 and dS halved, then converges.  The GPU driver takes the CPU
    restatement's restarts, step sizes and states."""
    from newton_cpu import continuation_cpu
    from ve_flow_around_cylinder_fortran_amd.newton import ArclengthContinuation

    mesh = cylinder_channel(8, 4, 6, 8)
    x0 = initial_state(mesh, 0.1)
    print("[gpu newton] continuation with forced emergency restarts, small mesh", flush=True)
    cont = ArclengthContinuation(NewtonSolver(mesh, Params(**OLDROYD_B)), LN0=0.1, dArN=0.02)
    cont.max_iter = 2
    cont.start(x0)
    gpu_states, gpu_dS, gpu_em = [], [], []
    for _ in range(4):
        x, hist = cont.step()
        gpu_states.append(x.cpu().numpy())
        gpu_dS.append(cont.dS)
        gpu_em.append([(e[1], e[2]) for e in cont.s.emergencies])
        assert hist[-1].cor_norm < 1e-8 and hist[-1].res_norm < 1e-8
    em = []
    cpu_states, cpu_dS = continuation_cpu(oracle, mesh, OLDROYD_B, x0, 0.1, 0.02, 4, max_iter=2, emergencies=em)
    assert gpu_em == [[(e[1], e[2]) for e in step] for step in em]
    assert all(len(step) == 1 and step[0] == ("iterations", 0.5) for step in gpu_em)
    for g, c in zip(gpu_states, cpu_states):
        assert abs(g[-1] - c[-1]) <= 2e-8
        assert np.abs(g - c).max() <= 1e-7 * np.abs(c).max()
    # cpu_dS holds the step entering each increment; the GPU's the one after it
    assert np.allclose(gpu_dS[:-1], cpu_dS[1:], rtol=1e-4, atol=0)


def test_divergence_stops_after_seven_restarts():
    """A divergence norm below any correction: the solve restarts with xF
    halved each time and stops with NewtonError once xF < 1e-2 (the seventh
    restart), as the reference's general stop does."""
    from ve_flow_around_cylinder_fortran_amd.newton import NewtonError

    mesh = cylinder_channel(8, 4, 6, 8)
    solver = NewtonSolver(mesh, Params(**OLDROYD_B))
    print("[gpu newton] forced divergence, small mesh", flush=True)
    with pytest.raises(NewtonError, match="below 1e-2"):
        solver.solve(initial_state(mesh, 0.1), 0.1, cor_norm_div=1e-30)
    assert [e[2] for e in solver.emergencies] == [0.5 ** k for k in range(1, 8)]
    assert all(e[1] == "divergence" for e in solver.emergencies)


def test_inexact_newton_converges_to_reference_solutions():
    """The inexact Newton (the last factorisation carried across continuation
    steps) converges to the same solutions: after five steps, an FP64 Newton
    solve with the reference policy at the final Wi, started from the inexact
    state, takes one iteration and moves the drag by less than 1e-10
    relative; fewer factorisations than steps were needed."""
    from ve_flow_around_cylinder_fortran_amd.newton import ArclengthContinuation

    mesh = cylinder_channel(8, 4, 6, 8)
    print("[gpu newton] inexact Newton continuation, small mesh", flush=True)
    solver = NewtonSolver(mesh, Params(**OLDROYD_B), reuse_factors=True)
    cont = ArclengthContinuation(solver, LN0=0.1, dArN=0.02)
    cont.start(initial_state(mesh, 0.1))
    flags = []
    for _ in range(5):
        x, hist = cont.step()
        flags.append("".join(h.flag[0] for h in hist))
        assert hist[-1].cor_norm < 1e-8 and hist[-1].res_norm < 1e-8
    assert any(f.startswith("M") for f in flags[1:])  # steps that start on the previous factorisation
    ref = NewtonSolver(mesh, Params(**OLDROYD_B))
    xr, hr = ref.solve(x, float(x[-1]))
    assert len(hr) <= 2
    edges = boundary_edges(mesh, "Cylinder")
    solver.asm.set_drag_edges(edges)
    ref.asm.set_drag_edges(edges)
    Fd, Fr = solver.asm.drag(x).item(), ref.asm.drag(xr).item()
    assert abs(Fd - Fr) <= 1e-10 * abs(Fr)


def _continue_to(mesh, target_EU, predictor, max_steps=25):
    from ve_flow_around_cylinder_fortran_amd.newton import ArclengthContinuation

    cont = ArclengthContinuation(NewtonSolver(mesh, Params(**OLDROYD_B)), LN0=0.1, dArN=0.02, predictor=predictor)
    cont.start(initial_state(mesh, 0.1))
    iters = 0
    for k in range(max_steps):
        x, hist = cont.step()
        if k >= 2:
            iters += len(hist)
        if float(x[-1]) >= target_EU:
            return k + 1, iters, cont
    raise AssertionError(f"EU {float(x[-1])} below {target_EU} after {max_steps} steps")


def test_learned_predictor_gains_over_tangent():
    """From Wi = 0.1 to Wi = 0.145 on the small mesh: with the online-trained
    per-node correction of the tangent predictor the continuation needs fewer
    steps and fewer Newton iterations than with the tangent predictor alone
    (the learned proposal is taken only where its residual is smaller)."""
    from ve_flow_around_cylinder_fortran_amd.learned import LearnedPredictor

    mesh = cylinder_channel(8, 4, 6, 8)
    print("[gpu newton] continuation to Wi 0.145, tangent predictor", flush=True)
    steps_t, iters_t, _ = _continue_to(mesh, 0.145, None)
    print("[gpu newton] continuation to Wi 0.145, learned predictor", flush=True)
    steps_l, iters_l, cont = _continue_to(mesh, 0.145, LearnedPredictor(mesh, seed=0))
    print(f"[gpu newton] tangent: {steps_t} steps, {iters_t} Newton iterations; learned: {steps_l} steps, "
          f"{iters_l} iterations; choices {cont.choices}", flush=True)
    assert "learned" in cont.choices
    assert steps_l < steps_t and iters_l < iters_t


def test_learned_predictor_overlapped_training():
    """Training on a second stream during the next Newton solve
    (LearnedPredictor(overlap=True)): every step converges, the network is
    trained once per observation after the first proposal, and its
    proposals are taken."""
    from ve_flow_around_cylinder_fortran_amd.learned import LearnedPredictor

    mesh = cylinder_channel(8, 4, 6, 8)
    print("[gpu newton] continuation to Wi 0.145, learned predictor trained on a second stream", flush=True)
    pred = LearnedPredictor(mesh, seed=0, overlap=True)
    steps_l, iters_l, cont = _continue_to(mesh, 0.145, pred)
    print(f"[gpu newton] overlapped: {steps_l} steps, {iters_l} iterations; choices {cont.choices}", flush=True)
    assert "learned" in cont.choices
    assert pred._rounds >= len(pred.losses) >= pred._rounds - 1 >= 1
    assert all(np.isfinite(pred.losses))


def test_band_lu_workload_mesh_and_partial_last_block():
    """BandLU on the 1,600-element Newton matrix (39,792 node unknowns: the
    last 64-column block is partial) against scipy, and a Newton solve with it
    against the default solver's (same flags, states within the Newton
    tolerance, drag within 1e-10)."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as sla

    from ve_flow_around_cylinder_fortran_amd.linsolve import BandLU
    from ve_flow_around_cylinder_fortran_amd.mesh import boundary_edges, workload_mesh

    mesh = workload_mesh(1600)
    x0 = initial_state(mesh, 0.1)
    print("[gpu newton] band LU, 1600 elements", flush=True)
    sb = NewtonSolver(mesh, Params(**OLDROYD_B), factorization="band")
    assert isinstance(sb.lu, BandLU) and (sb.N - 1) % 64 != 0
    xb, hb = sb.solve(x0, 0.1)
    sh = NewtonSolver(mesh, Params(**OLDROYD_B), factorization="hybrid")
    xh, hh = sh.solve(x0, 0.1)
    assert [h.flag for h in hb] == [h.flag for h in hh] and sb.lu.host_fallbacks == 0
    assert float((xb - xh).abs().max()) < 1e-8
    edges = boundary_edges(mesh, "Cylinder")
    sb.asm.set_drag_edges(edges)
    sh.asm.set_drag_edges(edges)
    Fb, Fh = sb.asm.drag(xb).item(), sh.asm.drag(xh).item()
    assert abs(Fb - Fh) <= 1e-10 * abs(Fh)
    v = sb.values
    rowptr, col = sb.asm.pattern()
    A = sp.csr_matrix((v.cpu().numpy(), col, rowptr), shape=(sb.N, sb.N))
    rhs = np.random.default_rng(5).standard_normal(sb.N)
    ref = sla.spsolve(A.tocsc(), rhs)
    sb.lu.factor(v)
    s = sb.lu.solve(torch.from_numpy(rhs).cuda()).cpu().numpy()
    assert np.abs(s - ref).max() / np.abs(ref).max() < 1e-9
