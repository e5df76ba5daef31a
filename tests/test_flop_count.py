"""The algorithmic operation count bench.py uses for the roofline: counted by
the operation-counting build of the oracle on one element residual of the
workload configuration (Oldroyd-B, Re = 0, DfN = 1e-3), independent of the
state values (no data-dependent branches in this configuration)."""
import importlib.util
import os

import numpy as np

from test_oracle_element import curved_xy, random_state

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def bench_constant():
    spec = importlib.util.spec_from_file_location("bench", os.path.join(ROOT, "bench.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod.FLOPS_PER_RESIDUAL_EVAL


def test_bench_flop_constant_matches_oracle_count(oracle):
    rng = np.random.default_rng(0)
    prm = oracle.params(ReN=0.0, BvN=0.59, VeN=0.0, DfN=1e-3, WiN0=1.0, model="L-PTT")
    counts = set()
    for _ in range(3):
        xy = curved_xy(rng)
        counts.add(oracle.count_flops(prm, xy, random_state(rng, xy), 0.5)["total"])
    assert counts == {bench_constant()}


def test_counter_sees_model_terms(oracle):
    """The counter is not blind: Giesekus adds the S^3 and C:C work, inertia the
    convective terms."""
    rng = np.random.default_rng(1)
    xy = curved_xy(rng)
    TL = random_state(rng, xy)
    base = oracle.count_flops(oracle.params(), xy, TL, 0.5)["total"]
    gies = oracle.count_flops(oracle.params(VeN=0.1, model="GIESEKUS"), xy, TL, 0.5)["total"]
    inert = oracle.count_flops(oracle.params(ReN=1.0), xy, TL, 0.5)["total"]
    dcs = oracle.count_flops(oracle.params(DfN=0.0), xy, TL, 0.5)["total"]
    assert gies > base and inert > base and dcs > base


def test_stability_flop_constants_match_oracle_count(oracle):
    """bench.py's per-evaluation and base-state counts of the stability
    assembly, for the --stability workload (L-PTT epsilon 0.25, Re = 0,
    DfN = 1e-3), at every perturbed unknown of an element."""
    import importlib.util

    spec = importlib.util.spec_from_file_location("bench", os.path.join(ROOT, "bench.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    rng = np.random.default_rng(3)
    prm = oracle.params(ReN=0.0, BvN=0.59, VeN=0.25, DfN=1e-3, WiN0=1.0, model="L-PTT")
    seen = set()
    for p in (0, 7, 20, 33, 53):
        xy = curved_xy(rng)
        TLd = np.zeros((9, 6))
        TLd.flat[p] = 1e-8
        full, base = oracle.count_flops_stability(prm, xy, random_state(rng, xy), 0.6, TLd)
        seen.add((full["total"], base["total"]))
    assert seen == {(mod.STAB_FLOPS_PER_EVAL, mod.STAB_FLOPS_BASE)}


def test_geometry_flop_constant_matches_oracle_count(oracle):
    """The geometry-only part of an evaluation (basis2D, Helem**2, BvN/H2,
    DfN/H2 at the 9 Gauss points) that the Jacobian pass takes from the
    residual pass: bench.py's constant, the same for any element shape, and
    the derived kernel and step counts."""
    spec = importlib.util.spec_from_file_location("bench", os.path.join(ROOT, "bench.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    rng = np.random.default_rng(4)
    prm = oracle.params(ReN=0.0, BvN=0.59, VeN=0.0, DfN=1e-3, WiN0=1.0, model="L-PTT")
    counts = {oracle.count_flops_geometry(prm, curved_xy(rng))["total"] for _ in range(3)}
    assert counts == {mod.FLOPS_GEOMETRY_PER_EVAL}
    g = oracle.count_flops_geometry(prm, curved_xy(rng))
    assert g["sqrt"] == 9 and g["div"] == 9 * 7  # 1/sqrt(Helem), 4 in basis2D, BvN/H2, DfN/H2 per point
    E, G = mod.FLOPS_PER_RESIDUAL_EVAL, mod.FLOPS_GEOMETRY_PER_EVAL
    assert mod.JACOBIAN_KERNEL_FLOPS == 55 * (E - G) + 54 * 55 * 3
    assert mod.STEP_FLOPS == G + 56 * (E - G) + 54 * 55 * 3 + 2916 + 108
