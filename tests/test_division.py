"""The GPU kernel divides many numerators by one denominator as
q1 = RN(q0 - RN(b*q0 - a)*y) with y = RN(1/b), q0 = RN(a*y).  These checks
pin that this equals the IEEE quotient: exhaustively over all significand
pairs at small precisions (exact integer arithmetic), and on binary64 pairs
drawn uniformly and near the hard cases, inside the exponent window the kernel
accepts (outside it the kernel repeats the evaluation with IEEE division)."""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("division") / "division_check")
    src = os.path.join(HERE, "division", "division_check.c")
    subprocess.run(["gcc", "-O2", "-ffp-contract=off", "-o", exe, src, "-lm"], check=True)
    return exe


@pytest.mark.parametrize("p", [4, 6, 8, 10, 12])
def test_exhaustive_small_precisions(checker, p):
    r = subprocess.run([checker, "exhaustive", str(p)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr


def test_binary64_samples(checker):
    r = subprocess.run([checker, "binary64", "4000000", "20261203"], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
