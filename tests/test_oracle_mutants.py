"""Mutation check of the oracle's pinning: each build of the oracle with a
planted plausible mistake (swe_oracle.cpp, ORC_MUTANT = 1..8: a dropped
term, a flipped sign, a swapped operand or coefficient) must fail at least
one of the pinning / reference tests.  The unmodified oracle passes them all
(test_oracle_pinning.py, test_oracle_reference.py)."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE = os.path.join(ROOT, "oracle")

# which pinning test is expected to catch each planted mistake
EXPECTED = {
    1: "test_uniform_flow_closed_forms",            # PV drops a term
    2: "test_linear_thickness_closed_forms",        # SSH gradient sign
    3: "test_momentum_del2_summation_by_parts",     # dc for dv in del2
    4: "test_tracer_flux_consistency",              # tracer edge value x2
    5: "test_rk4_amplification_polynomial",         # RK4 weights swapped
    6: "test_lsrk5_fourth_order",                   # LSRK5 coefficient sign
    7: "test_uniform_flow_closed_forms",            # KE without 1/4
    8: "test_momentum_del2_summation_by_parts",     # vertex circulation sign
}


@pytest.fixture(scope="module")
def mutants():
    subprocess.run(["make", "-s", "-C", ORACLE, "-j4", "mutants"], check=True)
    return {n: os.path.join(ORACLE, "_mutants", f"liboracle_mut{n}.so") for n in EXPECTED}


@pytest.mark.parametrize("n", sorted(EXPECTED))
def test_planted_mistake_is_caught(mutants, n):
    env = dict(os.environ, ORACLE_LIB=mutants[n])
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-p", "no:cacheprovider",
                        os.path.join(ROOT, "tests", "test_oracle_pinning.py"),
                        "-k", EXPECTED[n]], cwd=ROOT, env=env, capture_output=True, text=True)
    assert r.returncode == 1, f"mutant {n} not caught by {EXPECTED[n]}:\n{r.stdout[-2000:]}"
    assert "failed" in r.stdout
