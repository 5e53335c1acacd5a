"""Mutation check of the oracle's pinning: each build of the oracle with a
planted plausible mistake (swe_oracle.cpp, ORC_MUTANT = 1..12: a dropped
term, a flipped sign, a swapped operand or coefficient) must fail at least
one of the pinning tests.  The unmodified oracle passes them all
(test_oracle_pinning.py, test_oracle_reference.py, test_linear_model.py)."""
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
    9: ("test_linear_model.py", "test_linear_energy_is_conserved"),     # linear SSH sign
    10: ("test_linear_model.py", "test_linear_equals_nonlinear_at_rest"),  # h0 dropped
    11: ("test_linear_model.py", "test_linear_energy_uses_h0"),         # energy with h
    12: "test_bottom_drag_uniform_flow_decay",      # drag on the top level
}


@pytest.fixture(scope="module")
def mutants():
    subprocess.run(["make", "-s", "-C", ORACLE, "-j4", "mutants"], check=True)
    return {n: os.path.join(ORACLE, "_mutants", f"liboracle_mut{n}.so") for n in EXPECTED}


@pytest.mark.parametrize("n", sorted(EXPECTED))
def test_planted_mistake_is_caught(mutants, n):
    env = dict(os.environ, ORACLE_LIB=mutants[n])
    fname, k = EXPECTED[n] if isinstance(EXPECTED[n], tuple) else ("test_oracle_pinning.py",
                                                                  EXPECTED[n])
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-p", "no:cacheprovider",
                        os.path.join(ROOT, "tests", fname), "-k", k],
                       cwd=ROOT, env=env, capture_output=True, text=True)
    assert r.returncode == 1, f"mutant {n} not caught by {EXPECTED[n]}:\n{r.stdout[-2000:]}"
    assert "failed" in r.stdout
