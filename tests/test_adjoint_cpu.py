"""CPU checks for the adjoint's test infrastructure and host logic.

- The torch float64 restatement (tests/torch_swe.py), whose autograd gives
  the reference gradients of the GPU adjoint tests, is pinned against the
  C++ oracle: tendencies and steps agree to rounding.
- The binomial checkpointing cost reported by the library equals an
  independent dynamic program over all split points and reproduces the
  textbook count (200 steps, 10 snapshots: 3*200 - C(13,2) = 522).
No kernel runs here.
"""
import math

import numpy as np
import pytest
import torch

from oracle import pyoracle as po
from omega_playground_amd import swe

import torch_swe as ts
from helpers import ALL_TERMS, params_dict, rel_err, rich_state

TOL = 1e-13


def _pair(nx, ny, L, ntr, dc=10e3, seed=0):
    pd = params_dict(dc, ntr)
    om = po.Mesh(nx, ny, dc, L)
    omodel = po.Model(om, **pd)
    state = rich_state(om, L, ntr, seed)
    return om, omodel, pd, state


@pytest.mark.parametrize("terms", [1, 2, 4, 8, 16, 32, 64, ALL_TERMS])
def test_torch_tendency_matches_oracle(terms):
    om, omodel, pd, state = _pair(8, 6, 5, 2)
    omodel.terms = terms
    tend = omodel.new_state()
    omodel.compute_tendency(tend, state, 0.0)
    tm = ts.TorchMesh(om, 5)
    out = ts.tendency(tm, pd, terms, *(torch.from_numpy(a) for a in state))
    for name, a, b in zip(("h", "vn", "tr"), out, tend):
        assert rel_err(a.numpy(), b) <= TOL, name


@pytest.mark.parametrize("kind,nstages", [("rk4", 4), ("lsrk", 5), ("lsrk", 1)])
def test_torch_step_matches_oracle(kind, nstages):
    om, omodel, pd, state = _pair(8, 8, 3, 2, seed=4)
    st = po.Stepper(omodel, kind, nstages)
    dt = 20.0
    y = tuple(torch.from_numpy(a.copy()) for a in state)
    if kind == "rk4":
        out = ts.rk4_step(ts.TorchMesh(om, 3), pd, ALL_TERMS, dt, *y)
    else:
        a, b, _ = st.coeffs()
        out = ts.lsrk_step(ts.TorchMesh(om, 3), pd, ALL_TERMS, dt, a, b, *y)
    st.do_step(0.0, dt, state)
    for name, x, r in zip(("h", "vn", "tr"), out, state):
        assert rel_err(x.numpy(), r) <= TOL, name


def test_torch_rk4_coefficients_match_oracle():
    om, omodel, _, _ = _pair(4, 4, 1, 0)
    a, b, _ = po.Stepper(omodel, "rk4").coeffs()
    assert tuple(a[:3]) == ts.RK4_A
    assert np.allclose(b, ts.RK4_B, rtol=0, atol=1e-16)


def _dp_costs(T, C):
    """F[l][c]: forward steps to reverse l steps from a stored state with c
    snapshots (the stored one included), minimised over every split."""
    F = [[0] * (C + 1) for _ in range(T + 1)]
    for l in range(2, T + 1):
        F[l][1] = l * (l - 1) // 2
    for c in range(2, C + 1):
        for l in range(2, T + 1):
            F[l][c] = min(m + F[l - m][c - 1] + F[m][c] for m in range(1, l))
    return F


def test_revolve_cost_matches_dynamic_program():
    T, C = 60, 7
    F = _dp_costs(T, C)
    for l in range(1, T + 1):
        for c in range(1, C + 1):
            assert swe.revolve_cost(l, c) == F[l][c], (l, c)


def test_revolve_cost_textbook_values():
    # r = 3 repetitions: C(13, 10) = 286 >= 200 > C(12, 10) = 66
    assert swe.revolve_cost(200, 10) == 3 * 200 - math.comb(13, 2) == 522
    # one snapshot: restart from y_0 for every step
    assert swe.revolve_cost(50, 1) == 50 * 49 // 2
    # a snapshot per step: one forward sweep
    assert swe.revolve_cost(50, 50) == 49


def test_adjoint_entry_points_reject_null_handles():
    L = swe.lib()
    assert L.omega_tape_create(None, 10, 3, 0, None) != 0
    assert L.omega_model_tendency_vjp(None, None, None, None, None, None) != 0
    assert L.omega_revolve_cost(-1, 3) == -1
