"""GPU tests of the discrete adjoint.

Reference gradients come from torch.autograd on the float64 CPU restatement
in tests/torch_swe.py (itself pinned against the oracle by
tests/test_adjoint_cpu.py); the adjoint is also checked without torch,
against central finite differences of the library's own forward path.

Tolerances (relative max-norm unless stated):
  - tendency / step adjoint vs autograd reference: 1e-10 (both are exact
    derivatives; only the summation order differs),
  - dot-product test with an exact tangent (torch.func.jvp): 1e-11,
  - central finite differences: 1e-7 at the tendency and step level and,
    with one Richardson extrapolation step, for the checkpointed
    multi-step gradient,
  - checkpointed multi-step gradient vs autograd through the same steps:
    1e-8 (the misfit residual y_T - obs cancels about five digits, so the
    two forward paths' roundings show up at ~1e-10),
  - checkpointed gradients for different snapshot counts: bit-identical.
"""
import numpy as np
import pytest
import torch

from oracle import pyoracle as po
from omega_playground_amd import swe

import torch_swe as ts
from helpers import ALL_TERMS, params_dict, rel_err, rich_state, same_bits

pytestmark = pytest.mark.gpu

TOL_REF = 1e-10
TOL_DOT = 1e-11
TOL_FD = 1e-7
TOL_FD_RUN = 1e-7
TOL_RUN_REF = 1e-8
DC = 10e3


def setup(nx, ny, L, ntr, ordering="hilbert", seed=0, terms=ALL_TERMS):
    pd = params_dict(DC, ntr)
    gm = swe.PlanarHexagonalMesh(nx, ny, DC, L, ordering=ordering)
    gmodel = swe.ShallowWaterModel(gm, swe.ShallowWaterParams(**pd))
    gmodel.terms = terms
    state = rich_state(gm, L, ntr, seed)
    return gm, gmodel, pd, state


def upload(gm, ntr, arrays):
    s = swe.ShallowWaterState(gm, ntr)
    s.upload(*arrays)
    return s


def random_like(state, seed, scale=(1.0, 1.0, 1.0)):
    rng = np.random.default_rng(seed)
    return tuple(s * rng.standard_normal(a.shape) for a, s in zip(state, scale))


def dot(a, b):
    return float(sum(np.sum(x * y) for x, y in zip(a, b)))


def torch_params(pd):
    P = {k: torch.tensor(float(pd.get(k, 0.0)), dtype=torch.float64, requires_grad=True)
         for k in swe.PARAM_NAMES}
    params = dict(pd)
    params.update(P)
    return params, [P[k] for k in swe.PARAM_NAMES]


def torch_grad(fn, state, cot):
    """Gradient of <cot, fn(y, params)> with respect to y and the four
    mixing coefficients."""
    y = [torch.tensor(a, requires_grad=True) for a in state]
    params, plist = fn[1]()
    out = fn[0](params, *y)
    loss = sum((o * torch.from_numpy(np.ascontiguousarray(c))).sum() for o, c in zip(out, cot))
    g = torch.autograd.grad(loss, y + plist, allow_unused=True)
    gs = [x.detach().numpy() if x is not None else np.zeros(a.shape) for x, a in zip(g[:3], state)]
    gp = np.array([float(x) if x is not None else 0.0 for x in g[3:]])
    return gs, gp


def check_pgrad(gp, ref, tol):
    scale = np.max(np.abs(ref))
    for i, name in enumerate(swe.PARAM_NAMES):
        assert abs(gp[i] - ref[i]) <= tol * max(abs(ref[i]), 1e-300) or \
            (ref[i] == 0.0 and abs(gp[i]) <= tol * scale), (name, gp[i], ref[i])


# ---------------------------------------------------------------------------
# Tendency adjoint
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("terms", [1, 2, 4, 8, 16, 32, 64, ALL_TERMS])
def test_tendency_vjp_matches_autograd(terms):
    gm, gmodel, pd, state = setup(10, 8, 6, 2, terms=terms)
    print("terms", terms, flush=True)
    cot = random_like(state, 11)
    gs, gc = upload(gm, 2, state), upload(gm, 2, cot)
    out = swe.ShallowWaterState(gm, 2)
    gp = swe.tendency_vjp(gmodel, gs, gc, out)
    tm = ts.TorchMesh(gm, 6)
    ref, refp = torch_grad((lambda p, *y: ts.tendency(tm, p, terms, *y),
                            lambda: torch_params(pd)), state, cot)
    for name, a, b in zip(("h", "vn", "tr"), out.download(), ref):
        assert rel_err(a, b) <= TOL_REF, name
    check_pgrad(gp, refp, TOL_REF)


@pytest.mark.parametrize("ordering", ["original", "morton"])
@pytest.mark.parametrize("disable", [(False, False), (True, False), (False, True)])
def test_tendency_vjp_orderings_and_disabled(ordering, disable):
    pd = params_dict(DC, 3)
    pd.update(disable_h_tendency=disable[0], disable_vn_tendency=disable[1])
    gm = swe.PlanarHexagonalMesh(9, 7, DC, 5, ordering=ordering)
    gmodel = swe.ShallowWaterModel(gm, swe.ShallowWaterParams(**pd))
    state = rich_state(gm, 5, 3, 2)
    print("case", ordering, disable, flush=True)
    cot = random_like(state, 5)
    out = upload(gm, 3, random_like(state, 6))  # accumulates onto existing values
    before = out.download()
    gp = swe.tendency_vjp(gmodel, upload(gm, 3, state), upload(gm, 3, cot), out)
    tm = ts.TorchMesh(gm, 5)
    ref, refp = torch_grad((lambda p, *y: ts.tendency(tm, p, ALL_TERMS, *y),
                            lambda: torch_params(pd)), state, cot)
    for name, a, b0, b in zip(("h", "vn", "tr"), out.download(), before, ref):
        assert rel_err(a - b0, b) <= TOL_REF, name
    check_pgrad(gp, refp, TOL_REF)


def test_tendency_vjp_specialised_forward_path():
    # nlayers = 64 on a regular hexagonal mesh: the diagnostics the adjoint
    # consumes come from the pattern-specialised forward kernels
    gm, gmodel, pd, state = setup(8, 8, 64, 1)
    assert gmodel.kernel_variant() == 2
    cot = random_like(state, 3)
    out = swe.ShallowWaterState(gm, 1)
    gp = swe.tendency_vjp(gmodel, upload(gm, 1, state), upload(gm, 1, cot), out)
    tm = ts.TorchMesh(gm, 64)
    ref, refp = torch_grad((lambda p, *y: ts.tendency(tm, p, ALL_TERMS, *y),
                            lambda: torch_params(pd)), state, cot)
    for name, a, b in zip(("h", "vn", "tr"), out.download(), ref):
        assert rel_err(a, b) <= TOL_REF, name
    check_pgrad(gp, refp, TOL_REF)


def test_tendency_dot_product_exact_tangent():
    """<J v, w> with J v from forward-mode differentiation of the torch
    restatement equals <v, J^T w> from the library."""
    gm, gmodel, pd, state = setup(10, 10, 4, 2)
    v = random_like(state, 21, (1.0, 0.1, 1.0))
    w = random_like(state, 22)
    tm = ts.TorchMesh(gm, 4)
    _, jv = torch.func.jvp(lambda *y: ts.tendency(tm, pd, ALL_TERMS, *y),
                           tuple(torch.from_numpy(a) for a in state),
                           tuple(torch.from_numpy(a) for a in v))
    lhs = dot([x.numpy() for x in jv], w)
    out = swe.ShallowWaterState(gm, 2)
    swe.tendency_vjp(gmodel, upload(gm, 2, state), upload(gm, 2, w), out)
    rhs = dot(v, out.download())
    assert abs(lhs - rhs) <= TOL_DOT * abs(lhs), (lhs, rhs)


def _tend(gm, gmodel, ntr, y):
    t = swe.ShallowWaterState(gm, ntr)
    gmodel.compute_tendency(t, upload(gm, ntr, y), 0.0)
    return t.download()


def test_tendency_dot_product_finite_differences():
    """<(F(y+eps v) - F(y-eps v)) / 2eps, w> against <v, J^T w>, all on the GPU."""
    gm, gmodel, pd, state = setup(12, 10, 4, 2)
    v = random_like(state, 31, (1e-3, 1e-6, 1e-3))
    w = random_like(state, 32)
    yp = tuple(a + b for a, b in zip(state, v))
    ym = tuple(a - b for a, b in zip(state, v))
    fp, fm = _tend(gm, gmodel, 2, yp), _tend(gm, gmodel, 2, ym)
    lhs = dot([(a - b) / 2 for a, b in zip(fp, fm)], w)
    out = swe.ShallowWaterState(gm, 2)
    swe.tendency_vjp(gmodel, upload(gm, 2, state), upload(gm, 2, w), out)
    rhs = dot(v, out.download())
    assert abs(lhs - rhs) <= TOL_FD * abs(rhs), (lhs, rhs)


# ---------------------------------------------------------------------------
# Step adjoints
# ---------------------------------------------------------------------------
STEPPERS = [("rk4", 4), ("lsrk", 5), ("lsrk", 1)]


def make_stepper(gmodel, kind, nstages):
    return swe.RK4Stepper(gmodel) if kind == "rk4" else swe.LSRKStepper(gmodel, nstages)


def torch_step(kind, nstages, tm, dt, ntr):
    if kind == "rk4":
        return lambda p, *y: ts.rk4_step(tm, p, ALL_TERMS, dt, *y)
    om = po.Mesh(4, 4, DC, 1)
    a, b, _ = po.Stepper(po.Model(om, ntracers=0), "lsrk", nstages).coeffs()
    return lambda p, *y: ts.lsrk_step(tm, p, ALL_TERMS, dt, a, b, *y)


@pytest.mark.parametrize("kind,nstages", STEPPERS)
def test_step_vjp_matches_autograd(kind, nstages):
    gm, gmodel, pd, state = setup(8, 8, 4, 2, seed=3)
    print("stepper", kind, nstages, flush=True)
    st = make_stepper(gmodel, kind, nstages)
    dt = 30.0
    cot = random_like(state, 41)
    gc = upload(gm, 2, cot)
    gp = swe.step_vjp(st, 0.0, dt, upload(gm, 2, state), gc)
    tm = ts.TorchMesh(gm, 4)
    ref, refp = torch_grad((torch_step(kind, nstages, tm, dt, 2), lambda: torch_params(pd)),
                           state, cot)
    for name, a, b in zip(("h", "vn", "tr"), gc.download(), ref):
        assert rel_err(a, b) <= TOL_REF, name
    check_pgrad(gp, refp, TOL_REF)


@pytest.mark.parametrize("kind,nstages", STEPPERS)
def test_step_vjp_finite_differences(kind, nstages):
    gm, gmodel, pd, state = setup(10, 10, 3, 2, seed=5)
    st = make_stepper(gmodel, kind, nstages)
    dt = 30.0
    w = random_like(state, 51)
    gc = upload(gm, 2, w)
    swe.step_vjp(st, 0.0, dt, upload(gm, 2, state), gc)
    grad = gc.download()

    def phi(y):
        s = upload(gm, 2, y)
        st.do_step(0.0, dt, s)
        return dot(s.download(), w)

    for k in range(3):
        v = random_like(state, 60 + k, (1e-3, 1e-6, 1e-3))
        fd = (phi(tuple(a + b for a, b in zip(state, v))) -
              phi(tuple(a - b for a, b in zip(state, v)))) / 2
        ad = dot(grad, v)
        assert abs(fd - ad) <= TOL_FD * abs(ad), (k, fd, ad)


# ---------------------------------------------------------------------------
# Checkpointed reversal and the misfit
# ---------------------------------------------------------------------------
def _schedule_forward_steps(nsteps, nsnap):
    """Forward steps of the library's schedule: one sweep, then the binomial
    reversal minus the advances of its first descent (stored by the sweep)."""
    nsnap = min(nsnap, max(1, nsteps))
    a, c = 0, nsnap
    while c > 1 and nsteps - a > 1:
        l = nsteps - a
        best = min(range(1, l), key=lambda m: (m + swe.revolve_cost(l - m, c - 1) +
                                               swe.revolve_cost(m, c), m))
        a += best
        c -= 1
    return nsteps + swe.revolve_cost(nsteps, nsnap) - a


@pytest.mark.parametrize("kind,nstages", [("rk4", 4), ("lsrk", 5)])
def test_tape_independent_of_snapshot_count(kind, nstages):
    gm, gmodel, pd, state = setup(8, 6, 3, 1, seed=8)
    st = make_stepper(gmodel, kind, nstages)
    dt, n = 25.0, 13
    w = random_like(state, 71)
    results = []
    for nsnap in (1, 3, 5, n):
        print("snapshots", nsnap, flush=True)
        tape = swe.Tape(st, n, nsnap)
        s = upload(gm, 1, state)
        tape.forward(0.0, dt, s)
        cot = upload(gm, 1, w)
        gp = tape.reverse(cot)
        info = tape.info()
        assert info["n_snapshots"] == nsnap
        assert info["forward_steps"] == _schedule_forward_steps(n, nsnap)
        results.append((cot.download(), gp, s.download()))
    # the same trajectory and the same adjoint, bit for bit
    for g, gp, yT in results[1:]:
        for a, b in zip(g, results[0][0]):
            assert same_bits(a, b)
        assert same_bits(gp, results[0][1])
        for a, b in zip(yT, results[0][2]):
            assert same_bits(a, b)
    # and equal to composing step adjoints over stored states
    ys = [upload(gm, 1, state)]
    for i in range(n):
        nxt = swe.ShallowWaterState(gm, 1)
        nxt.copy_from(ys[-1])
        st.do_step(i * dt, dt, nxt)
        ys.append(nxt)
    cot = upload(gm, 1, w)
    gp = np.zeros(4)
    for i in reversed(range(n)):
        gp = gp + swe.step_vjp(st, i * dt, dt, ys[i], cot)
    for a, b in zip(cot.download(), results[0][0]):
        assert same_bits(a, b)


def test_tape_byte_cap_limits_snapshots():
    gm, gmodel, pd, state = setup(6, 6, 2, 1)
    st = swe.RK4Stepper(gmodel)
    nbytes = sum(a.nbytes for a in state)
    tape = swe.Tape(st, 10, 8, byte_cap=3 * nbytes + 10)
    tape.forward(0.0, 10.0, upload(gm, 1, state))
    assert tape.info()["n_snapshots"] == 3
    assert tape.info()["snapshot_bytes"] == 3 * nbytes
    tiny = swe.Tape(st, 10, 8, byte_cap=nbytes // 2)
    with pytest.raises(swe.OmegaError):
        tiny.forward(0.0, 10.0, upload(gm, 1, state))


def test_misfit_matches_numpy():
    gm, gmodel, pd, state = setup(9, 7, 3, 2)
    obs = random_like(state, 81, (1000.0, 1.0, 1000.0))
    w = (2.0, 0.5, 3.0)
    cot = swe.ShallowWaterState(gm, 2)
    J = swe.misfit(gmodel, upload(gm, 2, state), upload(gm, 2, obs), cot, w)
    ac = np.asarray(gm.area_cell)
    ae = 0.5 * np.asarray(gm.dc_edge) * np.asarray(gm.dv_edge)
    areas = (ac[:, None], ae[:, None], ac[None, :, None])
    ref_cot = [(wf * a) * (y - o) for wf, a, y, o in zip(w, areas, state, obs)]
    ref_J = sum(np.sum(0.5 * (wf * a) * (y - o) ** 2) for wf, a, y, o in zip(w, areas, state, obs))
    for a, b in zip(cot.download(), ref_cot):
        assert same_bits(a, b)
    assert abs(J - ref_J) <= 1e-13 * abs(ref_J)


def _c3_like_case(L=4):
    gm, gmodel, pd, state = setup(12, 12, L, 2, seed=9)
    st = swe.RK4Stepper(gmodel)
    dt, n, w = 40.0, 16, (1.0, 0.0, 1.0)
    # observations: the forward run from a perturbed initial thickness
    rng = np.random.default_rng(91)
    pert = (state[0] + 1e-2 * rng.standard_normal(state[0].shape), state[1], state[2])
    obs = upload(gm, 2, pert)
    st.run(0.0, dt, n, obs)
    return gm, gmodel, pd, state, st, dt, n, w, obs


def _richardson(f, h):
    """Central difference with one Richardson step: error O(h^4)."""
    d1 = (f(h) - f(-h)) / (2 * h)
    d2 = (f(h / 2) - f(-h / 2)) / h
    return (4 * d2 - d1) / 3


def test_adjoint_run_finite_differences():
    """Multi-step misfit gradient (initial state and coefficients) against
    central differences of the forward model, as in the C3 workload but
    small: RK4, 16 steps, 4 snapshots."""
    gm, gmodel, pd, state, st, dt, n, w, obs = _c3_like_case()
    y0 = upload(gm, 2, state)
    grad = swe.ShallowWaterState(gm, 2)
    J, gp = swe.adjoint_run(st, 0.0, dt, n, y0, obs, grad, w, n_snapshots=4)
    for a, b in zip(y0.download(), state):   # the initial state is kept
        assert same_bits(a, b)
    g = grad.download()

    def J_of(y, model=gmodel):
        s = upload(gm, 2, y)
        stp = swe.RK4Stepper(model)
        stp.run(0.0, dt, n, s)
        c = swe.ShallowWaterState(gm, 2)
        return swe.misfit(model, s, obs, c, w)

    assert abs(J_of(state) - J) <= 1e-14 * J
    for k in range(4):
        d = random_like(state, 100 + k, (1e-3, 1e-6, 1e-3))
        fd = _richardson(lambda e: J_of(tuple(a + e * b for a, b in zip(state, d))), 1.0)
        ad = dot(g, d)
        print("direction", k, fd, ad, (fd - ad) / ad, flush=True)
        assert abs(fd - ad) <= TOL_FD_RUN * abs(ad), (k, fd, ad)
    # coefficient gradients: visc_del2 and eddy_diff4
    for idx, name in ((0, "visc_del2"), (3, "eddy_diff4")):
        def J_param(e):
            q = dict(pd)
            q[name] = pd[name] * (1 + e)
            return J_of(state, swe.ShallowWaterModel(gm, swe.ShallowWaterParams(**q)))
        # relative steps much below 1e-3 drown in the forward run's rounding
        # noise amplified by the misfit's cancellation
        fd = _richardson(J_param, 1e-3) / pd[name]
        print(name, fd, gp[idx], (fd - gp[idx]) / gp[idx], flush=True)
        assert abs(fd - gp[idx]) <= TOL_FD_RUN * abs(gp[idx]), (name, fd, gp[idx])


def test_adjoint_run_finite_differences_32_layers():
    """The same check with 32 layers, where the forward steps (and the
    adjoint's stage recomputation) run the pattern kernels with four cells
    per wavefront."""
    gm, gmodel, pd, state, st, dt, n, w, obs = _c3_like_case(32)
    assert gmodel.kernel_variant() == 2
    grad = swe.ShallowWaterState(gm, 2)
    J, gp = swe.adjoint_run(st, 0.0, dt, n, upload(gm, 2, state), obs, grad, w, n_snapshots=4)
    g = grad.download()

    def J_of(y):
        s = upload(gm, 2, y)
        swe.RK4Stepper(gmodel).run(0.0, dt, n, s)
        return swe.misfit(gmodel, s, obs, swe.ShallowWaterState(gm, 2), w)

    assert abs(J_of(state) - J) <= 1e-14 * J
    for k in range(2):
        d = random_like(state, 200 + k, (1e-3, 1e-6, 1e-3))
        fd = _richardson(lambda e: J_of(tuple(a + e * b for a, b in zip(state, d))), 1.0)
        ad = dot(g, d)
        print("direction", k, fd, ad, (fd - ad) / ad, flush=True)
        assert abs(fd - ad) <= TOL_FD_RUN * abs(ad), (k, fd, ad)


def test_adjoint_run_matches_autograd():
    """The same gradient against torch.autograd through 16 RK4 steps of the
    CPU restatement and the misfit (observations shared)."""
    gm, gmodel, pd, state, st, dt, n, w, obs = _c3_like_case()
    grad = swe.ShallowWaterState(gm, 2)
    J, gp = swe.adjoint_run(st, 0.0, dt, n, upload(gm, 2, state), obs, grad, w, n_snapshots=4)
    tm = ts.TorchMesh(gm, 4)
    o = [torch.from_numpy(a) for a in obs.download()]
    ac = torch.from_numpy(np.asarray(gm.area_cell))
    y = [torch.tensor(a, requires_grad=True) for a in state]
    params, plist = torch_params(pd)
    yt = tuple(y)
    for _ in range(n):
        yt = ts.rk4_step(tm, params, ALL_TERMS, dt, *yt)
    Jt = 0.5 * (w[0] * ac[:, None] * (yt[0] - o[0]) ** 2).sum() + \
        0.5 * (w[2] * ac[None, :, None] * (yt[2] - o[2]) ** 2).sum()
    gr = torch.autograd.grad(Jt, y + plist)
    assert abs(float(Jt.detach()) - J) <= 1e-10 * J
    for name, a, b in zip(("h", "vn", "tr"), grad.download(), gr[:3]):
        e = rel_err(a, b.numpy())
        print(name, e, flush=True)
        assert e <= TOL_RUN_REF, name
    for i in range(4):
        print(swe.PARAM_NAMES[i], gp[i], float(gr[3 + i]), flush=True)
        assert abs(gp[i] - float(gr[3 + i])) <= TOL_RUN_REF * abs(float(gr[3 + i])), i


def test_adjoint_is_run_to_run_reproducible():
    gm, gmodel, pd, state = setup(10, 10, 8, 2, seed=12)
    st = swe.LSRKStepper(gmodel)
    obs = upload(gm, 2, random_like(state, 5, (1.0, 0.0, 1.0)))
    outs = []
    for _ in range(2):
        grad = swe.ShallowWaterState(gm, 2)
        J, gp = swe.adjoint_run(st, 0.0, 20.0, 6, upload(gm, 2, state), obs, grad, n_snapshots=3)
        outs.append((J, gp, grad.download()))
    assert outs[0][0] == outs[1][0]
    assert same_bits(outs[0][1], outs[1][1])
    for a, b in zip(outs[0][2], outs[1][2]):
        assert same_bits(a, b)


def test_adjoint_rejects_hook():
    gm, gmodel, pd, state = setup(6, 6, 2, 1)

    class Forced(swe.ShallowWaterModel):
        def additional_tendency(self, h_tend, vn_tend, h, vn, t, stream):
            return 0

    m = Forced(gm, swe.ShallowWaterParams(**pd))
    s = upload(gm, 1, state)
    with pytest.raises(swe.OmegaError, match="hook"):
        swe.tendency_vjp(m, s, s, swe.ShallowWaterState(gm, 1))


# ---------------------------------------------------------------------------
# torch.autograd.Function over the library (test harness only)
# ---------------------------------------------------------------------------
class LibraryStep(torch.autograd.Function):
    """One library time step as an autograd node: forward runs the HIP step,
    backward the library's step adjoint (CPU float64 tensors in and out,
    original numbering)."""

    @staticmethod
    def forward(ctx, h, vn, tr, stepper, gm, dt):
        ntr = tr.shape[0]
        s = upload(gm, ntr, [x.detach().numpy() for x in (h, vn, tr)])
        ctx.saved = (stepper, gm, dt, [x.detach().numpy().copy() for x in (h, vn, tr)])
        stepper.do_step(0.0, dt, s)
        return tuple(torch.from_numpy(a) for a in s.download())

    @staticmethod
    def backward(ctx, gh, gvn, gtr):
        stepper, gm, dt, y = ctx.saved
        ntr = y[2].shape[0]
        cot = upload(gm, ntr, [g.contiguous().numpy() for g in (gh, gvn, gtr)])
        swe.step_vjp(stepper, 0.0, dt, upload(gm, ntr, y), cot)
        out = [torch.from_numpy(a) for a in cot.download()]
        return out[0], out[1], out[2], None, None, None


def test_autograd_function_wrapper_matches_torch_reference():
    gm, gmodel, pd, state = setup(8, 8, 3, 2, seed=14)
    st = swe.LSRKStepper(gmodel)
    dt = 25.0
    # two library steps chained through autograd, loss on the result
    y = [torch.tensor(a, requires_grad=True) for a in state]
    z = LibraryStep.apply(*y, st, gm, dt)
    z = LibraryStep.apply(*z, st, gm, dt)
    loss = (z[0] ** 2).sum() * 1e-6 + (z[1] * z[1]).sum() + (z[2] ** 3).sum() * 1e-9
    g_lib = torch.autograd.grad(loss, y)
    # the same through the torch restatement
    tm = ts.TorchMesh(gm, 3)
    om = po.Mesh(4, 4, DC, 1)
    a, b, _ = po.Stepper(po.Model(om, ntracers=0), "lsrk", 5).coeffs()
    y2 = [torch.tensor(s, requires_grad=True) for s in state]
    w = ts.lsrk_step(tm, pd, ALL_TERMS, dt, a, b, *y2)
    w = ts.lsrk_step(tm, pd, ALL_TERMS, dt, a, b, *w)
    loss2 = (w[0] ** 2).sum() * 1e-6 + (w[1] * w[1]).sum() + (w[2] ** 3).sum() * 1e-9
    g_ref = torch.autograd.grad(loss2, y2)
    assert abs(float(loss) - float(loss2)) <= 1e-12 * abs(float(loss2))
    for name, p, q in zip(("h", "vn", "tr"), g_lib, g_ref):
        assert rel_err(p.numpy(), q.numpy()) <= TOL_REF, name


ADVECTION_ONLY = ALL_TERMS & ~(swe.TDEL2 | swe.TDEL4)


@pytest.mark.parametrize("L", [4, 3])
@pytest.mark.parametrize("kind,nstages", [("rk4", 4), ("lsrk", 5)])
def test_step_vjp_tracer_advection_only(kind, nstages, L):
    """With no tracer diffusion term the edge pass hands the cell pass each
    edge's tracer advection share (advection-only kernels); the step adjoint
    still matches autograd, and the parameter gradient (visc_del2 formed in
    the vertex/cell pass) too."""
    gm, gmodel, pd, state = setup(8, 8, L, 2, seed=6, terms=ADVECTION_ONLY)
    st = make_stepper(gmodel, kind, nstages)
    dt = 30.0
    cot = random_like(state, 43)
    gc = upload(gm, 2, cot)
    gp = swe.step_vjp(st, 0.0, dt, upload(gm, 2, state), gc)
    tm = ts.TorchMesh(gm, L)
    if kind == "rk4":
        step = lambda p, *y: ts.rk4_step(tm, p, ADVECTION_ONLY, dt, *y)
    else:
        om = po.Mesh(4, 4, DC, 1)
        a, b, _ = po.Stepper(po.Model(om, ntracers=0), "lsrk", nstages).coeffs()
        step = lambda p, *y: ts.lsrk_step(tm, p, ADVECTION_ONLY, dt, a, b, *y)
    ref, refp = torch_grad((step, lambda: torch_params(pd)), state, cot)
    for name, x, y in zip(("h", "vn", "tr"), gc.download(), ref):
        assert rel_err(x, y) <= TOL_REF, name
    check_pgrad(gp, refp, TOL_REF)


def test_forward_unchanged_after_step_vjp():
    """A step adjoint keeps the stage diagnostics in sets of its own: a
    forward step and a tendency evaluated afterwards give the same bits as
    before it, and a repeated step adjoint gives the same bits."""
    gm, gmodel, pd, state = setup(8, 8, 4, 2, seed=8)
    st = swe.RK4Stepper(gmodel)
    dt = 30.0

    def forward():
        y = upload(gm, 2, state)
        st.do_step(0.0, dt, y)
        t = swe.ShallowWaterState(gm, 2)
        gmodel.compute_tendency(t, upload(gm, 2, state), 0.0)
        return y.download() + t.download()

    before = forward()
    cot = random_like(state, 47)
    outs = []
    for _ in range(2):
        gc = upload(gm, 2, cot)
        swe.step_vjp(st, 0.0, dt, upload(gm, 2, state), gc)
        outs.append(gc.download())
    after = forward()
    for a, b in zip(before, after):
        assert same_bits(a, b)
    for a, b in zip(*outs):
        assert same_bits(a, b)


@pytest.mark.parametrize("terms,ntr", [(ALL_TERMS, 2), (1 | 2 | 4 | 8, 0), (1 | 2 | 4 | 8 | 32, 3),
                                       (ALL_TERMS, 0), (1 | 2 | 4 | 8 | 64, 2), (1 | 2 | 4, 1)])
@pytest.mark.parametrize("L", [6, 64])
def test_tendency_vjp_slot_uniform_form_same_bits(terms, ntr, L):
    """On the regular hexagonal mesh the VJP reads its per-slot coefficients
    once per wavefront (kernel flag bit 2), and for the term sets 7, 15, 79
    and 127 has the terms compiled in (forms bit 2); forbidding those forms
    runs the run-time-terms and then the general gathers.  Same expressions
    and the same entities per thread, so the same bits, coefficient
    gradients included."""
    print("case", terms, ntr, L, flush=True)
    gm, gmodel, pd, state = setup(12, 10, L, ntr, terms=terms)
    cot = random_like(state, 21)
    gs, gc = upload(gm, ntr, state), upload(gm, ntr, cot)
    base = random_like(state, 22)
    outs, gps, flags = [], [], []
    for forms in (7, 3, 2):
        gmodel.set_kernel_forms(forms)
        out = upload(gm, ntr, base)
        gps.append(swe.tendency_vjp(gmodel, gs, gc, out))
        outs.append(out.download())
        flags.append(gmodel.kernel_flags())
    assert flags[0] & 4 and flags[1] & 4 and not flags[2] & 4, flags
    for o in outs[1:]:
        for name, a, b in zip(("h", "vn", "tr"), outs[0], o):
            assert same_bits(a, b), name
    assert same_bits(gps[0], gps[1]) and same_bits(gps[0], gps[2])
    # and the slot-uniform form against the torch reference
    gmodel.set_kernel_forms(3)
    tm = ts.TorchMesh(gm, L)
    ref, refp = torch_grad((lambda p, *y: ts.tendency(tm, p, terms, *y),
                            lambda: torch_params(pd)), state, cot)
    for name, a, b0, b in zip(("h", "vn", "tr"), outs[0], base, ref):
        if b.size:
            assert rel_err(a - b0, b) <= TOL_REF, name
    check_pgrad(gps[0], refp, TOL_REF)


@pytest.mark.parametrize("terms,ntr", [(ALL_TERMS, 2), (1 | 2 | 4 | 8, 0),
                                       (1 | 2 | 4 | 8 | 64, 2), (ADVECTION_ONLY, 1),
                                       (2 | 4 | 8 | 32, 1)])
@pytest.mark.parametrize("L,lpt", [(16, 2), (32, 2), (64, 1), (64, 2), (128, 2), (512, 2)])
@pytest.mark.parametrize("kind,nstages", [("rk4", 4), ("lsrk", 5), ("lsrk", 2)])
def test_step_vjp_stage_written_rows_same_bits(kind, nstages, L, lpt, terms, ntr):
    """A step adjoint's stage recomputation also writes the rows its tendency
    VJPs linearise about (kernel flag bit 4: edge PV sums, thickness fluxes
    and their edges-on-edge sums) with the one-cell-per-row tendency kernel
    (16 or 32 layers, one level per lane) and with the pair kernel (64k
    layers, its compiled term sets included), and those VJPs skip their
    pre-pass.  With that form forbidden (forms bit 3) the pre-pass forms the
    same rows: same bits for the state cotangents and the coefficient
    gradients."""
    print("case", kind, nstages, L, lpt, terms, ntr, flush=True)
    gm, gmodel, pd, state = setup(12, 10, L, ntr, seed=terms % 7, terms=terms)
    gmodel.set_launch(0, lpt)
    dt = 30.0
    cot = random_like(state, 51)
    outs, gps, flags = [], [], []
    for forms in (15, 7):
        gmodel.set_kernel_forms(forms)
        st = make_stepper(gmodel, kind, nstages)
        gc = upload(gm, ntr, cot)
        gps.append(swe.step_vjp(st, 0.0, dt, upload(gm, ntr, state), gc))
        outs.append(gc.download())
        flags.append(gmodel.kernel_flags())
    assert flags[0] & 16 and not flags[1] & 16, flags
    for name, a, b in zip(("h", "vn", "tr"), *outs):
        assert same_bits(a, b), name
    assert same_bits(gps[0], gps[1])


def test_stage_written_rows_off_where_not_written():
    """The rows are written by the one-cell-per-row kernel and by the pair
    kernel's shared-coefficient 32-bit form, with the vn tendency on: not by
    the pair kernel's per-cell or 64-bit forms, not with the vn tendency
    disabled, not on the general kernels."""
    gm, gmodel, pd, state = setup(8, 8, 64, 1)
    assert gmodel.kernel_flags() & 16
    for forms in (14, 13):
        gmodel.set_kernel_forms(forms)
        assert not gmodel.kernel_flags() & 16
        gmodel.set_launch(0, 1)
        assert gmodel.kernel_flags() & 16
        gmodel.set_launch(0, 2)
    gmodel.set_kernel_forms(15)
    gm, gmodel, pd, state = setup(8, 8, 32, 1)
    assert gmodel.kernel_flags() & 16
    gmodel.set_variant(1)
    assert not gmodel.kernel_flags() & 16
    pd = params_dict(DC, 1)
    pd["disable_vn_tendency"] = True
    gmodel = swe.ShallowWaterModel(gm, swe.ShallowWaterParams(**pd))
    assert not gmodel.kernel_flags() & 16


@pytest.mark.parametrize("kind,nstages", [("rk4", 4), ("lsrk", 5)])
def test_checkpointed_gradient_stage_written_rows_same_bits(kind, nstages):
    """A checkpointed 6-step gradient at 32 layers gives the same J, state
    gradient and coefficient gradient bits with the stage-written VJP rows
    as with the pre-pass."""
    gm, gmodel, pd, state = setup(12, 10, 32, 2, seed=9)
    obs = random_like(state, 61)
    res = []
    for forms in (15, 7):
        gmodel.set_kernel_forms(forms)
        st = make_stepper(gmodel, kind, nstages)
        grad = swe.ShallowWaterState(gm, 2)
        J, gp = swe.adjoint_run(st, 0.0, 30.0, 6, upload(gm, 2, state), upload(gm, 2, obs), grad,
                                (1.0, 0.5, 1.0), n_snapshots=3)
        res.append((J, gp, grad.download()))
    assert res[0][0] == res[1][0]
    assert same_bits(res[0][1], res[1][1])
    for name, a, b in zip(("h", "vn", "tr"), res[0][2], res[1][2]):
        assert same_bits(a, b), name


def test_step_vjp_without_coefficient_read_back():
    """step_vjp with pgrad=False does not read the coefficient gradient back
    (no wait for the stream); the state cotangent has the same bits."""
    gm, gmodel, pd, state = setup(8, 8, 32, 2, seed=3)
    st = swe.RK4Stepper(gmodel)
    cot = random_like(state, 71)
    outs = []
    for pg in (True, False):
        gc = upload(gm, 2, cot)
        r = swe.step_vjp(st, 0.0, 30.0, upload(gm, 2, state), gc, pgrad=pg)
        assert (r is None) == (not pg)
        swe.synchronize()
        outs.append(gc.download())
    for name, a, b in zip(("h", "vn", "tr"), *outs):
        assert same_bits(a, b), name
