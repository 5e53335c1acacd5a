"""The trajectory adjoint at the size of the adjoint workload, checked
against the forward model.

Workload: doubly periodic planar hexagonal mesh nx=ny=256 (65,536 cells),
dc = 10 km, 32 layers, 2 tracers, FP64, RK4, 200 steps, del2 viscosity
1e3 m^2/s, inertia-gravity-wave initial state with seeded noise; cost
J = 1/2 sum area (h_T - h_obs)^2 + 1/2 sum area (tr_T - tr_obs)^2 with the
observations from a run whose initial h was perturbed by a Gaussian random
field (1e-2 m, correlation length 5 dc); gradient with respect to the
initial h, normal velocity and tracers and to visc_del2, binomial
checkpointing with 10 snapshots.

Check: along 32 seeded random directions d (h and tracers with standard
deviation 1e-3, normal velocity 1e-6) the directional derivative
<grad, d> agrees with a central finite difference of J(y0 + e d) with one
Richardson step (e = 1 and 1/2) to 1e-7 relative (measured 6e-11 to
1.4e-8).  dJ/dvisc_del2 is checked the same way over visc_del2 +- 50 and
+- 25 m^2/s, to 2e-6 (measured 2.7e-7): J (~5e9, a sum over 4 million cell-layers after 200
steps) carries ~1e-3 of rounding noise while a shift of 1 m^2/s moves it by
only ~66, so a finite difference in the viscosity cannot resolve more at
this size (with +- 1 m^2/s it scatters by ~3e-5); the small-mesh tests in
test_gpu_adjoint.py pin the coefficient gradient to 1e-7 and better.  The
finite differences run the library's own forward steps, so the check needs
no CPU reference at this size.
"""
import numpy as np
import pytest

from omega_playground_amd import swe, workloads as wl

pytestmark = pytest.mark.gpu

TOL = 1e-7
TOL_VISC = 2e-6
VISC_STEP = 50.0
NDIR = 32


def _richardson(f, h):
    d1 = (f(h) - f(-h)) / (2 * h)
    d2 = (f(h / 2) - f(-h / 2)) / h
    return (4 * d2 - d1) / 3


def test_adjoint_workload_gradient_vs_finite_differences():
    cfg = dict(wl.CONFIGS["c3"])
    ntr, n = cfg["ntracers"], cfg["nsteps"]
    print("adjoint workload: nx", cfg["nx"], "layers", cfg["nlayers"], "steps", n, flush=True)
    mesh = swe.PlanarHexagonalMesh(cfg["nx"], cfg["nx"], cfg["dc"], cfg["nlayers"])

    def model_with(visc):
        p = swe.ShallowWaterParams(f0=cfg["f0"], grav=cfg["grav"], visc_del2=visc,
                                   eddy_diff4=cfg["eddy_diff4"], ntracers=ntr)
        m = swe.ShallowWaterModel(mesh, p)
        st = swe.RK4Stepper(m)
        st.set_graph(True)
        return m, st

    model, st = model_with(cfg["visc_del2"])
    dt = wl.dt_rule(cfg["dc"], cfg["grav"])
    y0 = wl.make_state(mesh, cfg, seed=0)
    obs = swe.ShallowWaterState(mesh, ntr)
    obs.upload(*wl.adjoint_observations(mesh, cfg, seed=0))
    st.run(0.0, dt, n, obs)

    state = swe.ShallowWaterState(mesh, ntr)
    state.upload(*y0)
    grad = swe.ShallowWaterState(mesh, ntr)
    J, gp = swe.adjoint_run(st, 0.0, dt, n, state, obs, grad, (1.0, 0.0, 1.0), n_snapshots=10)
    g = grad.download()
    scratch = swe.ShallowWaterState(mesh, ntr)
    work = swe.ShallowWaterState(mesh, ntr)

    def J_of(y, m=model, s=st):
        work.upload(*y)
        s.run(0.0, dt, n, work)
        return swe.misfit(m, work, obs, scratch, (1.0, 0.0, 1.0))

    assert abs(J_of(y0) - J) <= 1e-14 * J
    worst = 0.0
    for k in range(NDIR):
        rng = np.random.default_rng(1000 + k)
        d = tuple(s * rng.standard_normal(a.shape) for a, s in zip(y0, (1e-3, 1e-6, 1e-3)))
        fd = _richardson(lambda e: J_of(tuple(a + e * b for a, b in zip(y0, d))), 1.0)
        ad = sum(float(np.dot(x.ravel(), y.ravel())) for x, y in zip(g, d))
        rel = abs(fd - ad) / abs(ad)
        worst = max(worst, rel)
        print(f"direction {k}: fd {fd:.12e} adjoint {ad:.12e} rel {rel:.2e}", flush=True)
        assert rel <= TOL, (k, fd, ad)

    # the viscosity gradient: forward runs with visc_del2 shifted
    def J_visc(e):
        m, s = model_with(cfg["visc_del2"] + e)
        return J_of(y0, m, s)

    fdv = _richardson(J_visc, VISC_STEP)
    relv = abs(fdv - gp[0]) / abs(gp[0])
    print(f"visc_del2: fd {fdv:.12e} adjoint {gp[0]:.12e} rel {relv:.2e}; "
          f"worst direction {worst:.2e}", flush=True)
    assert relv <= TOL_VISC, (fdv, gp[0])
