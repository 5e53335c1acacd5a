"""The throughput workloads at their configured shapes, checked against the
oracle on the GPU:

* nx=64, 64 layers, 1 tracer, inertia-gravity wave with seeded noise,
  10 RK4 steps (graph-captured): bit-for-bit.
* nx=512, 64 layers, 4 tracers, Gaussian double vortex, del2 viscosity 1e3,
  tracer del4 1e9, 2 RK4 steps (graph-captured): bit-for-bit, and two runs
  from the same state give the same bits.
* the seeded 48-member ensemble (nx in {32..192}, 16/32/64 layers, 1-4
  tracers, double vortices with offset centres), 2 LSRK5 steps in one
  autotuned, graph-captured ensemble: every member bit-for-bit.
* 16 members at nx=128, 64 layers, 2 tracers, RK4, a short checkpointed
  trajectory under a memory cap: each member's gradient has the bits of its
  own single-member adjoint run, and the gradient agrees with central
  differences of the forward ensemble along 2 random directions.

The oracle runs on the host's cores (OpenMP); the sizes are the workloads'
own, the step counts cut to what the oracle finishes in seconds.
"""
import os

import numpy as np
import pytest

from oracle import pyoracle as po
from omega_playground_amd import swe
from omega_playground_amd import workloads as wl

from helpers import rel_err, same_bits

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def oracle_threads():
    po.set_num_threads(min(16, os.cpu_count() or 1))


def oracle_case(cfg, seed=0):
    om = po.Mesh(cfg["nx"], cfg["nx"], cfg["dc"], cfg["nlayers"])
    omodel = po.Model(om, f0=cfg["f0"], grav=cfg["grav"], visc_del2=cfg["visc_del2"],
                      eddy_diff4=cfg["eddy_diff4"], ntracers=cfg["ntracers"])
    return om, omodel, tuple(np.ascontiguousarray(a) for a in wl.make_state(om, cfg, seed=seed))


def gpu_case(cfg):
    gm = swe.PlanarHexagonalMesh(cfg["nx"], cfg["nx"], cfg["dc"], cfg["nlayers"])
    params = swe.ShallowWaterParams(f0=cfg["f0"], grav=cfg["grav"], visc_del2=cfg["visc_del2"],
                                    eddy_diff4=cfg["eddy_diff4"], ntracers=cfg["ntracers"])
    return gm, swe.ShallowWaterModel(gm, params)


def run_config_steps(name, nsteps):
    cfg = dict(wl.CONFIGS[name])
    print(f"config {name}: nx={cfg['nx']} L={cfg['nlayers']} ntr={cfg['ntracers']} "
          f"{nsteps} {cfg['stepper']} steps", flush=True)
    om, omodel, state = oracle_case(cfg)
    gm, gmodel = gpu_case(cfg)
    dt = wl.dt_rule(cfg["dc"], cfg["grav"])
    gs = swe.ShallowWaterState(gm, cfg["ntracers"])
    gs.upload(*state)
    st = swe.RK4Stepper(gmodel)
    st.set_graph(True)
    st.run(0.0, dt, nsteps, gs)
    got = gs.download()
    ref = tuple(a.copy() for a in state)
    po.Stepper(omodel, "rk4").run(0.0, dt, nsteps, ref)
    return cfg, gm, gmodel, st, state, got, ref, dt


def test_c1_inertia_gravity_wave_10_rk4_steps():
    cfg, gm, gmodel, st, state, got, ref, dt = run_config_steps("c1", 10)
    assert gmodel.kernel_variant() == 2
    for name, a, b in zip(("h", "vn", "tr"), got, ref):
        assert rel_err(a, b) <= 1e-12, name
        assert same_bits(a, b), name


def test_c2_double_vortex_2_rk4_steps_and_reproducible():
    cfg, gm, gmodel, st, state, got, ref, dt = run_config_steps("c2", 2)
    assert gmodel.kernel_variant() == 2
    for name, a, b in zip(("h", "vn", "tr"), got, ref):
        assert rel_err(a, b) <= 1e-12, name
        assert same_bits(a, b), name
    # run to run: the same two steps from the same state give the same bits
    gs2 = swe.ShallowWaterState(gm, cfg["ntracers"])
    gs2.upload(*state)
    st.run(0.0, dt, 2, gs2)
    for a, b in zip(gs2.download(), got):
        assert np.array_equal(a.view(np.int64), b.view(np.int64))


def test_c4_ensemble_48_members_2_lsrk5_steps():
    specs = wl.ensemble_members(48, seed=0)
    batch = swe.Batch("lsrk", 5)
    keep = []
    for sp in specs:
        gm = swe.PlanarHexagonalMesh(sp["nx"], sp["nx"], sp["dc"], sp["nlayers"])
        pd = dict(f0=1e-4, grav=9.81, visc_del2=1e3, eddy_diff4=1e9, ntracers=sp["ntracers"])
        gmodel = swe.ShallowWaterModel(gm, swe.ShallowWaterParams(**pd))
        om = po.Mesh(sp["nx"], sp["nx"], sp["dc"], sp["nlayers"])
        y0 = tuple(np.ascontiguousarray(a) for a in wl.gaussian_double_vortex(
            om, sp["nlayers"], sp["ntracers"], seed=sp["seed"], dh=sp["dh"],
            centres=sp["centres"]))
        gs = swe.ShallowWaterState(gm, sp["ntracers"])
        gs.upload(*y0)
        dt = wl.dt_rule(sp["dc"])
        batch.add(gmodel, gs, dt)
        keep.append((sp, gm, gmodel, gs, om, pd, y0, dt))
    batch.set_graph(True)
    batch.autotune(reps=1)      # restores the states
    print("ensemble launch", batch.get_launch(), flush=True)
    batch.run(2)
    for i, (sp, gm, gmodel, gs, om, pd, y0, dt) in enumerate(keep):
        ref = tuple(a.copy() for a in y0)
        po.Stepper(po.Model(om, **pd), "lsrk", 5).run(0.0, dt, 2, ref)
        for name, a, b in zip(("h", "vn", "tr"), gs.download(), ref):
            assert same_bits(a, b), (i, sp["nx"], sp["nlayers"], sp["ntracers"], name)


def test_c5_ensemble_adjoint_16_members():
    cfg = dict(wl.CONFIGS["c5"])
    M, L, ntr, nsteps = cfg["nmembers"], cfg["nlayers"], cfg["ntracers"], 6
    print(f"ensemble adjoint: {M} members nx={cfg['nx']} L={L} ntr={ntr} {nsteps} steps",
          flush=True)
    gm, _ = gpu_case(cfg)
    params = swe.ShallowWaterParams(f0=cfg["f0"], grav=cfg["grav"], visc_del2=cfg["visc_del2"],
                                    eddy_diff4=cfg["eddy_diff4"], ntracers=ntr)
    dt = wl.dt_rule(cfg["dc"], cfg["grav"])
    w = (1.0, 0.0, 1.0)
    models = [swe.ShallowWaterModel(gm, params) for _ in range(M)]
    y0 = [wl.make_state(gm, cfg, seed=m) for m in range(M)]
    obs, states = [], []
    batch = swe.Batch("rk4", 4)
    for m in range(M):
        o = swe.ShallowWaterState(gm, ntr)
        o.upload(*wl.adjoint_observations(gm, cfg, seed=m))
        swe.RK4Stepper(models[m]).run(0.0, dt, nsteps, o)
        obs.append(o)
        s = swe.ShallowWaterState(gm, ntr)
        s.upload(*y0[m])
        batch.add(models[m], s, dt)
        states.append(s)
    ens_bytes = M * (gm.ncells * L * (1 + ntr) + gm.nedges * L) * 8
    cap = int(2.5 * ens_bytes)   # two ensemble snapshots fit under the cap
    grads = [swe.ShallowWaterState(gm, ntr) for _ in range(M)]
    J, pgrad = batch.adjoint_run(0.0, nsteps, obs, grads, w, n_snapshots=10, byte_cap=cap)
    # each member against its own single-member adjoint run (weight 1/M)
    for m in range(M):
        g1 = swe.ShallowWaterState(gm, ntr)
        Jm, _ = swe.adjoint_run(swe.RK4Stepper(models[m]), 0.0, dt, nsteps, states[m], obs[m],
                                g1, tuple(x / M for x in w), n_snapshots=3)
        for name, a, b in zip(("h", "vn", "tr"), grads[m].download(), g1.download()):
            assert same_bits(a, b), (m, name)

    # central differences of the ensemble-mean misfit along random directions
    # of every member's initial state (h and tracers with standard deviation
    # 1e-3, normal velocity 1e-6), one Richardson step (e = 1 and 1/2)
    scratch = swe.ShallowWaterState(gm, ntr)
    work = swe.ShallowWaterState(gm, ntr)

    def ens_J(e, dirs):
        tot = 0.0
        for m in range(M):
            work.upload(*(a + e * d for a, d in zip(y0[m], dirs[m])))
            swe.RK4Stepper(models[m]).run(0.0, dt, nsteps, work)
            tot += swe.misfit(models[m], work, obs[m], scratch, tuple(x / M for x in w))
        return tot

    assert abs(ens_J(0.0, [tuple(np.zeros_like(a) for a in y) for y in y0]) - J) <= 1e-12 * J
    gall = [g.download() for g in grads]
    for r in range(2):
        rng = np.random.default_rng(1000 + r)
        dirs = [tuple(sc * rng.standard_normal(a.shape) for a, sc in zip(y, (1e-3, 1e-6, 1e-3)))
                for y in y0]
        ad = sum(float(np.dot(g.ravel(), d.ravel()))
                 for gg, dd in zip(gall, dirs) for g, d in zip(gg, dd))
        d1 = (ens_J(1.0, dirs) - ens_J(-1.0, dirs)) / 2
        d2 = (ens_J(0.5, dirs) - ens_J(-0.5, dirs)) / 1.0
        fd = (4 * d2 - d1) / 3
        rel = abs(ad - fd) / abs(ad)
        print(f"direction {r}: adjoint {ad:.12e} fd {fd:.12e} rel {rel:.2e}", flush=True)
        assert rel <= 1e-7, (r, ad, fd)
