#!/usr/bin/env python3
"""Throughput benchmark of the shallow-water hot path on one MI355X.

Workload (default): the forward drop-in at scale -- doubly periodic planar
hexagonal mesh nx = ny = 512 (262,144 cells, 786,432 edges), dc = 5 km,
64 layers, 4 tracers, FP64, RK4, Gaussian double vortex in geostrophic
balance, momentum del2 viscosity 1e3 m^2/s, tracer del4 1e9 m^4/s, the step
captured in a HIP graph.  A step is one full RK4 step (4 tendency
evaluations + stage updates); the metric counts cell-layer-steps/s.

Prints one JSON line.  `roofline` is measured on one RK stage (its
diagnostics launch and its fused tendency + stage-update launch) with HIP
events on their launch stream: the stage's algorithmic bytes over the two
kernels' summed average durations; `step_roofline` is the same bytes over
the whole graph-replayed step.  `cpu_baseline` times the CPU oracle (OpenMP,
the job's cores, spread binding) on a bounded sample of the same workload.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

# the CPU baseline's OpenMP threads are spread over the job's cores; libgomp
# reads these when the oracle library first loads
os.environ.setdefault("OMP_PROC_BIND", "spread")
os.environ.setdefault("OMP_PLACES", "threads")

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from omega_playground_amd import swe  # noqa: E402
from omega_playground_amd import workloads as wl  # noqa: E402

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, published


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def build_case(cfg_name: str, ordering: str = "hilbert"):
    cfg = dict(wl.CONFIGS[cfg_name])
    mesh = swe.PlanarHexagonalMesh(cfg["nx"], cfg["nx"], cfg["dc"], cfg["nlayers"], ordering=ordering)
    params = swe.ShallowWaterParams(f0=cfg["f0"], grav=cfg["grav"], visc_del2=cfg["visc_del2"],
                                    eddy_diff4=cfg["eddy_diff4"], ntracers=cfg["ntracers"])
    model = swe.ShallowWaterModel(mesh, params)
    state0 = wl.make_state(mesh, cfg, seed=0)
    return cfg, mesh, params, model, state0


def alg_bytes_per_cell_layer_step(ntr: int, stepper: str) -> float:
    """B_alg: compulsory state traffic per cell-layer-step, S = 8*(1+3+ntr)
    bytes (one h, three vn, ntr tracers per cell); RK4 16*S, LSRK5 18*S."""
    S = 8.0 * (4 + ntr)
    return (16.0 if stepper == "rk4" else 18.0) * S


def host_cpu():
    """CPU model, sockets and physical cores of the host (lscpu)."""
    info = {}
    try:
        import subprocess
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            k, _, v = line.partition(":")
            info[k.strip()] = v.strip()
    except (OSError, ValueError):
        pass
    sockets = int(info.get("Socket(s)", "0") or 0)
    per = int(info.get("Core(s) per socket", "0") or 0)
    return {"model": info.get("Model name", "unknown"), "sockets": sockets,
            "physical_cores": sockets * per, "logical_cpus": os.cpu_count()}


def cpu_baseline(cfg, sample_steps: int, threads: int, runs: int = 3):
    """The oracle (C++/OpenMP) on the same workload: the reference's protocol
    (one warm-up step, then the timed steps), the median of `runs` timings,
    threads spread over the job's cores (OMP_PROC_BIND=spread,
    OMP_PLACES=threads, set before the OpenMP runtime starts, as
    scripts/run_thread_scaling.sh:12-17 runs the reference)."""
    from oracle import pyoracle as po

    po.set_num_threads(threads)
    om = po.Mesh(cfg["nx"], cfg["nx"], cfg["dc"], cfg["nlayers"])
    omodel = po.Model(om, f0=cfg["f0"], grav=cfg["grav"], visc_del2=cfg["visc_del2"],
                      eddy_diff4=cfg["eddy_diff4"], ntracers=cfg["ntracers"])
    st = po.Stepper(omodel, cfg["stepper"])
    state = tuple(np.ascontiguousarray(a) for a in wl.make_state(om, cfg, seed=0))
    dt = wl.dt_rule(cfg["dc"], cfg["grav"])
    st.run(0.0, dt, 1, state)  # warm-up step (reference protocol)
    times = []
    for r in range(runs):
        t0 = time.perf_counter()
        st.run((1 + r * sample_steps) * dt, dt, sample_steps, state)
        times.append(time.perf_counter() - t0)
    el = float(np.median(times))
    work = om.ncells * om.nlayers * sample_steps
    cpu = host_cpu()
    return {
        "value": work / el,
        "unit": "cell-layer-steps/s",
        "cores": threads,
        "kind": "port",
        "sample": f"median of {runs} runs of {sample_steps} RK4 steps of the same "
                  f"nx={cfg['nx']} workload after one warm-up step (runs "
                  f"{', '.join(f'{t:.2f}' for t in times)} s), oracle C++/OpenMP, {threads} "
                  f"threads with OMP_PROC_BIND={os.environ.get('OMP_PROC_BIND')} "
                  f"OMP_PLACES={os.environ.get('OMP_PLACES')} on the job's CPU share "
                  f"(host: {cpu['model']}, {cpu['sockets']} socket(s), "
                  f"{cpu['physical_cores']} physical cores)",
        "host_cpu": cpu,
        "runs_s": times,
    }


def reference_benchmark(nx: int = 512, nlayers: int = 64, ntracers: int = 1, nsteps: int = 20,
                        graph: bool = True):
    """The reference's own benchmark shape (benchmark/benchmark.cpp:83-164):
    double vortex on lx = 5000 km, ly = lx*sqrt(3)/2 (dc = lx/nx), tracers
    1000, all four mixing coefficients 1e-5 (so del2 and del4 momentum and
    tracer mixing all run), LSRK5, dt = 0.1 dc / sqrt(g h0); one untimed step,
    then nsteps timed steps between device synchronisations."""
    dv = wl.DoubleVortexTest(benchmark=True)
    dc = dv.lx / nx
    mesh = swe.PlanarHexagonalMesh(nx, nx, dc, nlayers)
    params = swe.ShallowWaterParams(f0=dv.coriolis, grav=dv.g, visc_del2=1e-5, visc_del4=1e-5,
                                    eddy_diff2=1e-5, eddy_diff4=1e-5, ntracers=ntracers)
    model = swe.ShallowWaterModel(mesh, params)
    xc, yc = np.asarray(mesh.x_cell), np.asarray(mesh.y_cell)
    xe, ye, ae = np.asarray(mesh.x_edge), np.asarray(mesh.y_edge), np.asarray(mesh.angle_edge)
    h = np.repeat(dv.h(xc, yc)[:, None], nlayers, 1)
    vn = np.repeat(wl.normal_velocity(ae, dv.vx(xe, ye)[:, None], dv.vy(xe, ye)[:, None]),
                   nlayers, 1)
    tr = np.full((ntracers, mesh.ncells, nlayers), 1000.0)
    state = swe.ShallowWaterState(mesh, ntracers)
    state.upload(h, vn, tr)
    dt = 0.1 * dc / math.sqrt(dv.g * dv.h0)
    st = swe.LSRKStepper(model, 5)
    st.set_graph(graph)
    st.run(0.0, dt, 1, state)
    swe.synchronize()
    t0 = time.perf_counter()
    st.run(0.0, dt, nsteps, state, t_offset_steps=1)
    swe.synchronize()
    el = time.perf_counter() - t0
    # per-kernel times on eager steps
    st.set_graph(False)
    swe.timing_enable(True)
    swe.timing_reset()
    st.run(0.0, dt, 2, state)
    swe.synchronize()
    kern = {}
    for name in ("diag", "vdel4", "tend"):
        ms, n = swe.timing_read(name)
        kern[name] = {"avg_ms": ms / max(n, 1), "launches": n}
    swe.timing_enable(False)
    work = mesh.ncells * nlayers * nsteps
    B = alg_bytes_per_cell_layer_step(ntracers, "lsrk")
    value = work / el
    return {
        "workload": f"reference benchmark shape: double vortex, planar hex nx={nx} "
                    f"({mesh.ncells} cells), {nlayers} layers, {ntracers} tracer(s), LSRK5, "
                    f"visc_del2 = visc_del4 = eddy_diff2 = eddy_diff4 = 1e-5, {nsteps} steps",
        "seconds": el,
        "ms_per_step": 1e3 * el / nsteps,
        "cell_layer_steps_per_s": value,
        "B_alg_per_cell_layer_step": B,
        "roofline_frac": B * value / 1e9 / HBM_PEAK_GBS,
        "kernels": kern,
        "kernel_variant": model.kernel_variant(),
    }


def adjoint_bench(n_snapshots: int = 10, nsteps: int | None = None, graph: bool = True,
                  config: str = "c3"):
    """The adjoint workload: RK4 on nx=256, 32 layers, 2 tracers, del2
    viscosity on; misfit of h and tracers after the trajectory against
    observations from a run whose initial h was perturbed by a correlated
    Gaussian random field; gradient with respect to the initial state and
    the mixing coefficients with binomial checkpointing.  config "c2" runs
    the same on the forward workload's shape (nx=512, 64 layers, 4 tracers,
    del2 viscosity and tracer del4), for profiles of the 64-layer kernels."""
    cfg = dict(wl.CONFIGS[config])
    nsteps = nsteps or cfg["nsteps"]
    mesh = swe.PlanarHexagonalMesh(cfg["nx"], cfg["nx"], cfg["dc"], cfg["nlayers"])
    params = swe.ShallowWaterParams(f0=cfg["f0"], grav=cfg["grav"], visc_del2=cfg["visc_del2"],
                                    eddy_diff4=cfg["eddy_diff4"], ntracers=cfg["ntracers"])
    model = swe.ShallowWaterModel(mesh, params)
    ntr, L, nc = cfg["ntracers"], cfg["nlayers"], mesh.ncells
    dt = wl.dt_rule(cfg["dc"], cfg["grav"])
    st = swe.RK4Stepper(model)
    st.set_graph(graph)
    y0 = wl.make_state(mesh, cfg, seed=0)
    obs = swe.ShallowWaterState(mesh, ntr)
    obs.upload(*wl.adjoint_observations(mesh, cfg, seed=0))
    st.run(0.0, dt, nsteps, obs)
    state = swe.ShallowWaterState(mesh, ntr)
    state.upload(*y0)
    cot = swe.ShallowWaterState(mesh, ntr)
    cot.upload(*[np.ones_like(a) for a in y0])
    # per step: forward step vs step adjoint (stage recomputation included)
    st.run(0.0, dt, 2, state)
    swe.step_vjp(st, 0.0, dt, state, cot)
    swe.synchronize()
    reps = 20
    t0 = time.perf_counter()
    st.run(0.0, dt, reps, state)
    swe.synchronize()
    t_fwd = (time.perf_counter() - t0) / reps
    st.set_graph(False)
    # like the forward steps, the step adjoints are timed back to back: the
    # coefficient gradient is reduced on the device but not read back to the
    # host after every step (that read would wait for the stream each time)
    t0 = time.perf_counter()
    for _ in range(reps):
        swe.step_vjp(st, 0.0, dt, state, cot, pgrad=False)
    swe.synchronize()
    t_adj = (time.perf_counter() - t0) / reps
    st.set_graph(graph)
    # the whole trajectory: forward sweep + checkpointed reverse sweep
    state.upload(*y0)
    grad = swe.ShallowWaterState(mesh, ntr)
    swe.synchronize()
    t0 = time.perf_counter()
    J, pgrad = swe.adjoint_run(st, 0.0, dt, nsteps, state, obs, grad, (1.0, 0.0, 1.0),
                               n_snapshots=n_snapshots)
    swe.synchronize()
    t_run = time.perf_counter() - t0
    t0 = time.perf_counter()
    st.run(0.0, dt, nsteps, state)
    swe.synchronize()
    t_fwd_run = time.perf_counter() - t0
    work = nc * L * nsteps
    recompute = swe.revolve_cost(nsteps, n_snapshots)
    return {
        "workload": f"RK4 adjoint, planar hex nx={cfg['nx']} ({nc} cells), {L} layers, {ntr} "
                    f"tracers, del2 visc {cfg['visc_del2']:g}"
                    + (f", tracer del4 {cfg['eddy_diff4']:g}" if cfg["eddy_diff4"] else "")
                    + f", {nsteps} steps, {n_snapshots} snapshots, misfit of h and tracers vs "
                    f"perturbed-h0 observations",
        "fwd_ms_per_step": 1e3 * t_fwd,
        "step_adjoint_ms": 1e3 * t_adj,
        "ratio_per_step": t_adj / t_fwd,
        "forward_run_s": t_fwd_run,
        "adjoint_run_s": t_run,
        "ratio_total": t_run / t_fwd_run,
        "binomial_recompute_steps": recompute,
        "fwd_adj_cell_layer_steps_per_s": work / t_run,
        "J": J,
        "dJ_dvisc_del2": float(pgrad[0]),
    }


def batched_bench(nmembers: int = 48, nsteps: int = 50, specs=None, label=None, graph=True,
                  tune=True):
    """The ensemble workload: 48 seeded members (nx in {32..192}, 16/32/64
    layers, 1-4 tracers, double vortex with offset centres and random dH),
    LSRK5, 50 steps, del2 viscosity 1e3 and tracer del4 1e9 as the forward
    workload; one autotuned, graph-captured launch per stage phase for the
    whole ensemble against a sequential per-member loop (each member its own
    graph-captured stepper) on the same GPU.  Per-stage bandwidth: the
    members' algorithmic bytes (18 S per cell-layer-step for LSRK5, S =
    8 (4 + ntr) B) over the time, as a fraction of the 8 TB/s peak."""
    specs = specs if specs is not None else wl.ensemble_members(nmembers, seed=0)
    nmembers = len(specs)
    members = []
    for sp in specs:
        mesh = swe.PlanarHexagonalMesh(sp["nx"], sp["nx"], sp["dc"], sp["nlayers"])
        params = swe.ShallowWaterParams(f0=1e-4, grav=9.81, visc_del2=1e3, eddy_diff4=1e9,
                                        ntracers=sp["ntracers"])
        model = swe.ShallowWaterModel(mesh, params)
        y0 = wl.gaussian_double_vortex(mesh, sp["nlayers"], sp["ntracers"], seed=sp["seed"],
                                       dh=sp["dh"], centres=sp["centres"])
        members.append((mesh, model, sp, y0, wl.dt_rule(sp["dc"])))
    work = sum(m[0].ncells * m[2]["nlayers"] for m in members) * nsteps
    alg = sum(m[0].ncells * m[2]["nlayers"] * alg_bytes_per_cell_layer_step(m[2]["ntracers"], "lsrk")
              for m in members) * nsteps

    def fresh():
        out = []
        for mesh, model, sp, y0, dt in members:
            s = swe.ShallowWaterState(mesh, sp["ntracers"])
            s.upload(*y0)
            out.append(s)
        return out

    # batched
    states_b = fresh()
    batch = swe.Batch("lsrk", 5)
    for (mesh, model, sp, y0, dt), s in zip(members, states_b):
        batch.add(model, s, dt)
    batch.set_graph(graph)
    tune_ms = batch.autotune(reps=3) if tune else None
    batch.run(2)   # capture the graphs of both buffer parities
    swe.synchronize()
    for (mesh, model, sp, y0, dt), s in zip(members, states_b):
        s.upload(*y0)
    t0 = time.perf_counter()
    batch.run(nsteps)
    swe.synchronize()
    t_batch = time.perf_counter() - t0
    # looped: one member after another, each with its own graph-captured stepper
    states_l = fresh()
    steppers = []
    for mesh, model, sp, y0, dt in members:
        st = swe.LSRKStepper(model, 5)
        st.set_graph(graph)
        steppers.append(st)
    for st, (mesh, model, sp, y0, dt), s in zip(steppers, members, states_l):
        st.run(0.0, dt, 2, s)   # capture both buffer parities
        s.upload(*y0)
    swe.synchronize()
    t0 = time.perf_counter()
    for st, (mesh, model, sp, y0, dt), s in zip(steppers, members, states_l):
        st.run(0.0, dt, nsteps, s)
    swe.synchronize()
    t_loop = time.perf_counter() - t0
    # looped without graphs: each member's steps launched kernel by kernel,
    # as a loop over the reference's do_step calls would run them
    states_e = fresh()
    for st in steppers:
        st.set_graph(False)
    swe.synchronize()
    t0 = time.perf_counter()
    for st, (mesh, model, sp, y0, dt), s in zip(steppers, members, states_e):
        st.run(0.0, dt, nsteps, s)
    swe.synchronize()
    t_eager = time.perf_counter() - t0
    identical = all(
        all(np.array_equal(a, b) for a, b in zip(sb.download(), sl.download()))
        for sb, sl in zip(states_b, states_l))
    info = batch.info()
    return {
        "workload": label or (f"{nmembers} members, nx in {{32..192}}, 16/32/64 layers, 1-4 "
                              f"tracers, LSRK5, {nsteps} steps, double vortex, del2 visc 1e3, "
                              f"tracer del4 1e9"),
        "cell_layers": info["cell_layers"],
        "blocks_per_launch": info["blocks_per_launch"],
        "launch": list(batch.get_launch()),
        "persistent_blocks_per_cu": batch.get_persistent(),
        "batched_GBs": alg / t_batch / 1e9,
        "batched_frac": alg / t_batch / 1e9 / HBM_PEAK_GBS,
        "looped_GBs": alg / t_loop / 1e9,
        "looped_frac": alg / t_loop / 1e9 / HBM_PEAK_GBS,
        "autotune_ms_per_step": tune_ms,
        "batched_s": t_batch,
        "looped_s": t_loop,
        "speedup": t_loop / t_batch,
        "looped_eager_s": t_eager,
        "speedup_vs_eager_loop": t_eager / t_batch,
        "batched_cell_layer_steps_per_s": work / t_batch,
        "looped_cell_layer_steps_per_s": work / t_loop,
        "bit_identical": bool(identical),
    }


def ensemble_adjoint_bench(nsteps: int | None = None, n_snapshots: int = 10, overrides=None):
    """The ensemble adjoint workload: 16 members on one nx=128 mesh, 64
    layers, 2 tracers, RK4, shared del2 viscosity 1e3 and tracer del4 1e9;
    per-member seeded initial states and observations (forward runs from
    initial h perturbed by correlated Gaussian random fields); gradient of
    the ensemble-mean misfit with respect to every member's initial state
    and the shared coefficients, checkpoints capped at 8 GB.  One stacked
    sweep for the ensemble against a loop of per-member adjoint runs with
    the same snapshot count (the one the cap leaves the ensemble) and cap."""
    cfg = dict(wl.CONFIGS["c5"])
    cfg.update(overrides or {})
    nsteps = nsteps or cfg["nsteps"]
    M, ntr, L = cfg["nmembers"], cfg["ntracers"], cfg["nlayers"]
    mesh = swe.PlanarHexagonalMesh(cfg["nx"], cfg["nx"], cfg["dc"], L)
    params = swe.ShallowWaterParams(f0=cfg["f0"], grav=cfg["grav"], visc_del2=cfg["visc_del2"],
                                    eddy_diff4=cfg["eddy_diff4"], ntracers=ntr)
    dt = wl.dt_rule(cfg["dc"], cfg["grav"])
    w = (1.0, 0.0, 1.0)
    models = [swe.ShallowWaterModel(mesh, params) for _ in range(M)]
    y0 = [wl.make_state(mesh, cfg, seed=m) for m in range(M)]
    obs, states, obatch, batch = [], [], swe.Batch("rk4", 4), swe.Batch("rk4", 4)
    for m in range(M):
        o = swe.ShallowWaterState(mesh, ntr)
        o.upload(*wl.adjoint_observations(mesh, cfg, seed=m))
        obatch.add(models[m], o, dt)
        obs.append(o)
        s = swe.ShallowWaterState(mesh, ntr)
        s.upload(*y0[m])
        batch.add(models[m], s, dt)
        states.append(s)
    obatch.run(nsteps)
    swe.synchronize()
    grads = [swe.ShallowWaterState(mesh, ntr) for _ in range(M)]
    t0 = time.perf_counter()
    J, pgrad = batch.adjoint_run(0.0, nsteps, obs, grads, w, n_snapshots=n_snapshots,
                                 byte_cap=cfg["byte_cap"])
    swe.synchronize()
    t_batch = time.perf_counter() - t0
    ens_bytes = M * (mesh.ncells * L * (1 + ntr) + mesh.nedges * L) * 8
    snaps = max(1, min(n_snapshots, cfg["byte_cap"] // ens_bytes))
    # per-member loop with the ensemble's snapshot count (the cap leaves the
    # ensemble 9 of the 10 snapshots; a member alone would get all 10 and so
    # recompute fewer steps), hence the same checkpoint schedule per member
    J_loop, pg_loop, max_diff = 0.0, np.zeros(4), 0.0
    t0 = time.perf_counter()
    for m in range(M):
        st = swe.RK4Stepper(models[m])
        st.set_graph(True)
        g = swe.ShallowWaterState(mesh, ntr)
        Jm, pgm = swe.adjoint_run(st, 0.0, dt, nsteps, states[m], obs[m], g,
                                  tuple(x / M for x in w), n_snapshots=int(snaps),
                                  byte_cap=cfg["byte_cap"])
        J_loop += Jm
        pg_loop += pgm
        if m == 0:
            g0 = g
    swe.synchronize()
    t_loop = time.perf_counter() - t0
    a, b = g0.download(), grads[0].download()
    max_diff = max(float(np.max(np.abs(x - y))) for x, y in zip(a, b))
    return {
        "workload": f"ensemble adjoint, {M} members, planar hex nx={cfg['nx']} ({mesh.ncells} "
                    f"cells), {L} layers, {ntr} tracers, RK4, {nsteps} steps, del2 visc "
                    f"{cfg['visc_del2']:g} and tracer del4 {cfg['eddy_diff4']:g} shared, "
                    f"checkpoints <= {cfg['byte_cap'] / 1e9:g} GB",
        "snapshots": int(snaps),
        "forward_steps_recomputed": swe.revolve_cost(nsteps, int(snaps)),
        "batched_s": t_batch,
        "looped_s": t_loop,
        "speedup": t_loop / t_batch,
        "J": J,
        "J_loop_rel_diff": abs(J - J_loop) / abs(J_loop),
        "dJ_dvisc_del2": float(pgrad[0]),
        "dJ_deddy_diff4": float(pgrad[3]),
        "pgrad_loop_rel_diff": float(np.max(np.abs(pgrad - pg_loop) / np.maximum(np.abs(pg_loop), 1e-300))),
        "member0_grad_max_abs_diff": max_diff,
    }


def copy_calibration():
    """A plain device copy of known size, for scaling the HBM byte counters:
    1 GiB read + 1 GiB written per copy kernel."""
    import torch

    n = (1 << 30) // 8
    a = torch.ones(n, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read": n * 8, "copy_bytes_written": n * 8, "copies": 3}))


def load_traffic(kernel: str, workload: str):
    """HBM bytes per launch of a kernel from the committed rocprofv3 counter
    summary, measured on this workload: at this launch shape when recorded,
    else at another launch shape of the same workload and ordering (the
    kernels and their data are the same; `launch` says which was measured)."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as f:
            rec = json.load(f)
    except (OSError, ValueError):
        return None
    key = workload if kernel in rec.get(workload, {}) else next(
        (k for k in rec if k.rsplit(":", 1)[0] == workload.rsplit(":", 1)[0]
         and kernel in rec[k]), None)
    if key is None:
        return None
    ent = dict(rec[key][kernel])
    ent["launch"] = key.rsplit(":", 1)[1]
    return ent


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--config", default="c2")
    ap.add_argument("--ordering", default="hilbert")
    ap.add_argument("--no-graph", action="store_true")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=4)
    ap.add_argument("--skip-adjoint", action="store_true")
    ap.add_argument("--skip-batched", action="store_true")
    ap.add_argument("--skip-ensemble-adjoint", action="store_true")
    ap.add_argument("--c5-steps", type=int, default=500,
                    help="steps of the ensemble adjoint leg (the workload's length)")
    ap.add_argument("--skip-reference-shape", action="store_true")
    ap.add_argument("--batched-steps", type=int, default=50)
    ap.add_argument("--adjoint-config", choices=("c3", "c2"), default="c3",
                    help="adjoint leg on its own workload (c3) or on the forward workload's "
                         "shape (c2: nx=512, 64 layers, 4 tracers; for profiles)")
    ap.add_argument("--adjoint-steps", type=int, default=None,
                    help="trajectory length of the adjoint leg (default: the workload's 200)")
    ap.add_argument("--cpu-runs", type=int, default=3)
    ap.add_argument("--copy-calibration", action="store_true")
    ap.add_argument("--launch", default="", help="rows,lpt,swizzle,levels_per_block")
    ap.add_argument("--no-tune", action="store_true",
                    help="keep the default launch shapes (model and ensembles)")
    ap.add_argument("--variant", type=int, default=None,
                    help="0 specialised if possible, 1 general")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: there is no multi-GPU path (layers shard with no "
              "exchange, the horizontal split needs one halo exchange per stage); run with "
              "--gpus 1",
              file=sys.stderr)
        sys.exit(2)

    if args.copy_calibration:
        copy_calibration()
        return

    cfg, mesh, params, model, state0 = build_case(args.config, args.ordering)
    if args.launch:
        model.set_launch(*(int(x) for x in args.launch.split(",")))
    if args.variant is not None:
        model.set_variant(args.variant)
    ntr, L, nc = cfg["ntracers"], cfg["nlayers"], mesh.ncells
    dt = wl.dt_rule(cfg["dc"], cfg["grav"])
    state = swe.ShallowWaterState(mesh, ntr)
    state.upload(*state0)
    stepper = swe.RK4Stepper(model) if cfg["stepper"] == "rk4" else swe.LSRKStepper(model, 5)
    stepper.set_graph(not args.no_graph)
    tune_ms = None
    if not args.launch and not args.no_tune:
        tune_ms = model.autotune(state, reps=10)
        log(f"[bench] autotuned launch {model.get_launch()} ({tune_ms:.3f} ms per tendency)")
    log(f"[bench] {args.config}: nx={cfg['nx']} cells={nc} L={L} ntr={ntr} dt={dt:.4f}s "
        f"launch={model.get_launch()} kernels={model.kernel_variant()} graph={not args.no_graph}")

    # warm-up
    stepper.run(0.0, dt, args.warmup, state)
    swe.synchronize()
    # timed region: K steps, device-synchronised on both sides
    t0 = time.perf_counter()
    stepper.run(args.warmup * dt, dt, args.steps, state)
    swe.synchronize()
    elapsed = time.perf_counter() - t0
    ms_per_step = 1e3 * elapsed / max(args.steps, 1)
    work = nc * L * args.steps
    value = work / elapsed
    h, vn, tr = state.download()
    finite = bool(np.isfinite(h).all() and np.isfinite(vn).all() and np.isfinite(tr).all())
    log(f"[bench] {args.steps} steps in {elapsed:.4f}s = {ms_per_step:.3f} ms/step, "
        f"{value:.4e} cell-layer-steps/s, finite={finite}")

    # per-kernel durations with HIP events on the launch stream (eager steps)
    ksteps = max(2, min(10, args.steps))
    stepper.set_graph(False)
    swe.timing_enable(True)
    swe.timing_reset()
    stepper.run(0.0, dt, ksteps, state)
    swe.synchronize()
    tend_ms, tend_n = swe.timing_read("tend")
    diag_ms, diag_n = swe.timing_read("diag")
    swe.timing_enable(False)
    avg_tend_ms = tend_ms / max(tend_n, 1)
    avg_diag_ms = diag_ms / max(diag_n, 1)
    B = alg_bytes_per_cell_layer_step(ntr, cfg["stepper"])
    # one RK stage is the unit of launch: its diagnostics launch and its fused
    # tendency + update launch carry the stage's share of B_alg (a quarter of
    # the step's for RK4); the headline roofline charges those bytes to the
    # two kernels' summed average durations
    nlaunch = 4 if cfg["stepper"] == "rk4" else 5
    alg_bytes_launch = B * nc * L / nlaunch
    stage_ms = avg_tend_ms + avg_diag_ms
    achieved = alg_bytes_launch / (stage_ms * 1e-3) / 1e9
    tend_only = alg_bytes_launch / (avg_tend_ms * 1e-3) / 1e9
    step_achieved = B * value / 1e9
    log(f"[bench] tend kernel avg {avg_tend_ms:.4f} ms x{tend_n}, diag avg {avg_diag_ms:.4f} ms "
        f"x{diag_n}; B_alg {B:.0f} B/cell-layer-step")

    adj = None
    if not args.skip_adjoint:
        log("[bench] adjoint workload")
        adj = adjoint_bench(nsteps=args.adjoint_steps, graph=not args.no_graph,
                            config=args.adjoint_config)
        log(f"[bench] adjoint: step ratio {adj['ratio_per_step']:.2f}, trajectory ratio "
            f"{adj['ratio_total']:.2f} ({adj['adjoint_run_s']:.3f} s)")

    bat = None
    if not args.skip_batched:
        log("[bench] ensemble workload")
        bat = batched_bench(nsteps=args.batched_steps, graph=not args.no_graph,
                            tune=not args.no_tune)
        log(f"[bench] ensemble: batched {bat['batched_s']:.3f} s vs looped {bat['looped_s']:.3f} s "
            f"= {bat['speedup']:.2f}x, bit-identical {bat['bit_identical']}, launch {bat['launch']}, "
            f"{bat['batched_frac']:.3f} vs {bat['looped_frac']:.3f} of peak")
        small = [dict(nx=32, dc=10e3, nlayers=16, ntracers=2,
                      centres=((0.4, 0.5), (0.6, 0.5)), dh=75.0, seed=i) for i in range(48)]
        bat["small_ensemble"] = batched_bench(
            nsteps=args.batched_steps, specs=small, graph=not args.no_graph,
            tune=not args.no_tune,
            label=f"48 members, nx=32 (1024 cells), 16 layers, 2 tracers, LSRK5, "
                  f"{args.batched_steps} steps, double vortex, del2 visc 1e3, tracer del4 1e9")
        sm = bat["small_ensemble"]
        log(f"[bench] small ensemble: batched {sm['batched_s']:.4f} s vs looped "
            f"{sm['looped_s']:.4f} s = {sm['speedup']:.2f}x (eager loop {sm['speedup_vs_eager_loop']:.2f}x)")

    eadj = None
    if not args.skip_ensemble_adjoint:
        log("[bench] ensemble adjoint workload")
        eadj = ensemble_adjoint_bench(args.c5_steps)
        log(f"[bench] ensemble adjoint: batched {eadj['batched_s']:.3f} s vs looped "
            f"{eadj['looped_s']:.3f} s = {eadj['speedup']:.2f}x, {eadj['snapshots']} snapshots")
        # where stacking pays: 16 small members (nx=32, 16 layers), whose
        # per-member runs cannot fill the GPU
        sm = ensemble_adjoint_bench(100, overrides=dict(nx=32, nlayers=16))
        eadj["small_ensemble_adjoint"] = sm
        log(f"[bench] small ensemble adjoint: batched {sm['batched_s']:.4f} s vs looped "
            f"{sm['looped_s']:.4f} s = {sm['speedup']:.2f}x, {sm['snapshots']} snapshots")

    refb = None
    if not args.skip_reference_shape:
        log("[bench] reference benchmark shape")
        refb = reference_benchmark(graph=not args.no_graph)
        log(f"[bench] reference shape: {refb['ms_per_step']:.3f} ms/step, "
            f"{refb['cell_layer_steps_per_s']:.4e} cell-layer-steps/s")

    cpu = None
    if not args.no_cpu_baseline:
        # the job's CPU share: OMP_NUM_THREADS when set (the GPU box sets it to
        # the cores of its share), else every logical CPU
        threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or (os.cpu_count() or 1)
        log(f"[bench] CPU baseline: {args.cpu_runs} x {args.cpu_steps} steps, {threads} threads")
        cpu = cpu_baseline(cfg, args.cpu_steps, threads, args.cpu_runs)
        log(f"[bench] CPU {cpu['value']:.4e} cell-layer-steps/s")

    workload_key = f"{args.config}:{args.ordering}:" + ",".join(str(x) for x in model.get_launch())
    traffic = load_traffic("tend", workload_key)
    dtraffic = load_traffic("diag", workload_key)
    # measured HBM bytes per stage (rocprofv3 FETCH_SIZE + WRITE_SIZE of both
    # kernels, committed in profiles/traffic.json) over this run's durations
    stage_traffic = None
    if traffic and dtraffic:
        stage_traffic = traffic["bytes_per_launch"] + dtraffic["bytes_per_launch"]
    traffic_gbs = stage_traffic / (stage_ms * 1e-3) / 1e9 if stage_traffic else None
    step_traffic_gbs = nlaunch * stage_traffic / (ms_per_step * 1e-3) / 1e9 \
        if stage_traffic else None

    out = {
        "metric": "cell-layer-steps/s FP64 fwd+adjoint at 1 GPU; adjoint/fwd ratio; batched speedup",
        "value": value,
        "unit": "cell-layer-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"forward RK4, planar hex nx={cfg['nx']} ({nc} cells), dc="
                               f"{cfg['dc'] / 1e3:g} km, {L} layers, {ntr} tracers, Gaussian double "
                               f"vortex, del2 visc {cfg['visc_del2']:g}, tracer del4 "
                               f"{cfg['eddy_diff4']:g}, hipGraph step",
                   "ncells": nc, "nlayers": L, "ntracers": ntr, "stepper": cfg["stepper"],
                   "graph": not args.no_graph, "parallelism": "single GPU",
                   "launch": list(model.get_launch()), "autotune_ms_per_tendency": tune_ms},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": stage_traffic,
                     "traffic_source": "profiles/traffic.json (rocprofv3 --pmc FETCH_SIZE, "
                                       "WRITE_SIZE; scaled by a plain copy), measured at "
                                       f"launch {traffic['launch']}"
                     if stage_traffic else None,
                     "kernel": "one RK stage: k_diag (diagnostics) + k_tend (fused tendency + "
                               "RK4 stage update)",
                     "stage_avg_ms": stage_ms,
                     "alg_bytes_per_launch": alg_bytes_launch,
                     "traffic_GBs": traffic_gbs,
                     "traffic_frac": traffic_gbs / HBM_PEAK_GBS if traffic_gbs else None,
                     "kernels": {
                         "tend": {"avg_ms": avg_tend_ms, "launches": tend_n,
                                  "alg_frac_if_charged_alone": tend_only / HBM_PEAK_GBS,
                                  "traffic": traffic["bytes_per_launch"] if traffic else None},
                         "diag": {"avg_ms": avg_diag_ms, "launches": diag_n,
                                  "traffic": dtraffic["bytes_per_launch"] if dtraffic else None}}},
        "step_roofline": {"B_alg_per_cell_layer_step": B, "achieved_GBs": step_achieved,
                          "frac": step_achieved / HBM_PEAK_GBS, "diag_kernel_avg_ms": avg_diag_ms,
                          "traffic_GBs": step_traffic_gbs,
                          "traffic_frac": step_traffic_gbs / HBM_PEAK_GBS
                          if step_traffic_gbs else None},
        "finite": finite,
        "adjoint": adj,
        "batched": bat,
        "ensemble_adjoint": eadj,
        "reference_shape": refb,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
