"""GPU tests of ensembles: members with different meshes, layer counts,
tracer counts, coefficients and time steps advanced together must end
bit-identical to stepping each member alone (which the parity tests tie to
the oracle), with and without graph capture, after autotuning, and when a
member is also stepped on its own in between."""
import numpy as np
import pytest

from oracle import pyoracle as po
from omega_playground_amd import swe

from helpers import params_dict, rich_state, same_bits

pytestmark = pytest.mark.gpu

# (nx, ny, nlayers, ntracers, dc, params overrides, dt)
MEMBERS = [
    (8, 6, 4, 1, 10e3, dict(eddy_diff2=0.0, eddy_diff4=0.0, visc_del4=0.0), 20.0),
    (10, 10, 6, 3, 10e3, {}, 15.0),
    (12, 8, 2, 0, 5e3, dict(visc_del4=0.0), 10.0),
    (6, 6, 8, 2, 20e3, dict(disable_vn_tendency=True), 30.0),
    (8, 8, 64, 1, 10e3, dict(visc_del2=0.0), 12.0),
    (10, 8, 32, 2, 10e3, {}, 14.0),
    (8, 6, 16, 1, 10e3, dict(visc_del4=0.0), 18.0),
]


def make_members(members, seed=0):
    out = []
    for i, (nx, ny, L, ntr, dc, over, dt) in enumerate(members):
        pd = params_dict(dc, ntr)
        pd.update(over)
        mesh = swe.PlanarHexagonalMesh(nx, ny, dc, L)
        model = swe.ShallowWaterModel(mesh, swe.ShallowWaterParams(**pd))
        y0 = rich_state(mesh, L, ntr, seed + i)
        out.append((mesh, model, ntr, y0, dt, pd))
    return out


def state_of(mesh, ntr, y):
    s = swe.ShallowWaterState(mesh, ntr)
    s.upload(*y)
    return s


def solo(members, kind, nstages, nsteps):
    res = []
    for mesh, model, ntr, y0, dt, _ in members:
        st = swe.RK4Stepper(model) if kind == "rk4" else swe.LSRKStepper(model, nstages)
        s = state_of(mesh, ntr, y0)
        st.run(0.0, dt, nsteps, s)
        res.append(s.download())
    return res


ODD = [(7, 5, 5, 2, 10e3, {}, 25.0)]


@pytest.mark.parametrize("kind,nstages", [("lsrk", 5), ("rk4", 4), ("lsrk", 1)])
@pytest.mark.parametrize("graph", [False, True])
@pytest.mark.parametrize("odd", [False, True])
def test_batch_matches_members_stepped_alone(kind, nstages, graph, odd):
    print("case", kind, nstages, graph, odd, flush=True)
    members = make_members(MEMBERS + (ODD if odd else []))
    ref = solo(members, kind, nstages, 3)
    b = swe.Batch(kind, nstages)
    states = []
    for mesh, model, ntr, y0, dt, _ in members:
        s = state_of(mesh, ntr, y0)
        b.add(model, s, dt)
        states.append(s)
    b.set_graph(graph)
    b.run(3)
    assert b.get_launch()[1] == (1 if odd else 2)
    for i, (s, r) in enumerate(zip(states, ref)):
        for name, a, c in zip(("h", "vn", "tr"), s.download(), r):
            assert same_bits(a, c), (i, name)


def test_batch_kernel_forms_per_member():
    """Two 64-layer members, one allowed the shared-coefficient and 32-bit
    forms of the pair kernels and one forced to the general forms, in one
    batch: each ends with the bits of its solo steps and of the oracle."""
    members = make_members([(8, 8, 64, 2, 10e3, {}, 12.0), (10, 8, 64, 1, 10e3, {}, 15.0)])
    members[1][1].set_kernel_forms(0)
    assert members[0][1].kernel_flags() & 3 == 3 and members[1][1].kernel_flags() == 0
    ref = solo(members, "rk4", 4, 3)
    b = swe.Batch("rk4", 4)
    states = []
    for mesh, model, ntr, y0, dt, _ in members:
        s = state_of(mesh, ntr, y0)
        b.add(model, s, dt)
        states.append(s)
    b.run(3)
    for i, (s, r) in enumerate(zip(states, ref)):
        for name, a, c in zip(("h", "vn", "tr"), s.download(), r):
            assert same_bits(a, c), (i, name)
    mesh, model, ntr, y0, dt, pd = members[1]
    oref = tuple(a.copy() for a in y0)
    po.Stepper(po.Model(po.Mesh(10, 8, 10e3, 64), **pd), "rk4").run(0.0, dt, 3, oref)
    for name, a, c in zip(("h", "vn", "tr"), states[1].download(), oref):
        assert same_bits(a, c), name


def test_batch_member_matches_oracle():
    members = make_members(MEMBERS[1:2])
    mesh, model, ntr, y0, dt, pd = members[0]
    b = swe.Batch("lsrk", 5)
    s = state_of(mesh, ntr, y0)
    b.add(model, s, dt)
    b.run(2)
    om = po.Mesh(10, 10, 10e3, 6)
    ref = tuple(a.copy() for a in y0)
    po.Stepper(po.Model(om, **pd), "lsrk", 5).run(0.0, dt, 2, ref)
    for name, a, c in zip(("h", "vn", "tr"), s.download(), ref):
        assert same_bits(a, c), name


def test_batch_autotune_and_interleaved_solo_steps():
    members = make_members(MEMBERS, seed=4)
    ref = solo(members, "lsrk", 5, 4)
    b = swe.Batch("lsrk", 5)
    b.set_graph(True)
    states = []
    for mesh, model, ntr, y0, dt, _ in members:
        s = state_of(mesh, ntr, y0)
        b.add(model, s, dt)
        states.append(s)
    ms = b.autotune(reps=2)
    assert ms > 0
    threads, lpt = b.get_launch()
    assert threads in (64, 128, 256) and lpt in (1, 2)
    b.step()
    # the second step: member 1 alone, the others in a second ensemble; the
    # first ensemble then finds the buffers flipped and rebuilds its tables
    mesh1, model1, ntr1, _, dt1, _ = members[1]
    swe.LSRKStepper(model1).do_step(0.0, dt1, states[1])
    b2 = swe.Batch("lsrk", 5)
    for i, (mesh, model, ntr, y0, dt, _) in enumerate(members):
        if i != 1:
            b2.add(model, states[i], dt)
    b2.step()
    b.run(2)   # the first ensemble rebuilds its tables for the new buffer parity
    for i, (s, r) in enumerate(zip(states, ref)):
        for name, a, c in zip(("h", "vn", "tr"), s.download(), r):
            assert same_bits(a, c), (i, name)


def test_batch_rejects_hook_and_shared_state():
    members = make_members(MEMBERS[:1])
    mesh, model, ntr, y0, dt, pd = members[0]

    class Forced(swe.ShallowWaterModel):
        def additional_tendency(self, h_tend, vn_tend, h, vn, t, stream):
            return 0

    b = swe.Batch()
    s = state_of(mesh, ntr, y0)
    with pytest.raises(swe.OmegaError, match="hook"):
        b.add(Forced(mesh, swe.ShallowWaterParams(**pd)), s, dt)
    b.add(model, s, dt)
    with pytest.raises(swe.OmegaError, match="one member"):
        b.add(model, s, dt)


@pytest.mark.parametrize("threads", [256, 128, 64])
@pytest.mark.parametrize("persistent", [0, 1, 4])
def test_batch_launch_shapes_and_persistent_grid(threads, persistent):
    """Every cells-per-block choice and the persistent grid (blocks walking
    runs of table entries; 1 block per CU is far fewer blocks than entries)
    give each member the bits of its solo run; general-kernel members (odd
    layer counts, 2-, 4-, 6- and 8-layer meshes) and pattern-specialised ones
    share each launch."""
    print("launch", threads, persistent, flush=True)
    members = make_members(MEMBERS)
    ref = solo(members, "lsrk", 5, 2)
    b = swe.Batch("lsrk", 5)
    states = []
    for mesh, model, ntr, y0, dt, _ in members:
        s = state_of(mesh, ntr, y0)
        b.add(model, s, dt)
        states.append(s)
    b.set_launch(threads, 2)
    b.set_persistent(persistent)
    b.set_graph(True)
    b.run(2)
    for i, (s, r) in enumerate(zip(states, ref)):
        for name, a, c in zip(("h", "vn", "tr"), s.download(), r):
            assert same_bits(a, c), (i, name)


def test_batch_autotune_cache(tmp_path, monkeypatch):
    """A tuned choice is written to OMEGA_TUNE_CACHE and applied from it by
    the next autotune of the same ensemble without timing."""
    cache = tmp_path / "tune.txt"
    monkeypatch.setenv("OMEGA_TUNE_CACHE", str(cache))
    members = make_members(MEMBERS[:3])

    def batch():
        b = swe.Batch("lsrk", 5)
        for mesh, model, ntr, y0, dt, _ in members:
            b.add(model, state_of(mesh, ntr, y0), dt)
        return b

    b1 = batch()
    ms1 = b1.autotune(reps=1)
    choice = (b1.get_launch(), b1.get_persistent())
    assert cache.read_text().count("\n") == 1
    b2 = batch()
    b2.set_launch(64, 1)
    ms2 = b2.autotune(reps=1)
    assert (b2.get_launch(), b2.get_persistent()) == choice
    assert ms2 == pytest.approx(ms1)
    assert cache.read_text().count("\n") == 1


def test_model_autotune_keeps_results():
    """The single-model tuner picks a launch shape; tendencies and steps stay
    the oracle's bits under it."""
    nx, L, ntr, dc = 12, 64, 2, 10e3
    pd = params_dict(dc, ntr)
    om = po.Mesh(nx, nx, dc, L)
    omodel = po.Model(om, **pd)
    gm = swe.PlanarHexagonalMesh(nx, nx, dc, L)
    gmodel = swe.ShallowWaterModel(gm, swe.ShallowWaterParams(**pd))
    y0 = rich_state(om, L, ntr, 3)
    gs = state_of(gm, ntr, y0)
    ms = gmodel.autotune(gs, reps=2)
    assert ms > 0
    print("model autotune", gmodel.get_launch(), ms, flush=True)
    st = swe.RK4Stepper(gmodel)
    st.set_graph(True)
    st.run(0.0, 20.0, 2, gs)
    ref = tuple(a.copy() for a in y0)
    po.Stepper(omodel, "rk4").run(0.0, 20.0, 2, ref)
    for a, c in zip(gs.download(), ref):
        assert same_bits(a, c)


@pytest.mark.parametrize("kind,nstages", [("lsrk", 5), ("rk4", 4)])
@pytest.mark.parametrize("forms", [7, 3])
def test_batch_compiled_term_set(kind, nstages, forms):
    """Pattern-specialised members that share a term set (here del2
    viscosity and tracer del4, the 64-layer members' pair operators get it
    compiled into the ensemble's tendency launch; 16/32-layer members keep
    reading theirs) end with the bits of their solo steps, as with that
    form forbidden."""
    over = dict(visc_del4=0.0, eddy_diff2=0.0)
    members = make_members([(8, 8, 64, 2, 10e3, over, 12.0), (10, 8, 64, 1, 10e3, over, 15.0),
                            (8, 6, 32, 3, 10e3, over, 14.0), (12, 10, 128, 2, 10e3, over, 9.0)])
    for m in members:
        m[1].set_kernel_forms(forms)
        assert m[1].terms == 79
    print("case", kind, forms, flush=True)
    ref = solo(members, kind, nstages, 3)
    b = swe.Batch(kind, nstages)
    states = []
    for mesh, model, ntr, y0, dt, _ in members:
        s = state_of(mesh, ntr, y0)
        b.add(model, s, dt)
        states.append(s)
    b.set_graph(True)
    b.run(3)
    for i, (s, r) in enumerate(zip(states, ref)):
        for name, a, c in zip(("h", "vn", "tr"), s.download(), r):
            assert same_bits(a, c), (i, name)


@pytest.mark.parametrize("rows,drows", [(4, 2), (2, 0), (0, 4)])
def test_model_tuned_launch_from_cache_keeps_results(tmp_path, monkeypatch, rows, drows):
    """A recorded model choice (OMEGA_TUNE_CACHE: rows per block, levels per
    thread, swizzle and the diagnostics launch's own rows) is applied without
    timing; tendencies and steps under it stay the oracle's bits."""
    nx, L, ntr, dc = 12, 64, 2, 10e3
    pd = params_dict(dc, ntr)
    om = po.Mesh(nx, nx, dc, L)
    omodel = po.Model(om, **pd)
    gm = swe.PlanarHexagonalMesh(nx, nx, dc, L)
    gmodel = swe.ShallowWaterModel(gm, swe.ShallowWaterParams(**pd))
    cache = tmp_path / "tune.txt"
    cache.write_text(f"model;{gm.ncells}:{L}:{ntr}:{gmodel.terms}:h {rows} 2 1 0.5 {drows}\n")
    monkeypatch.setenv("OMEGA_TUNE_CACHE", str(cache))
    y0 = rich_state(om, L, ntr, 4)
    gs = state_of(gm, ntr, y0)
    assert gmodel.autotune(gs, reps=1) == pytest.approx(0.5)
    assert gmodel.get_launch()[:3] == (rows, 2, 1)
    print("case", rows, drows, flush=True)
    st = swe.RK4Stepper(gmodel)
    st.set_graph(True)
    st.run(0.0, 20.0, 2, gs)
    ref = tuple(a.copy() for a in y0)
    po.Stepper(omodel, "rk4").run(0.0, 20.0, 2, ref)
    for a, c in zip(gs.download(), ref):
        assert same_bits(a, c)
