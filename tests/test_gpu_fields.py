"""GPU tests of the per-field tendency calls on device arrays, the auxiliary
state, the opt-in bottom drag, and two hazards of captured steps: a model
setting changed after a graph was captured, and tracer rows of a state
beyond the model's tracer count.

Oracle: oracle/swe_oracle.cpp (compute_{h,vn,tr}_tendency after aux_compute,
shallow_water_model.cpp:43-311; the drag term restates the comment at
shallow_water_hierarchical.cpp:479-485).  Bar: bit-for-bit.
"""
import numpy as np
import pytest

from oracle import pyoracle as po
from omega_playground_amd import swe

from helpers import params_dict, rich_state, same_bits

pytestmark = pytest.mark.gpu


def pair(nx, L, ntr, dc=10e3, **extra):
    pd = params_dict(dc, ntr, True)
    pd.update(extra)
    om = po.Mesh(nx, nx, dc, L)
    omodel = po.Model(om, **pd)
    gm = swe.PlanarHexagonalMesh(nx, nx, dc, L)
    gmodel = swe.ShallowWaterModel(gm, swe.ShallowWaterParams(**pd))
    return om, omodel, gm, gmodel, rich_state(om, L, ntr, seed=nx + L)


def upload(gm, ntr, arrays):
    s = swe.ShallowWaterState(gm, ntr)
    s.upload(*arrays)
    return s


# --- compute_{h,vn,tr}_tendency(Real2d/3d tend, ..., AddMode) on device arrays
@pytest.mark.parametrize("L", [6, 64])
def test_device_field_tendencies_match_oracle(L):
    om, omodel, gm, gmodel, state = pair(10, L, 3)
    print("device fields L", L, flush=True)
    gs = upload(gm, 3, state)
    rng = np.random.default_rng(L)
    junk = tuple(rng.standard_normal(a.shape) for a in state)
    part = upload(gm, 3, junk)
    h_p, vn_p, tr_p = gs.device_ptrs()
    th, tv, tt = part.device_ptrs()
    gmodel.compute_aux(gs)
    gmodel.compute_h_tendency_dev(th, h_p, vn_p)
    gmodel.compute_vn_tendency_dev(tv, h_p, vn_p)
    gmodel.compute_tr_tendency_dev(tt, tr_p, vn_p)
    h, vn, tr = state
    omodel.aux_compute(h, vn, tr)
    ref = [j.copy() for j in junk]
    omodel.compute_field_tendency(0, ref[0], h, vn)
    omodel.compute_field_tendency(1, ref[1], h, vn)
    omodel.compute_field_tendency(2, ref[2], tr, vn)
    for f, (a, b) in enumerate(zip(part.download(), ref)):
        assert same_bits(a, b), f
    # increment mode
    gmodel.compute_vn_tendency_dev(tv, h_p, vn_p, swe.ADD_INCREMENT)
    omodel.compute_field_tendency(1, ref[1], h, vn, increment=True)
    assert same_bits(part.download()[1], ref[1])


def test_device_field_tendencies_read_the_auxiliary_state():
    """As in the reference, compute_h_tendency takes its thickness flux from
    the auxiliary state, not from its h argument (thickness_hadv_cell.hpp:
    48-72 reads flux_h_edge), and compute_tr_tendency its normalised tracers
    (tr_cell is not read)."""
    om, omodel, gm, gmodel, state = pair(8, 4, 2)
    other = rich_state(om, 4, 2, seed=99)
    gs = upload(gm, 2, state)
    go = upload(gm, 2, other)
    out = swe.ShallowWaterState(gm, 2)
    gmodel.compute_aux(gs)
    h_p, vn_p, tr_p = gs.device_ptrs()
    oh, ovn, otr = go.device_ptrs()
    th, tv, tt = out.device_ptrs()
    gmodel.compute_h_tendency_dev(th, oh, vn_p)       # h argument unused
    gmodel.compute_tr_tendency_dev(tt, otr, vn_p)     # tr argument unused
    h, vn, tr = state
    omodel.aux_compute(h, vn, tr)
    rh, rt = np.zeros_like(h), np.zeros_like(tr)
    omodel.compute_field_tendency(0, rh, other[0], vn)
    omodel.compute_field_tendency(2, rt, other[2], vn)
    got = out.download()
    assert same_bits(got[0], rh)
    assert same_bits(got[2], rt)
    # the vn tendency reads h for its SSH gradient and the thickness flux:
    # another array than the auxiliary state's is refused
    with pytest.raises(swe.OmegaError):
        gmodel.compute_vn_tendency_dev(tv, oh, vn_p)


def test_device_field_tendency_needs_an_auxiliary_state():
    om, omodel, gm, gmodel, state = pair(6, 2, 1)
    gs = upload(gm, 1, state)
    out = swe.ShallowWaterState(gm, 1)
    h_p, vn_p, _ = gs.device_ptrs()
    with pytest.raises(swe.OmegaError):
        gmodel.compute_h_tendency_dev(out.device_ptrs()[0], h_p, vn_p)


@pytest.mark.parametrize("L", [6, 64])
def test_auxiliary_arrays_match_oracle(L):
    """m_aux_state fields after compute_aux (shallow_water_auxstate.cpp:12-50)."""
    om, omodel, gm, gmodel, state = pair(10, L, 2)
    gs = upload(gm, 2, state)
    gmodel.compute_aux(gs)
    omodel.aux_compute(*state)
    assert same_bits(gmodel.aux_array("rvort_vertex"), omodel.rvort_vertex)
    assert same_bits(gmodel.aux_array("ke_cell"), omodel.ke_cell)
    assert same_bits(gmodel.aux_array("vel_div_cell"), omodel.vel_div_cell)
    # normalised PV at vertices is f * (1 / thickness)
    inv = gmodel.aux_array("inv_thick_vertex")
    assert same_bits(inv * omodel.params.f0, omodel.norm_pvort_vertex)


# --- opt-in bottom drag ------------------------------------------------------
@pytest.mark.parametrize("L", [5, 64])
def test_bottom_drag_tendency_matches_oracle(L):
    om, omodel, gm, gmodel, state = pair(10, L, 2, drag_coeff=3e-3)
    # off by default even with drag_coeff set, as in the reference
    assert not gmodel.terms & swe.DRAG
    t_off = swe.ShallowWaterState(gm, 2)
    gs = upload(gm, 2, state)
    gmodel.compute_tendency(t_off, gs, 0.0)
    ref_off = omodel.new_state()
    omodel.compute_tendency(ref_off, state, 0.0)
    for a, b in zip(t_off.download(), ref_off):
        assert same_bits(a, b)
    gmodel.terms = gmodel.terms | swe.DRAG
    omodel.terms = omodel.terms | po.DRAG
    assert gmodel.kernel_variant() == 1  # general kernels carry the term
    t_on = swe.ShallowWaterState(gm, 2)
    gmodel.compute_tendency(t_on, gs, 0.0)
    ref_on = omodel.new_state()
    omodel.compute_tendency(ref_on, state, 0.0)
    got = t_on.download()
    for a, b in zip(got, ref_on):
        assert same_bits(a, b)
    # only the deepest level changes
    dv = got[1] - t_off.download()[1]
    assert np.all(dv[:, :L - 1] == 0) and np.any(dv[:, L - 1] != 0)


@pytest.mark.parametrize("kind", ["rk4", "lsrk"])
@pytest.mark.parametrize("graph", [False, True])
def test_bottom_drag_steps_match_oracle(kind, graph):
    om, omodel, gm, gmodel, state = pair(8, 4, 1, drag_coeff=2e-3)
    gmodel.terms = gmodel.terms | swe.DRAG
    omodel.terms = omodel.terms | po.DRAG
    ref = tuple(a.copy() for a in state)
    po.Stepper(omodel, kind, 5).run(0.0, 30.0, 3, ref)
    gs = upload(gm, 1, state)
    st = swe.RK4Stepper(gmodel) if kind == "rk4" else swe.LSRKStepper(gmodel, 5)
    st.set_graph(graph)
    st.run(0.0, 30.0, 3, gs)
    for a, b in zip(gs.download(), ref):
        assert same_bits(a, b)


def test_bottom_drag_has_no_adjoint():
    om, omodel, gm, gmodel, state = pair(6, 2, 1, drag_coeff=1e-3)
    gmodel.terms = gmodel.terms | swe.DRAG
    gs = upload(gm, 1, state)
    cot = upload(gm, 1, state)
    out = swe.ShallowWaterState(gm, 1)
    with pytest.raises(swe.OmegaError):
        swe.tendency_vjp(gmodel, gs, cot, out)


# --- captured steps after a model setting changes ---------------------------
@pytest.mark.parametrize("kind", ["rk4", "lsrk"])
def test_graph_replays_follow_model_settings(kind):
    """A graph captured before set_f_vertex / set_terms is not replayed after
    them: each step matches the oracle with the settings of that step."""
    om, omodel, gm, gmodel, state = pair(10, 64, 2)
    gs = upload(gm, 2, state)
    ref = tuple(a.copy() for a in state)
    st = swe.RK4Stepper(gmodel) if kind == "rk4" else swe.LSRKStepper(gmodel, 5)
    ost = po.Stepper(omodel, kind, 5)
    st.set_graph(True)
    dt = 20.0
    st.run(0.0, dt, 2, gs)
    ost.run(0.0, dt, 2, ref)
    # a non-uniform Coriolis parameter (the pair kernels then read it per cell)
    xv = np.asarray(om.x_vertex)
    f = 1e-4 * (1 + 0.5 * np.sin(2 * np.pi * xv / om.period_x))
    gmodel.set_f_vertex(f)
    omodel.f_vertex[:] = f
    st.run(2 * dt, dt, 2, gs)
    ost.run(2 * dt, dt, 2, ref)
    for a, b in zip(gs.download(), ref):
        assert same_bits(a, b)
    # switch the KE gradient off
    gmodel.terms = gmodel.terms & ~swe.KE
    omodel.terms = omodel.terms & ~po.KE
    st.run(4 * dt, dt, 2, gs)
    ost.run(4 * dt, dt, 2, ref)
    for a, b in zip(gs.download(), ref):
        assert same_bits(a, b)


def test_batch_tables_follow_model_settings():
    om, omodel, gm, gmodel, state = pair(10, 64, 2)
    gs = upload(gm, 2, state)
    ref = tuple(a.copy() for a in state)
    ost = po.Stepper(omodel, "lsrk", 5)
    b = swe.Batch("lsrk", 5)
    b.add(gmodel, gs, 20.0)
    b.set_graph(True)
    b.run(2)
    ost.run(0.0, 20.0, 2, ref)
    gmodel.terms = gmodel.terms & ~swe.SSH
    omodel.terms = omodel.terms & ~po.SSH
    b.run(3)
    ost.run(0.0, 20.0, 3, ref)
    for a, c in zip(gs.download(), ref):
        assert same_bits(a, c)


# --- state tracers beyond the model's tracer count --------------------------
@pytest.mark.parametrize("graph", [False, True])
def test_lsrk_keeps_extra_state_tracers(graph):
    """A 3-tracer state stepped by a 1-tracer model: the two extra tracer
    rows are not stepped and stay as they were (lsrk.cpp:47-87 loops over the
    model's tracers only), after odd and even numbers of five-stage steps,
    alone and in an ensemble."""
    om, omodel, gm, gmodel, st3 = pair(8, 64, 1)
    h, vn, _ = st3
    tr = rich_state(om, 64, 3, seed=5)[2]
    gs = swe.ShallowWaterState(gm, 3)
    gs.upload(h, vn, tr)
    st = swe.LSRKStepper(gmodel, 5)
    st.set_graph(graph)
    ref = (h.copy(), vn.copy(), np.ascontiguousarray(tr[:1]).copy())
    ost = po.Stepper(omodel, "lsrk", 5)
    for n in (1, 2):
        st.run(0.0, 20.0, n, gs)
        ost.run(0.0, 20.0, n, ref)
        got = gs.download()
        assert same_bits(got[2][1:], tr[1:])
        assert same_bits(got[2][:1], ref[2])
        assert same_bits(got[0], ref[0])
    # ensemble
    gb = swe.ShallowWaterState(gm, 3)
    gb.upload(h, vn, tr)
    b = swe.Batch("lsrk", 5)
    b.add(gmodel, gb, 20.0)
    b.set_graph(graph)
    b.run(3)
    assert same_bits(gb.download()[2][1:], tr[1:])


# --- the numbering hooks see: device geometry and the permutations ---------
def test_device_geometry_is_in_device_numbering():
    """The arrays an additional-tendency hook receives are in device
    numbering; the mesh's device geometry (what a hook's own kernel indexes)
    is the host geometry permuted by the same map, and is available before
    any model or state exists."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    gm = swe.PlanarHexagonalMesh(12, 10, 10e3, 4)
    for name, kind in (("x_cell", 0), ("y_cell", 0), ("area_cell", 0), ("x_edge", 1),
                       ("y_edge", 1), ("angle_edge", 1), ("x_vertex", 2), ("y_vertex", 2)):
        ptr, n = gm.device_array(name)
        assert ptr and n == (gm.ncells, gm.nedges, gm.nvertices)[kind], name
        out = np.empty(n)
        assert hip.hipMemcpy(C.c_void_p(out.ctypes.data), C.c_void_p(ptr), C.c_size_t(8 * n), 2) == 0
        host = np.asarray(getattr(gm, name), dtype=np.float64)
        assert same_bits(out, host[gm.permutation(kind)]), name
