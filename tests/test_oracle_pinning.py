"""Pins the CPU oracle with facts the mathematics fixes independently of the
code: closed forms on the regular hexagonal mesh, discrete conservation
identities of the TRiSK scheme, manufactured-solution convergence for the
terms no reference test exercises (momentum del2/del4, tracer del2 with
variable thickness), and the exact amplification polynomials of the time
steppers.  Each check is chosen so that a plausible slip (a dropped term, a
wrong sign or index, swapped operands or coefficients) fails it;
test_oracle_mutants.py demonstrates that on planted mistakes.
"""
import math

import numpy as np
import pytest

from oracle import pyoracle as po

TERMS_ALL = 127


def state_zero(m, L, ntr=0):
    return (np.zeros((m.ncells, L)), np.zeros((m.nedges, L)), np.zeros((ntr, m.ncells, L)))


def tendency(model, state):
    t = model.new_state()
    model.compute_tendency(t, state, 0.0)
    return t


def uniform_vn(m, L, u0, v0):
    ae = np.asarray(m.angle_edge)
    return np.ascontiguousarray(np.repeat((np.cos(ae) * u0 + np.sin(ae) * v0)[:, None], L, 1))


# --- closed forms for uniform flow ----------------------------------------
def test_uniform_flow_closed_forms():
    m = po.Mesh(14, 10, 3.0, 2)
    u0, v0, H, f = 0.7, -0.3, 5.0, 2.0
    model = po.Model(m, f0=f, grav=9.81)
    h = np.full((m.ncells, 2), H)
    vn = uniform_vn(m, 2, u0, v0)
    model.terms = po.PV
    t = tendency(model, (h, vn, np.zeros((0, m.ncells, 2))))
    ae = np.asarray(m.angle_edge)
    ut = (-np.sin(ae) * u0 + np.cos(ae) * v0)[:, None]
    # TRiSK PV flux of a uniform flow with uniform h and f: f times the
    # tangential velocity, exactly
    assert np.max(np.abs(t[1] - f * ut)) < 1e-14
    assert np.max(np.abs(t[0])) < 1e-14  # no thickness flux divergence
    model.aux_compute(h, vn, np.zeros((0, m.ncells, 2)))
    assert np.allclose(model.ke_cell, 0.5 * (u0 * u0 + v0 * v0), rtol=1e-14, atol=0)
    assert np.max(np.abs(model.vel_div_cell)) < 1e-14
    assert np.max(np.abs(model.rvort_vertex)) < 1e-14
    assert np.allclose(model.norm_pvort_vertex, f / H, rtol=1e-14)
    assert np.allclose(model.mean_h_edge, H, rtol=0, atol=0)


def test_linear_thickness_closed_forms():
    """h linear in x, y: the vertex thickness is the value at the vertex,
    the SSH gradient is -g grad(h).n, and h_tend = -u.grad(h) for a uniform
    flow, all exactly (midpoint and centroid rules integrate linear fields)."""
    m = po.Mesh(12, 12, 2.0, 1)
    a, b, c0, g = 0.3, -0.2, 50.0, 9.81
    xc, yc = np.asarray(m.x_cell), np.asarray(m.y_cell)
    # keep h linear across the periodic seam by using a mesh far larger than
    # the stencil and testing away from the seam
    h = (c0 + a * xc + b * yc)[:, None].copy()
    u0, v0 = 0.4, 0.9
    vn = uniform_vn(m, 1, u0, v0)
    model = po.Model(m, f0=0.0, grav=g)
    model.terms = po.SSH
    t = tendency(model, (h, vn, np.zeros((0, m.ncells, 1))))
    ae = np.asarray(m.angle_edge)
    lx, ly = m.period_x, m.period_y
    xe, ye = np.asarray(m.x_edge), np.asarray(m.y_edge)
    inner_e = (xe > 3 * m.dc) & (xe < lx - 3 * m.dc) & (ye > 3 * m.dc) & (ye < ly - 3 * m.dc)
    assert np.max(np.abs(t[1][inner_e, 0] + g * (np.cos(ae) * a + np.sin(ae) * b)[inner_e])) < 1e-11
    inner_c = (xc > 3 * m.dc) & (xc < lx - 3 * m.dc) & (yc > 3 * m.dc) & (yc < ly - 3 * m.dc)
    assert np.max(np.abs(t[0][inner_c, 0] + (u0 * a + v0 * b))) < 1e-12
    model.aux_compute(h, vn, np.zeros((0, m.ncells, 1)))
    xv, yv = np.asarray(m.x_vertex), np.asarray(m.y_vertex)
    inner_v = (xv > 3 * m.dc) & (xv < lx - 3 * m.dc) & (yv > 3 * m.dc) & (yv < ly - 3 * m.dc)
    rv_norm = model.norm_rvort_vertex[:, 0]
    # rvort = 0 for uniform flow, so read the thickness through npv with f=1
    model2 = po.Model(m, f0=1.0, grav=g)
    model2.aux_compute(h, vn, np.zeros((0, m.ncells, 1)))
    thick = 1.0 / model2.norm_pvort_vertex[:, 0]
    assert np.max(np.abs(thick[inner_v] - (c0 + a * xv + b * yv)[inner_v])) < 1e-11
    assert np.max(np.abs(rv_norm)) < 1e-15


# --- discrete conservation identities -------------------------------------
def test_pv_flux_is_energy_neutral():
    """sum_e dc dv F_e qt_e = 0 for any state (the antisymmetry of the TRiSK
    weights); a wrong weight sign, slot or edges-on-edge entry breaks it."""
    m = po.Mesh(16, 12, 1.0, 3)
    rng = np.random.default_rng(0)
    model = po.Model(m, f0=1.3)
    model.terms = po.PV
    h = 10 + rng.random((m.ncells, 3))
    vn = rng.standard_normal((m.nedges, 3))
    t = tendency(model, (h, vn, np.zeros((0, m.ncells, 3))))
    F = model.flux_h_edge * vn
    area = (np.asarray(m.dc_edge) * np.asarray(m.dv_edge))[:, None]
    assert abs(np.sum(area * F * t[1])) < 1e-13 * np.sum(np.abs(area * F * t[1]))


def test_curl_of_gradient_and_mass_flux_conservation():
    m = po.Mesh(10, 8, 1.0, 2)
    rng = np.random.default_rng(1)
    model = po.Model(m, f0=0.0)
    phi = rng.standard_normal((m.ncells, 2))
    c0, c1 = m.cells_on_edge[:, 0], m.cells_on_edge[:, 1]
    vn = np.ascontiguousarray((phi[c1] - phi[c0]) / np.asarray(m.dc_edge)[:, None])
    model.aux_compute(np.ones((m.ncells, 2)), vn, np.zeros((0, m.ncells, 2)))
    assert np.max(np.abs(model.rvort_vertex)) < 1e-12
    # total mass tendency vanishes (flux form)
    h = 5 + rng.random((m.ncells, 2))
    vn = rng.standard_normal((m.nedges, 2))
    t = tendency(model, (h, vn, np.zeros((0, m.ncells, 2))))
    A = np.asarray(m.area_cell)[:, None]
    assert abs(np.sum(A * t[0])) < 1e-13 * np.sum(np.abs(A * t[0]))


def test_tracer_flux_consistency():
    """tr = c*h advects like h: the tracer advection tendency is c times the
    thickness tendency (same fluxes, norm_tr = c)."""
    m = po.Mesh(12, 10, 1.0, 2)
    rng = np.random.default_rng(2)
    model = po.Model(m, f0=1e-4, ntracers=1)
    h = 5 + rng.random((m.ncells, 2))
    vn = rng.standard_normal((m.nedges, 2))
    tr = (2.5 * h)[None].copy()
    t = tendency(model, (h, vn, tr))
    assert np.max(np.abs(t[2][0] - 2.5 * t[0])) < 1e-14 * np.max(np.abs(t[2][0]))


# --- manufactured solutions for terms no reference test covers -------------
def momentum_mixing(m, vn, term, nu=1.0):
    kw = dict(visc_del2=nu) if term == po.VDEL2 else dict(visc_del4=nu)
    model = po.Model(m, f0=0.0, **kw)
    model.terms = term
    t = tendency(model, (np.ones((m.ncells, 1)), vn, np.zeros((0, m.ncells, 1))))
    return t[1][:, 0], model


def test_momentum_del2_summation_by_parts():
    """The del2 term is nu*(grad(div) - curl^T(curl)) of the mimetic C-grid:
    sum_e dc dv u L(u) = -(sum_c A div^2 + sum_v A_tri vort^2) (Green's
    identity, which fixes the sign and metric of both halves) and L is
    self-adjoint in the dc*dv inner product.  (On the hexagonal C-grid the
    vertex vorticity is only first-order accurate, so the curl half does
    not converge pointwise to the vector Laplacian; the reference operator
    is pinned by its algebra instead.)"""
    m = po.Mesh(12, 10, 1.0, 1)
    rng = np.random.default_rng(0)
    wgt = np.asarray(m.dc_edge) * np.asarray(m.dv_edge)
    u = rng.standard_normal((m.nedges, 1))
    w = rng.standard_normal((m.nedges, 1))
    Lu, mod = momentum_mixing(m, u, po.VDEL2)
    Lw, _ = momentum_mixing(m, w, po.VDEL2)
    d, z = mod.vel_div_cell[:, 0], mod.rvort_vertex[:, 0]
    energy = np.sum(np.asarray(m.area_cell) * d * d) + np.sum(np.asarray(m.area_triangle) * z * z)
    assert abs(np.sum(wgt * u[:, 0] * Lu) + energy) < 1e-13 * energy
    a, b = np.sum(wgt * w[:, 0] * Lu), np.sum(wgt * u[:, 0] * Lw)
    assert abs(a - b) < 1e-13 * abs(a)


def test_momentum_del2_fourier_closed_form():
    """For the discrete gradient of a Fourier mode the curl vanishes and
    del2 = nu * lambda * vn with the 7-point hexagonal Laplacian symbol
    lambda = 2/(3 dc^2) sum_j (cos(k.d_j) - 1)."""
    m = po.Mesh(12, 12, 1.0, 1)
    xc, yc = np.asarray(m.x_cell), np.asarray(m.y_cell)
    kx, ky = 2 * math.pi / m.period_x, 2 * math.pi * 2 / m.period_y
    phi = np.cos(kx * xc + ky * yc)
    c0, c1 = m.cells_on_edge[:, 0], m.cells_on_edge[:, 1]
    vn = ((phi[c1] - phi[c0]) / np.asarray(m.dc_edge))[:, None].copy()
    Lu, mod = momentum_mixing(m, vn, po.VDEL2, nu=0.7)
    dc = m.dc
    lam = 2 / (3 * dc * dc) * sum(math.cos(kx * dc * math.cos(a) + ky * dc * math.sin(a)) - 1
                                  for a in np.arange(6) * math.pi / 3)
    assert np.max(np.abs(Lu - 0.7 * lam * vn[:, 0])) < 1e-13 * np.max(np.abs(lam * vn))
    assert np.max(np.abs(mod.rvort_vertex)) < 1e-14


def test_momentum_del4_is_minus_del2_squared():
    """del4 = -nu4 L(L(u)) with L the del2 operator (its max(dv, dc/4)
    guard is inert on the regular mesh): sum_e dc dv u del4 = -nu4 sum dc dv
    (L u)^2 with L u the pre-pass vel_del2_edge, and the pre-pass equals the
    del2 operator applied to u."""
    m = po.Mesh(12, 10, 1.0, 1)
    rng = np.random.default_rng(4)
    wgt = np.asarray(m.dc_edge) * np.asarray(m.dv_edge)
    u = rng.standard_normal((m.nedges, 1))
    t4, mod = momentum_mixing(m, u, po.VDEL4, nu=0.3)
    d2 = mod.vel_del2_edge[:, 0]
    L1, _ = momentum_mixing(m, u, po.VDEL2, nu=1.0)
    assert np.max(np.abs(d2 - L1)) < 1e-13 * np.max(np.abs(L1))
    lhs, rhs = np.sum(wgt * u[:, 0] * t4), -0.3 * np.sum(wgt * d2 * d2)
    assert abs(lhs - rhs) < 1e-13 * abs(rhs)


def rate(a, b):
    return math.log2(a / b)


def tracer_del2_error(nx):
    """div(h grad(phi)) with variable h, through both the del2 tendency and
    the del4 pre-pass (tracer_del2_cell)."""
    m = po.Mesh(nx, nx)
    lx, ly = m.period_x, m.period_y
    kx, ky = 2 * math.pi / lx, 2 * math.pi / ly
    xc, yc = np.asarray(m.x_cell), np.asarray(m.y_cell)
    h = 2 + np.sin(kx * xc) * np.cos(ky * yc)
    phi = np.cos(kx * xc + 2 * ky * yc)
    hx, hy = kx * np.cos(kx * xc) * np.cos(ky * yc), -ky * np.sin(kx * xc) * np.sin(ky * yc)
    px, py = -kx * np.sin(kx * xc + 2 * ky * yc), -2 * ky * np.sin(kx * xc + 2 * ky * yc)
    lap = -(kx * kx + 4 * ky * ky) * phi
    exact = hx * px + hy * py + h * lap
    model = po.Model(m, f0=0.0, ntracers=1, eddy_diff2=1.0, eddy_diff4=1.0)
    model.terms = po.TDEL2
    st = (h[:, None].copy(), np.zeros((m.nedges, 1)), (h * phi)[None, :, None].copy())
    t = tendency(model, st)
    e1 = np.max(np.abs(t[2][0, :, 0] - exact)) / np.max(np.abs(exact))
    model.terms = po.TDEL4
    tendency(model, st)
    e2 = np.max(np.abs(model.tracer_del2_cell[:, 0] - exact)) / np.max(np.abs(exact))
    return e1, e2


def test_tracer_del2_variable_thickness_manufactured():
    a = tracer_del2_error(16)
    b = tracer_del2_error(32)
    for i in range(2):
        assert b[i] < 0.02 and rate(a[i], b[i]) > 1.8


# --- time steppers: exact amplification polynomials ------------------------
def eigen_setup(kind, nstages):
    m = po.Mesh(16, 16, 1.0, 1)
    model = po.Model(m, ntracers=1, eddy_diff2=1.0, disable_h_tendency=True,
                     disable_vn_tendency=True)
    xc, yc = np.asarray(m.x_cell), np.asarray(m.y_cell)
    phi = np.cos(2 * math.pi * (2 * xc / m.period_x + 1 * yc / m.period_y))
    st = (np.ones((m.ncells, 1)), np.zeros((m.nedges, 1)), phi[None, :, None].copy())
    t = tendency(model, st)
    i = int(np.argmax(np.abs(phi)))
    lam = t[2][0, i, 0] / phi[i]
    return m, model, phi, lam, i


def amplification(kind, nstages, z):
    m, model, phi, lam, i = eigen_setup(kind, nstages)
    dt = z / lam
    st = (np.ones((m.ncells, 1)), np.zeros((m.nedges, 1)), phi[None, :, None].copy())
    po.Stepper(model, kind, nstages).run(0.0, dt, 1, st)
    return st[2][0, i, 0] / phi[i]


@pytest.mark.parametrize("z", [-0.1, -0.7, -1.9])
def test_rk4_amplification_polynomial(z):
    r = amplification("rk4", 4, z)
    assert abs(r - (1 + z + z ** 2 / 2 + z ** 3 / 6 + z ** 4 / 24)) < 1e-13


def test_euler_amplification():
    z = -0.3
    assert abs(amplification("lsrk", 1, z) - (1 + z)) < 1e-14


def test_lsrk5_fourth_order():
    """Carpenter-Kennedy (5,4): R(z) = e^z + O(z^5) with the fifth-order
    coefficient of the published scheme, i.e. the error ratio between z
    and z/2 is 2^5."""
    e = [abs(amplification("lsrk", 5, z) - math.exp(z)) for z in (-0.4, -0.2)]
    assert 28 < e[0] / e[1] < 36
    # stable well beyond RK4's real-axis limit
    assert abs(amplification("lsrk", 5, -3.5)) < 1


# --- model bookkeeping ------------------------------------------------------
def test_increment_mode_adds_to_tendency():
    m = po.Mesh(10, 8, 1.0, 2)
    rng = np.random.default_rng(5)
    model = po.Model(m, f0=1e-4, ntracers=2, visc_del2=1e-2, eddy_diff4=1e-3)
    st = (5 + rng.random((m.ncells, 2)), rng.standard_normal((m.nedges, 2)),
          rng.random((2, m.ncells, 2)) + 1)
    base = tendency(model, st)
    prev = tuple(rng.standard_normal(a.shape) for a in base)
    inc = tuple(a.copy() for a in prev)
    model.compute_tendency(inc, st, 0.0, increment=True)
    # h adds one accumulated sum: exactly prev + replace value; tracers and
    # vn add their terms one by one
    assert np.array_equal(inc[0], prev[0] + base[0])
    for i in (1, 2):
        assert np.max(np.abs(inc[i] - (prev[i] + base[i]))) < 1e-14 * np.max(np.abs(inc[i]))


def test_diagnostics_closed_forms():
    m = po.Mesh(8, 6, 2.0, 3)
    g, f, H = 9.81, 1e-4, 7.0
    model = po.Model(m, f0=f, grav=g)
    h = np.full((m.ncells, 3), H)
    vn = np.zeros((m.nedges, 3))
    A = np.sum(m.area_cell)
    At = np.sum(m.area_triangle)
    assert math.isclose(model.mass(h), 3 * H * A, rel_tol=1e-14)
    assert math.isclose(model.energy(h, vn), 3 * g * H * H / 2 * A, rel_tol=1e-14)
    assert math.isclose(model.circulation(vn), 3 * f * At, rel_tol=1e-14)


def test_mesh_invariants():
    m = po.Mesh(9, 8, 1.0, 1)
    eoc, coe, voe = m.edges_on_cell, m.cells_on_edge, m.vertices_on_edge
    for c in range(m.ncells):
        for j in range(6):
            e = eoc[c, j]
            assert c in coe[e]
            assert c in m.cells_on_vertex[m.vertices_on_cell[c, j]]
    # every edge appears in exactly two cells with opposite signs
    s = np.zeros(m.nedges)
    np.add.at(s, eoc.ravel(), m.edge_sign_on_cell.ravel())
    assert np.all(s == 0)
    sv = np.zeros(m.nedges)
    np.add.at(sv, m.edges_on_vertex.ravel(), m.edge_sign_on_vertex.ravel())
    assert np.all(sv == 0)
    # closed-form geometry of the regular hexagon
    dc = m.dc
    assert np.allclose(m.area_cell, dc * dc * math.sqrt(3) / 2, rtol=1e-15)
    assert np.allclose(m.dv_edge, dc / math.sqrt(3), rtol=1e-15)
    assert np.allclose(m.area_triangle, 3 * m.kiteareas_on_vertex[:, 0], rtol=1e-15)
    # weights: each edge's five-weight blocks are antisymmetric under the
    # edge pairing (w_ee' = -w_e'e)
    W = {}
    for e in range(m.nedges):
        for j in range(10):
            W[(e, m.edges_on_edge[e, j])] = W.get((e, m.edges_on_edge[e, j]), 0) + m.weights_on_edge[e, j]
    for (e, f), w in W.items():
        assert abs(w + W.get((f, e), 0.0)) < 1e-15


# --- opt-in bottom drag (shallow_water_hierarchical.cpp:479-485) ------------
def test_bottom_drag_uniform_flow_decay():
    """A uniform flow of speed U in every layer with f = 0 and uniform h has
    no vorticity, divergence, or KE and SSH gradients, and KE = U^2/2 in
    every cell.  With the drag term on, the deepest layer's tendency is then
    -C_d |U| u_e / H on every edge, exactly to rounding, and the other layers
    get none; stepping the flow follows dU/dt = -C_d U^2 / H, whose solution
    U(t) = U0 / (1 + C_d U0 t / H) RK4 matches to its truncation error.  The
    term is off by default (the reference applies no drag)."""
    L = 3
    m = po.Mesh(12, 10, 1000.0, L)
    U0, ang, H, cd = 0.8, 0.4, 50.0, 2e-3
    vn = uniform_vn(m, L, U0 * math.cos(ang), U0 * math.sin(ang))
    h = np.full((m.ncells, L), H)
    trz = np.zeros((0, m.ncells, L))
    model = po.Model(m, f0=0.0, grav=9.81, drag_coeff=cd)
    t_off = tendency(model, (h, vn, trz))
    assert np.max(np.abs(t_off[1])) < 1e-15
    model.terms = model.terms | po.DRAG
    t = tendency(model, (h, vn, trz))
    expect = -cd * U0 * vn[:, L - 1] / H
    assert np.max(np.abs(t[1][:, L - 1] - expect)) <= 1e-12 * np.max(np.abs(expect))
    assert np.max(np.abs(t[1][:, :L - 1])) < 1e-15
    assert np.max(np.abs(t[0])) < 1e-15
    dt, n = 100.0, 50
    st = (h.copy(), vn.copy(), trz.copy())
    po.Stepper(model, "rk4").run(0.0, dt, n, st)
    U = U0 / (1 + cd * U0 * n * dt / H)
    assert np.max(np.abs(st[1][:, L - 1] - vn[:, L - 1] * (U / U0))) <= 1e-9 * U0
    assert np.max(np.abs(st[1][:, :L - 1] - vn[:, :L - 1])) < 1e-13


# --- per-field entry points (shallow_water_model.cpp:43-165) -----------------
def test_field_tendencies_are_the_fused_tendency_blocks():
    """compute_h/vn/tr_tendency after the aux pass give the three blocks of
    compute_tendency bit for bit (the reference's compute_tendency is those
    three calls after compute_auxiliary_variables), in replace and in
    increment mode; a field tendency that skipped a term or read the wrong
    auxiliary array would differ."""
    L, ntr = 4, 2
    m = po.Mesh(10, 8, 1e5, L)
    model = po.Model(m, f0=1e-4, grav=9.81, visc_del2=3.0, visc_del4=1e9, eddy_diff2=2.0,
                     eddy_diff4=1e8, ntracers=ntr)
    rng = np.random.default_rng(7)
    h = 100.0 + rng.random((m.ncells, L))
    vn = rng.standard_normal((m.nedges, L))
    tr = rng.random((ntr, m.ncells, L))
    fused = tendency(model, (h, vn, tr))
    model.aux_compute(h, vn, tr)
    parts = [np.full_like(fused[0], 3.0), np.full_like(fused[1], 3.0), np.full_like(fused[2], 3.0)]
    model.compute_field_tendency(0, parts[0], h, vn)
    model.compute_field_tendency(1, parts[1], h, vn)
    model.compute_field_tendency(2, parts[2], tr, vn)
    for a, b in zip(parts, fused):
        assert np.array_equal(a, b)
        assert np.max(np.abs(b)) > 0
    # increment mode adds term by term onto what is there, as compute_tendency
    # in increment mode does
    inc = [np.full_like(p, 0.5) for p in parts]
    whole = [np.full_like(p, 0.5) for p in parts]
    model.compute_tendency(whole, (h, vn, tr), 0.0, increment=True)
    model.aux_compute(h, vn, tr)
    for f, a in enumerate((h, h, tr)):
        model.compute_field_tendency(f, inc[f], a, vn, increment=True)
        assert np.array_equal(inc[f], whole[f])
        assert np.allclose(inc[f], 0.5 + fused[f], rtol=0, atol=1e-14)
