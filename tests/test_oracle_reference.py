"""The reference's own self-checking tests, run on the CPU oracle.

Each test restates one of omega_playground's test drivers (test/*.cpp) and
asserts the same convergence rate or conservation bound; they pin the
oracle's mesh, operators, tendencies and time steppers.
"""
import math

import numpy as np
import pytest

from oracle import pyoracle as po
from omega_playground_amd import workloads as wl


def rate(e0, e1):
    return math.log2(e0 / e1)


# test/operator_convergence.cpp:11-224 -------------------------------------
def operator_errors(nx):
    m = po.Mesh(nx, nx)  # dc = 1/nx
    lx, ly = m.period_x, m.period_y
    xc, yc, xe, ye = m.x_cell, m.y_cell, m.x_edge, m.y_edge
    xv, yv = m.x_vertex, m.y_vertex
    ne, ce = np.cos(m.angle_edge), np.sin(m.angle_edge)
    k2x, k2y = 2 * math.pi / lx, 2 * math.pi / ly
    # gradient (cell -> edge)
    f = np.sin(k2x * xc) * np.sin(k2y * yc)
    c0, c1 = m.cells_on_edge[:, 0], m.cells_on_edge[:, 1]
    grad = (f[c1] - f[c0]) / m.dc_edge
    gx = k2x * np.cos(k2x * xe) * np.sin(k2y * ye)
    gy = k2y * np.sin(k2x * xe) * np.cos(k2y * ye)
    e_grad = np.max(np.abs(ne * gx + ce * gy - grad))
    # divergence (edge -> cell)
    vx = np.sin(k2x * xe) * np.cos(k2y * ye)
    vy = np.cos(k2x * xe) * np.sin(k2y * ye)
    vn = ne * vx + ce * vy
    eoc, sgn = m.edges_on_cell, m.edge_sign_on_cell
    div = np.sum(m.dv_edge[eoc] * sgn * vn[eoc], axis=1) / m.area_cell
    ediv = 2 * math.pi * (1 / lx + 1 / ly) * np.cos(k2x * xc) * np.cos(k2y * yc)
    e_div = np.max(np.abs(ediv - div))
    # curl (edge -> vertex)
    eov, sv = m.edges_on_vertex, m.edge_sign_on_vertex
    curl = np.sum(m.dc_edge[eov] * sv * vn[eov], axis=1) / m.area_triangle
    ecurl = 2 * math.pi * (-1 / lx + 1 / ly) * np.sin(k2x * xv) * np.sin(k2y * yv)
    e_curl = np.max(np.abs(ecurl - curl))
    # tangential reconstruction
    eoe, w = m.edges_on_edge[:, :10], m.weights_on_edge[:, :10]
    recon = np.sum(w * vn[eoe], axis=1)
    e_recon = np.max(np.abs((-ce * vx + ne * vy) - recon))
    return e_grad, e_div, e_curl, e_recon


def test_operator_convergence():
    errs = [operator_errors(6 * 2 ** l) for l in range(5)]
    rates = [rate(errs[-2][i], errs[-1][i]) for i in range(4)]
    # operator_convergence.cpp:256-301: grad 2, div 2, curl 1, recon 2 (+-0.05)
    for r, expected in zip(rates, (2, 2, 1, 2)):
        assert abs(r - expected) < 0.05, rates


# test/nonlinear_swe_convergence.cpp:115-212 -------------------------------
def swe_error(n):
    ms = wl.ManufacturedSolution()
    m = po.Mesh(n, n, ms.lx / n, 1)
    model = po.Model(m, f0=ms.f0, grav=ms.grav)
    xc, yc, xe, ye, ae = (np.array(a) for a in (m.x_cell, m.y_cell, m.x_edge, m.y_edge,
                                                 m.angle_edge))

    def forcing(ht, vt, h, vn, t):
        fh, fv = ms.forcing(xc, yc, xe, ye, ae, t)
        ht += fh[:, None]
        vt += fv[:, None]

    model.set_additional_tendency(forcing)
    st = po.Stepper(model, "rk4")
    timeend = 10 * 60 * 60
    dt = 3 * m.dc / 1e3
    ns = math.ceil(timeend / dt)
    dt = timeend / ns
    h = ms.h(xc, yc, 0)[:, None].copy()
    vn = wl.normal_velocity(ae, ms.vx(xe, ye, 0)[:, None], ms.vy(xe, ye, 0)[:, None])
    st.run(0.0, dt, ns, (h, np.ascontiguousarray(vn), np.zeros((0, m.ncells, 1))))
    return math.sqrt(((ms.h(xc, yc, timeend)[:, None] - h) ** 2).sum() / (n * n))


def test_nonlinear_swe_convergence():
    r = rate(swe_error(50), swe_error(100))
    assert abs(r - 2) < 0.05, r


# test/tracer_{advection,diffusion,hyperdiffusion}_convergence.cpp ----------
def tracer_error(kind, nx):
    lx = 1.0 if kind == "adv" else 10.0
    ly = math.sqrt(3) / 2 * lx
    m = po.Mesh(nx, nx, lx / nx, 1)
    kw = dict(disable_h_tendency=True, disable_vn_tendency=True, ntracers=1)
    if kind == "diff":
        kw["eddy_diff2"] = 1.0
    if kind == "hyper":
        kw["eddy_diff4"] = 1.0
    model = po.Model(m, **kw)
    st = po.Stepper(model, "lsrk", 5)
    xc, yc = np.array(m.x_cell), np.array(m.y_cell)
    if kind == "adv":
        T, dt = 1.0, 0.5 * m.dc
        h = np.exp(np.cos(2 * math.pi * xc / lx))
        exact = lambda t: np.sin(2 * math.pi * (xc - t) / lx) * np.sin(2 * math.pi * (yc - t) / ly)
        vx = vy = 1.0
    else:
        T = 2.0
        c = 4 * math.pi ** 2 * (1 / lx ** 2 + 1 / ly ** 2)
        decay = c if kind == "diff" else c * c
        dt = m.dc ** 2 / 4 if kind == "diff" else m.dc ** 4 / 8
        h = np.ones_like(xc)
        exact = lambda t: (np.sin(2 * math.pi * xc / lx) * np.sin(2 * math.pi * yc / ly)
                           * np.exp(-decay * t))
        vx = vy = 0.0
    ns = math.ceil(T / dt)
    dt = T / ns
    ae = np.array(m.angle_edge)
    vn = (np.cos(ae) * vx + np.sin(ae) * vy)[:, None].copy()
    tr = exact(0)[None, :, None].copy()
    st.run(0.0, dt, ns, (h[:, None].copy(), vn, tr))
    return math.sqrt(((exact(T)[None, :, None] - tr) ** 2).sum() / (nx * nx))


@pytest.mark.parametrize("kind,nx", [("adv", 50), ("diff", 16), ("hyper", 16)])
def test_tracer_convergence(kind, nx):
    r = rate(tracer_error(kind, nx), tracer_error(kind, 2 * nx))
    assert abs(r - 2) < 0.05, r


# test/double_vortex.cpp:79-168 ---------------------------------------------
def test_double_vortex_conservation():
    dv = wl.DoubleVortexTest()
    nx = 25
    dc = dv.lx / nx
    ny = math.ceil(dv.ly / (dc * math.sqrt(3) / 2))
    m = po.Mesh(nx, ny, dc, 1)
    model = po.Model(m, f0=dv.coriolis, grav=dv.g)
    st = po.Stepper(model, "lsrk", 5)
    h, vn, tr = dv.state(m, 1, 0)
    T = 200 * 1000
    dt = 0.1 * m.dc / math.sqrt(dv.g * dv.h0)
    ns = math.ceil(T / dt)
    dt = T / ns
    m0, c0, e0 = model.mass(h), model.circulation(vn), model.energy(h, vn)
    st.run(0.0, dt, ns, (h, vn, tr))
    m1, c1, e1 = model.mass(h), model.circulation(vn), model.energy(h, vn)
    assert abs((m1 - m0) / m0) < 5e-14
    assert abs((c1 - c0) / c0) < 5e-14
    assert abs((e1 - e0) / e0) < 1e-10
