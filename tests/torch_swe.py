"""Differentiable float64 restatement of the shallow-water tendency and of
the RK4 / low-storage RK steps in torch, on the CPU -- TEST INFRASTRUCTURE.

It is written from the operator formulas of the reference
(src/model/auxiliary_vars/{vorticity,kinetic,thickness,tracer}_aux.hpp,
src/model/tendency_terms/*.hpp, src/time_integration/{rk4,lsrk}.cpp) as
whole-array gathers, independently of the C++ oracle and of the HIP
kernels; torch.autograd then supplies reference gradients for the
library's discrete adjoint.  Its forward values are pinned against the
oracle in tests/test_adjoint_cpu.py.  Arrays use the reference's original
numbering: h[nc][L], vn[ne][L], tr[ntr][nc][L].
"""
from __future__ import annotations

import numpy as np
import torch

PV, KE, SSH, VDEL2, VDEL4, TDEL2, TDEL4 = 1, 2, 4, 8, 16, 32, 64

# rk4.cpp:8-25
RK4_A = (0.5, 0.5, 1.0)
RK4_B = (1.0 / 6.0, 1.0 / 3.0, 1.0 / 3.0, 1.0 / 6.0)


class TorchMesh:
    """Connectivity and metrics of an MPAS mesh object exposing the
    reference's field names (the library mesh or the oracle mesh)."""

    def __init__(self, mesh, nlayers: int):
        g = lambda n: np.asarray(getattr(mesh, n))
        L = nlayers
        self.L = L
        noc = g("nedges_on_cell").astype(np.int64)
        eoc = g("edges_on_cell").astype(np.int64)
        me = eoc.shape[1]
        cm = (np.arange(me)[None, :] < noc[:, None]).astype(np.float64)
        self.eoc = torch.from_numpy(np.where(cm > 0, eoc, 0))
        area = g("area_cell")
        dv = g("dv_edge")
        dc = g("dc_edge")
        sgn_c = g("edge_sign_on_cell") * cm
        eocn = self.eoc.numpy()
        self.cdiv = torch.from_numpy(dv[eocn] * sgn_c / area[:, None])            # [nc][me]
        self.cke = torch.from_numpy(dv[eocn] * dc[eocn] * cm / area[:, None])      # [nc][me]
        self.inv_area = torch.from_numpy(1.0 / area)
        self.dv_sgn = torch.from_numpy(dv[eocn] * sgn_c)                           # [nc][me]
        coe = g("cells_on_edge").astype(np.int64)
        voe = g("vertices_on_edge").astype(np.int64)
        self.c0, self.c1 = torch.from_numpy(coe[:, 0]), torch.from_numpy(coe[:, 1])
        self.v0, self.v1 = torch.from_numpy(voe[:, 0]), torch.from_numpy(voe[:, 1])
        noe = g("nedges_on_edge").astype(np.int64)
        eoe = g("edges_on_edge").astype(np.int64)
        me2 = eoe.shape[1]
        em = np.arange(me2)[None, :] < noe[:, None]
        self.eoe = torch.from_numpy(np.where(em, eoe, 0))
        self.woe = torch.from_numpy(np.where(em, g("weights_on_edge"), 0.0))
        self.inv_dc = torch.from_numpy(1.0 / dc)
        self.inv_dv = torch.from_numpy(1.0 / dv)
        self.inv_dv4 = torch.from_numpy(1.0 / np.maximum(dv, 0.25 * dc))
        self.s2 = torch.from_numpy(np.asarray(g("mesh_scaling_del2"), dtype=np.float64))
        self.s4 = torch.from_numpy(np.asarray(g("mesh_scaling_del4"), dtype=np.float64))
        top = g("max_level_edge_top").astype(np.int64)
        self.mask = torch.from_numpy((np.arange(L)[None, :] < top[:, None]).astype(np.float64))
        eov = g("edges_on_vertex").astype(np.int64)
        cov = g("cells_on_vertex").astype(np.int64)
        at = g("area_triangle")
        self.eov = torch.from_numpy(eov)
        self.cov = torch.from_numpy(cov)
        self.crv = torch.from_numpy(dc[eov] * g("edge_sign_on_vertex") / at[:, None])
        self.cth = torch.from_numpy(g("kiteareas_on_vertex") / at[:, None])
        self.nv = eov.shape[0]


def _cell_sum(coef, edge_field, eoc):
    # sum_j coef[c, j] * X[eoc[c, j], k]
    return (coef[:, :, None] * edge_field[eoc]).sum(dim=1)


def tendency(tm: TorchMesh, params: dict, terms: int, h, vn, tr, f_vertex=None):
    """(h_t, vn_t, tr_t) of compute_tendency (shallow_water_model.cpp:313-347).
    params: grav, visc_del2, visc_del4, eddy_diff2, eddy_diff4 (floats or
    tensors), disable_h_tendency, disable_vn_tendency, f0."""
    g = params["grav"]
    nu2, nu4 = params["visc_del2"], params["visc_del4"]
    ka2, ka4 = params["eddy_diff2"], params["eddy_diff4"]
    fv = f_vertex if f_vertex is not None else torch.full((tm.nv,), float(params["f0"]),
                                                          dtype=torch.float64)
    # vertex diagnostics (vorticity_aux.hpp:85-116)
    zeta = (tm.crv[:, :, None] * vn[tm.eov]).sum(dim=1)
    theta = (tm.cth[:, :, None] * h[tm.cov]).sum(dim=1)
    iota = 1.0 / theta
    q = iota * zeta
    p = iota * fv[:, None]
    # cell diagnostics (kinetic_aux.hpp:44-73)
    K = 0.25 * _cell_sum(tm.cke, vn * vn, tm.eoc)
    D = _cell_sum(tm.cdiv, vn, tm.eoc)
    # edge thickness (thickness_aux.hpp:33-52), edge PV (vorticity_aux.hpp:118-132)
    F = 0.5 * (h[tm.c0] + h[tm.c1])
    S = 0.5 * (q[tm.v0] + q[tm.v1]) + 0.5 * (p[tm.v0] + p[tm.v1])

    ht = -_cell_sum(tm.cdiv, F * vn, tm.eoc)   # thickness_hadv_cell.hpp:48-72
    if params.get("disable_h_tendency", False):
        ht = torch.zeros_like(h)

    ut = torch.zeros_like(vn)
    if not params.get("disable_vn_tendency", False):
        if terms & PV:   # pv_flux_edge.hpp:56-84
            Fu = F * vn
            ut = ut + (tm.woe[:, :, None] * Fu[tm.eoe] * 0.5 * (S[:, None, :] + S[tm.eoe])).sum(dim=1)
        if terms & KE:   # ke_grad_edge.hpp:37-51
            ut = ut - (K[tm.c1] - K[tm.c0]) * tm.inv_dc[:, None]
        if terms & SSH:  # ssh_grad_edge.hpp:37-50
            ut = ut - g * (h[tm.c1] - h[tm.c0]) * tm.inv_dc[:, None]
        if terms & VDEL2:  # vel_diff_edge.hpp:61-86
            lap = (D[tm.c1] - D[tm.c0]) * tm.inv_dc[:, None] - (zeta[tm.v1] - zeta[tm.v0]) * tm.inv_dv[:, None]
            ut = ut + tm.mask * nu2 * tm.s2[:, None] * lap
        if terms & VDEL4:  # vel_hyperdiff_edge.hpp:140-237
            E2 = (D[tm.c1] - D[tm.c0]) * tm.inv_dc[:, None] - (zeta[tm.v1] - zeta[tm.v0]) * tm.inv_dv4[:, None]
            Z2 = (tm.crv[:, :, None] * E2[tm.eov]).sum(dim=1)
            D2 = _cell_sum(tm.cdiv, E2, tm.eoc)
            lap = (D2[tm.c1] - D2[tm.c0]) * tm.inv_dc[:, None] - (Z2[tm.v1] - Z2[tm.v0]) * tm.inv_dv[:, None]
            ut = ut - tm.mask * nu4 * tm.s4[:, None] * lap

    ntr = tr.shape[0]
    tts = []
    for l in range(ntr):
        nu = tr[l] / h   # tracer_aux.hpp:31-51
        ne = 0.5 * (nu[tm.c0] + nu[tm.c1])
        grad = (nu[tm.c1] - nu[tm.c0]) * tm.inv_dc[:, None]
        tt = -_cell_sum(tm.cdiv, F * ne * vn, tm.eoc)       # tracer_hadv_cell.hpp:60-90
        if terms & TDEL2:                                     # tracer_diff_cell.hpp:61-92
            tt = tt + ka2 * _cell_sum(tm.cdiv, F * tm.s2[:, None] * grad, tm.eoc)
        if terms & TDEL4:                                     # tracer_hyperdiff_cell.hpp:94-161
            G = _cell_sum(tm.cdiv, F * grad, tm.eoc)
            gG = (G[tm.c1] - G[tm.c0]) * tm.inv_dc[:, None]
            tt = tt - ka4 * _cell_sum(tm.cdiv, tm.s4[:, None] * gG, tm.eoc)
        tts.append(tt)
    trt = torch.stack(tts) if tts else torch.zeros_like(tr)
    return ht, ut, trt


def rk4_step(tm, params, terms, dt, h, vn, tr):
    """rk4.cpp:28-104 (classical RK4)."""
    y = (h, vn, tr)
    k = tendency(tm, params, terms, *y)
    acc = tuple(a + dt * RK4_B[0] * b for a, b in zip(y, k))
    for s in range(1, 4):
        ys = tuple(a + dt * RK4_A[s - 1] * b for a, b in zip(y, k))
        k = tendency(tm, params, terms, *ys)
        acc = tuple(a + dt * RK4_B[s] * b for a, b in zip(acc, k))
    return acc


def lsrk_step(tm, params, terms, dt, rka, rkb, h, vn, tr):
    """lsrk.cpp:36-89: tend = a_s*tend + F(y); y += dt*b_s*tend."""
    y = (h, vn, tr)
    t = None
    for a, b in zip(rka, rkb):
        k = tendency(tm, params, terms, *y)
        t = k if t is None else tuple(a * ti + ki for ti, ki in zip(t, k))
        y = tuple(yi + dt * b * ti for yi, ti in zip(y, t))
    return y
