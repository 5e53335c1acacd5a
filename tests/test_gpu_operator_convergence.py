"""The reference's operator-convergence check (test/operator_convergence.cpp:
11-302) run through the HIP kernels: each discrete operator is isolated from
a tendency term or an auxiliary array of the library and compared with the
analytic operator on meshes nx = 6, 12, 24, 48, 96 (dc = 1/nx); the rate
between the two finest meshes must be 2 (gradient), 2 (divergence),
1 (curl) and 2 (tangential reconstruction), each within +-0.05 as in the
reference (operator_convergence.cpp:256-301).

* gradient: the SSH-gradient term alone with g = 1 on h = 2 + f gives
  vn_tend = -(f_c1 - f_c0) / dc (ssh_grad_edge.hpp:37-50);
* divergence: the thickness-flux divergence with h = 1 gives
  h_tend = -div(vn) (thickness_hadv_cell.hpp:48-72);
* curl: the auxiliary relative vorticity at vertices (vorticity_aux.hpp:
  94-106);
* reconstruction: the PV flux alone with h = 1 and a planetary vorticity
  f0 = 1e14 that dominates the relative one (|zeta| / f0 < 1e-12), so
  vn_tend / f0 is the TRiSK tangential reconstruction sum_j w_ej vn_ej
  (pv_flux_edge.hpp:56-84) to a relative 1e-12.

64 layers run the pattern-specialised kernels on the meshes that admit them,
1 layer the general ones.
"""
import math

import numpy as np
import pytest

from omega_playground_amd import swe

pytestmark = pytest.mark.gpu

F0 = 1e14


def errors(nx, L):
    m = swe.PlanarHexagonalMesh(nx, nx, None, L)
    lx, ly = m.period_x, m.period_y
    k2x, k2y = 2 * math.pi / lx, 2 * math.pi / ly
    xc, yc = np.asarray(m.x_cell), np.asarray(m.y_cell)
    xe, ye, ae = np.asarray(m.x_edge), np.asarray(m.y_edge), np.asarray(m.angle_edge)
    xv, yv = np.asarray(m.x_vertex), np.asarray(m.y_vertex)
    ne, ce = np.cos(ae), np.sin(ae)
    vx = np.sin(k2x * xe) * np.cos(k2y * ye)
    vy = np.cos(k2x * xe) * np.sin(k2y * ye)
    vn = np.repeat((ne * vx + ce * vy)[:, None], L, 1)
    ones = np.ones((m.ncells, L))

    def tendency(model, h, u):
        s = swe.ShallowWaterState(m, 0)
        s.upload(np.ascontiguousarray(h), np.ascontiguousarray(u))
        t = swe.ShallowWaterState(m, 0)
        model.compute_tendency(t, s, 0.0)
        return t.download(), s

    # gradient through the SSH term
    model = swe.ShallowWaterModel(m, swe.ShallowWaterParams(f0=0.0, grav=1.0))
    model.terms = swe.SSH
    f = np.sin(k2x * xc) * np.sin(k2y * yc)
    (ht, vt, _), _ = tendency(model, 2.0 + np.repeat(f[:, None], L, 1), np.zeros((m.nedges, L)))
    grad = -vt
    gx = k2x * np.cos(k2x * xe) * np.sin(k2y * ye)
    gy = k2y * np.sin(k2x * xe) * np.cos(k2y * ye)
    e_grad = np.max(np.abs((ne * gx + ce * gy)[:, None] - grad))
    # divergence through the thickness flux with h = 1
    model.terms = 0
    (ht, vt, _), _ = tendency(model, ones, vn)
    ediv = 2 * math.pi * (1 / lx + 1 / ly) * np.cos(k2x * xc) * np.cos(k2y * yc)
    e_div = np.max(np.abs(ediv[:, None] + ht))
    # curl: auxiliary relative vorticity
    model.terms = swe.PV
    s = swe.ShallowWaterState(m, 0)
    s.upload(ones, np.ascontiguousarray(vn))
    model.compute_aux(s)
    curl = model.aux_array("rvort_vertex")
    ecurl = 2 * math.pi * (-1 / lx + 1 / ly) * np.sin(k2x * xv) * np.sin(k2y * yv)
    e_curl = np.max(np.abs(ecurl[:, None] - curl))
    assert np.max(np.abs(curl)) < 1e-12 * F0
    # tangential reconstruction through the PV flux with a dominant f0
    model_r = swe.ShallowWaterModel(m, swe.ShallowWaterParams(f0=F0, grav=1.0))
    model_r.terms = swe.PV
    (ht, vt, _), _ = tendency(model_r, ones, vn)
    recon = vt / F0
    e_recon = np.max(np.abs((-ce * vx + ne * vy)[:, None] - recon))
    return (e_grad, e_div, e_curl, e_recon), model_r.kernel_variant()


@pytest.mark.parametrize("L", [1, 64])
def test_operator_convergence_on_gpu(L):
    errs, kinds = [], []
    for lev in range(5):
        nx = 6 * 2 ** lev
        e, kv = errors(nx, L)
        print(f"nx={nx} L={L} kernels={kv} grad {e[0]:.3e} div {e[1]:.3e} curl {e[2]:.3e} "
              f"recon {e[3]:.3e}", flush=True)
        errs.append(e)
        kinds.append(kv)
    if L == 64:
        assert kinds[-1] == 2  # the finest meshes ran the pattern-specialised kernels
    rates = [math.log2(errs[-2][i] / errs[-1][i]) for i in range(4)]
    print("rates", rates, flush=True)
    for r, expected in zip(rates, (2, 2, 1, 2)):
        assert abs(r - expected) < 0.05, rates
