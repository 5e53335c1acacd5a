// Per-entity device operators of the forward path, shared by the
// single-model kernels (kernels_fwd.hip) and the ensemble kernels
// (kernels_batch.hip).  Each function computes one entity at one
// level-vector k..k+N-1 and is free of thread indexing.
#pragma once

#include "device_common.hpp"

namespace omega_hip {
namespace fwd {

using namespace dev;
// ---------------------------------------------------------------------------
// Diagnostics: vertex vorticity / thickness, cell KE / divergence, tracer
// Laplacian for tracer del4.
// ---------------------------------------------------------------------------

// VorticityAuxVars::compute_vort_vertex (vorticity_aux.hpp:85-116); stores
// rvort and 1/thickness, from which nrv = inv*rvort and npv = inv*f are
// formed where needed.
template <int N>
__device__ __forceinline__ void vertex_diag(const MeshView &m, int v, int k,
                                            const double *__restrict__ h,
                                            const double *__restrict__ vn, const DiagOut &d) {
  const int L = m.L;
  Vd<N> rv = zero<N>(), th = zero<N>();
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    const int e = m.eov[3 * v + j];
    const int c = m.cov[3 * v + j];
    const double crv = m.crv[3 * v + j];
    const double cth = m.cth[3 * v + j];
    const Vd<N> ve = ld<N>(vn + (size_t)e * L + k);
    const Vd<N> hc = ld<N>(h + (size_t)c * L + k);
#pragma unroll
    for (int i = 0; i < N; ++i) {
      rv[i] = rv[i] + crv * ve[i];
      th[i] = th[i] + cth * hc[i];
    }
  }
  Vd<N> inv;
#pragma unroll
  for (int i = 0; i < N; ++i) inv[i] = 1.0 / th[i];
  st<N>(d.rvort_v + (size_t)v * L + k, rv);
  st<N>(d.inv_thick_v + (size_t)v * L + k, inv);
}

// KineticAuxVars::compute_kinetic_cell (kinetic_aux.hpp:44-73) and, for
// tracer del4, TracerHyperDiffusionOnCell::compute_tracer_del2
// (tracer_hyperdiff_cell.hpp:94-129) with norm_tr = tr*(1/h)
// (tracer_aux.hpp:31-51) formed on the fly.
template <int N>
__device__ __forceinline__ void cell_diag(const MeshView &m, int terms, int ntr, int c, int k,
                                          const double *__restrict__ h,
                                          const double *__restrict__ vn,
                                          const double *__restrict__ tr, const DiagOut &d) {
  const int L = m.L, me = m.me;
  const int n = m.noc[c];
  if (terms & (T_KE | T_VDEL2 | T_VDEL4 | T_DRAG)) {
    Vd<N> ke = zero<N>(), div = zero<N>();
    for (int j = 0; j < n; ++j) {
      const int e = m.eoc[c * me + j];
      const double cke = m.cke[c * me + j];
      const double cd = m.cdiv[c * me + j];
      const Vd<N> ve = ld<N>(vn + (size_t)e * L + k);
#pragma unroll
      for (int i = 0; i < N; ++i) {
        ke[i] = ke[i] + cke * ve[i] * ve[i] * 0.25;
        div[i] = div[i] + cd * ve[i];
      }
    }
    st<N>(d.ke_c + (size_t)c * L + k, ke);
    st<N>(d.div_c + (size_t)c * L + k, div);
  }
  if ((terms & T_TDEL4) && ntr > 0) {
    const Vd<N> hs = ld<N>(h + (size_t)c * L + k);
    Vd<N> inv_s;
#pragma unroll
    for (int i = 0; i < N; ++i) inv_s[i] = 1.0 / hs[i];
    for (int l0 = 0; l0 < ntr; l0 += kTrGroup) {
      Vd<N> d2[kTrGroup];
#pragma unroll
      for (int g = 0; g < kTrGroup; ++g) d2[g] = zero<N>();
      for (int j = 0; j < n; ++j) {
        const int e = m.eoc[c * me + j];
        const int a0 = m.coe[2 * e], a1 = m.coe[2 * e + 1];
        const double cd = m.cdiv[c * me + j];
        const double idc = m.inv_dc[e];
        const Vd<N> h0 = ld<N>(h + (size_t)a0 * L + k);
        const Vd<N> h1 = ld<N>(h + (size_t)a1 * L + k);
        Vd<N> inv0, inv1, cm;
#pragma unroll
        for (int i = 0; i < N; ++i) {
          inv0[i] = (a0 == c) ? inv_s[i] : 1.0 / h0[i];
          inv1[i] = (a1 == c) ? inv_s[i] : 1.0 / h1[i];
          cm[i] = cd * (0.5 * (h0[i] + h1[i]));
        }
#pragma unroll
        for (int g = 0; g < kTrGroup; ++g) {
          const int l = l0 + g;
          if (l < ntr) {
            const double *trl = tr + (size_t)l * m.nc * L;
            const Vd<N> t0 = ld<N>(trl + (size_t)a0 * L + k);
            const Vd<N> t1 = ld<N>(trl + (size_t)a1 * L + k);
#pragma unroll
            for (int i = 0; i < N; ++i) {
              const double grad = (t1[i] * inv1[i] - t0[i] * inv0[i]) * idc;
              d2[g][i] = d2[g][i] + cm[i] * grad;
            }
          }
        }
      }
#pragma unroll
      for (int g = 0; g < kTrGroup; ++g) {
        const int l = l0 + g;
        if (l < ntr) st<N>(d.trdel2_c + ((size_t)l * m.nc + c) * L + k, d2[g]);
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Momentum del4 pre-passes (vel_hyperdiff_edge.hpp:140-210)
// ---------------------------------------------------------------------------
// compute_vel_del2 (vel_hyperdiff_edge.hpp:140-164)
template <int N>
__device__ __forceinline__ void vdel4_edge(const MeshView &m, const DiagOut &d, int e, int k) {
  const int L = m.L;
  const int a0 = m.coe[2 * e], a1 = m.coe[2 * e + 1];
  const int b0 = m.voe[2 * e], b1 = m.voe[2 * e + 1];
  const double idc = m.inv_dc[e], idv = m.inv_dv4[e];
  const Vd<N> d0 = ld<N>(d.div_c + (size_t)a0 * L + k), d1 = ld<N>(d.div_c + (size_t)a1 * L + k);
  const Vd<N> r0 = ld<N>(d.rvort_v + (size_t)b0 * L + k), r1 = ld<N>(d.rvort_v + (size_t)b1 * L + k);
  Vd<N> out;
#pragma unroll
  for (int i = 0; i < N; ++i) out[i] = (d1[i] - d0[i]) * idc - (r1[i] - r0[i]) * idv;
  st<N>(d.vdel2_e + (size_t)e * L + k, out);
}

// compute_vel_del2_rvort (vel_hyperdiff_edge.hpp:166-187)
template <int N>
__device__ __forceinline__ void vdel4_vertex(const MeshView &m, const DiagOut &d, int v, int k) {
  const int L = m.L;
  Vd<N> rv = zero<N>();
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    const int e = m.eov[3 * v + j];
    const double crv = m.crv[3 * v + j];
    const Vd<N> x = ld<N>(d.vdel2_e + (size_t)e * L + k);
#pragma unroll
    for (int i = 0; i < N; ++i) rv[i] = rv[i] + crv * x[i];
  }
  st<N>(d.vdel2_rv_v + (size_t)v * L + k, rv);
}

// compute_vel_del2_div (vel_hyperdiff_edge.hpp:189-210)
template <int N>
__device__ __forceinline__ void vdel4_cell(const MeshView &m, const DiagOut &d, int c, int k) {
  const int L = m.L, me = m.me;
  Vd<N> acc = zero<N>();
  const int n = m.noc[c];
  for (int j = 0; j < n; ++j) {
    const int e = m.eoc[c * me + j];
    const double cd = m.cdiv[c * me + j];
    const Vd<N> x = ld<N>(d.vdel2_e + (size_t)e * L + k);
#pragma unroll
    for (int i = 0; i < N; ++i) acc[i] = acc[i] + cd * x[i];
  }
  st<N>(d.vdel2_div_c + (size_t)c * L + k, acc);
}

// ---------------------------------------------------------------------------
// Tendencies + stage update
// ---------------------------------------------------------------------------

// Edge-averaged normalised relative and planetary vorticity of edge e
// (vorticity_aux.hpp:108-131): nrv_v = inv_thick*rvort, npv_v = inv_thick*f.
template <int N>
__device__ __forceinline__ void edge_vort(const MeshView &m, const ModelCoef &mc, const DiagOut &d,
                                          int e, int k, Vd<N> &nrv, Vd<N> &npv) {
  const int L = m.L;
  const int b0 = m.voe[2 * e], b1 = m.voe[2 * e + 1];
  const double f0 = mc.f_v[b0], f1 = mc.f_v[b1];
  const Vd<N> r0 = ld<N>(d.rvort_v + (size_t)b0 * L + k), r1 = ld<N>(d.rvort_v + (size_t)b1 * L + k);
  const Vd<N> i0 = ld<N>(d.inv_thick_v + (size_t)b0 * L + k);
  const Vd<N> i1 = ld<N>(d.inv_thick_v + (size_t)b1 * L + k);
#pragma unroll
  for (int i = 0; i < N; ++i) {
    nrv[i] = 0.5 * (i0[i] * r0[i] + i1[i] * r1[i]);
    npv[i] = 0.5 * (i0[i] * f0 + i1[i] * f1);
  }
}

// ShallowWaterModel::compute_vn_tendency (shallow_water_model.cpp:186-217):
// PV flux (pv_flux_edge.hpp:56-84), KE gradient (ke_grad_edge.hpp:37-51),
// SSH gradient (ssh_grad_edge.hpp:37-50), del2 (vel_diff_edge.hpp:61-86),
// del4 (vel_hyperdiff_edge.hpp:212-237), in that order.
template <int N>
__device__ __forceinline__ void edge_tend(const MeshView &m, const ModelCoef &mc,
                                          const StageArgs &a, const DiagOut &d, int e, int k) {
  const int L = m.L;
  const size_t idx = (size_t)e * L + k;
  if (!a.do_vn) {
    if (a.epi == EPI_STORE) return; // tendency left untouched
    epilogue<N>(a, a.vn, idx, tend_base<N>(a, a.vn, idx));
    return;
  }
  const double *__restrict__ h = a.h.in;
  const double *__restrict__ vn = a.vn.in;
  Vd<N> T = tend_base<N>(a, a.vn, idx);
  const int c0 = m.coe[2 * e], c1 = m.coe[2 * e + 1];
  const double idc = m.inv_dc[e];
  if (a.terms & T_PV) {
    Vd<N> nrv_e, npv_e;
    edge_vort<N>(m, mc, d, e, k, nrv_e, npv_e);
    Vd<N> qs;
#pragma unroll
    for (int i = 0; i < N; ++i) qs[i] = nrv_e[i] + npv_e[i];
    Vd<N> qt = zero<N>();
    const int n = m.noe[e];
    for (int j = 0; j < n; ++j) {
      const int ej = m.eoe[e * m.me2 + j];
      const double wj = m.w[e * m.me2 + j];
      const int a0 = m.coe[2 * ej], a1 = m.coe[2 * ej + 1];
      const Vd<N> h0 = ld<N>(h + (size_t)a0 * L + k), h1 = ld<N>(h + (size_t)a1 * L + k);
      const Vd<N> vj = ld<N>(vn + (size_t)ej * L + k);
      Vd<N> nrv_j, npv_j;
      edge_vort<N>(m, mc, d, ej, k, nrv_j, npv_j);
#pragma unroll
      for (int i = 0; i < N; ++i) {
        const double fh = 0.5 * (h0[i] + h1[i]);
        const double nv = (qs[i] + nrv_j[i] + npv_j[i]) * 0.5;
        qt[i] = qt[i] + wj * fh * vj[i] * nv;
      }
    }
#pragma unroll
    for (int i = 0; i < N; ++i) T[i] = T[i] + qt[i];
  }
  if (a.terms & T_KE) {
    const Vd<N> k0 = ld<N>(d.ke_c + (size_t)c0 * L + k), k1 = ld<N>(d.ke_c + (size_t)c1 * L + k);
#pragma unroll
    for (int i = 0; i < N; ++i) T[i] = T[i] - (k1[i] - k0[i]) * idc;
  }
  if (a.terms & T_SSH) {
    const Vd<N> h0 = ld<N>(h + (size_t)c0 * L + k), h1 = ld<N>(h + (size_t)c1 * L + k);
#pragma unroll
    for (int i = 0; i < N; ++i) T[i] = T[i] - mc.grav * (h1[i] - h0[i]) * idc;
  }
  if (a.terms & T_DRAG) {
    // bottom drag on the edge's deepest active level, the formula left as
    // a comment at shallow_water_hierarchical.cpp:479-485:
    // -drag_coeff * sqrt(ke_c0 + ke_c1) * vn * (1 / h_drag), with h_drag the
    // edge-mean thickness 0.5 * (h_c0 + h_c1)
    const int ml = m.maxlev_e[e];
    if (ml - 1 >= k && ml - 1 < k + N) {
      const int i = ml - 1 - k;
      const double k0 = d.ke_c[(size_t)c0 * L + k + i], k1 = d.ke_c[(size_t)c1 * L + k + i];
      const double h0 = h[(size_t)c0 * L + k + i], h1 = h[(size_t)c1 * L + k + i];
      const double u = vn[idx + i];
      const double inv_h = 1.0 / (0.5 * (h0 + h1));
      const double drag = -mc.drag * sqrt(k0 + k1) * u * inv_h;
#pragma unroll
      for (int q = 0; q < N; ++q)
        if (q == i) T[q] = T[q] + drag;
    }
  }
  if (a.terms & (T_VDEL2 | T_VDEL4)) {
    const int b0 = m.voe[2 * e], b1 = m.voe[2 * e + 1];
    const double idv = m.inv_dv[e];
    const int ml = m.maxlev_e[e];
    if (a.terms & T_VDEL2) {
      const Vd<N> d0 = ld<N>(d.div_c + (size_t)c0 * L + k), d1 = ld<N>(d.div_c + (size_t)c1 * L + k);
      const Vd<N> r0 = ld<N>(d.rvort_v + (size_t)b0 * L + k), r1 = ld<N>(d.rvort_v + (size_t)b1 * L + k);
      const double cv = mc.cv2[e];
#pragma unroll
      for (int i = 0; i < N; ++i) {
        const double del2u = (d1[i] - d0[i]) * idc - (r1[i] - r0[i]) * idv;
        const double coef = (k + i < ml) ? cv : 0.0;
        T[i] = T[i] + coef * del2u;
      }
    }
    if (a.terms & T_VDEL4) {
      const Vd<N> d0 = ld<N>(d.vdel2_div_c + (size_t)c0 * L + k);
      const Vd<N> d1 = ld<N>(d.vdel2_div_c + (size_t)c1 * L + k);
      const Vd<N> r0 = ld<N>(d.vdel2_rv_v + (size_t)b0 * L + k);
      const Vd<N> r1 = ld<N>(d.vdel2_rv_v + (size_t)b1 * L + k);
      const double cv = mc.cv4[e];
#pragma unroll
      for (int i = 0; i < N; ++i) {
        const double del2u = (d1[i] - d0[i]) * idc - (r1[i] - r0[i]) * idv;
        const double coef = (k + i < ml) ? cv : 0.0;
        T[i] = T[i] - coef * del2u;
      }
    }
  }
  epilogue<N>(a, a.vn, idx, T);
}

// compute_h_tendency (shallow_water_model.cpp:71-84, thickness_hadv_cell.hpp:48-72)
// and compute_tr_tendency (shallow_water_model.cpp:285-309): tracer advection
// (tracer_hadv_cell.hpp:60-90), del2 (tracer_diff_cell.hpp:61-92) and del4
// (tracer_hyperdiff_cell.hpp:131-161), each summed separately then added.
template <int N>
__device__ __forceinline__ void cell_tend(const MeshView &m, const ModelCoef &mc,
                                          const StageArgs &a, const DiagOut &d, int c, int k) {
  const int L = m.L, me = m.me, NC = m.nc;
  const double *__restrict__ h = a.h.in;
  const double *__restrict__ vn = a.vn.in;
  const int n = m.noc[c];
  const size_t hidx = (size_t)c * L + k;
  if (a.do_h) {
    Vd<N> T = tend_base<N>(a, a.h, hidx);
    Vd<N> acc = zero<N>();
    for (int j = 0; j < n; ++j) {
      const int e = m.eoc[c * me + j];
      const int a0 = m.coe[2 * e], a1 = m.coe[2 * e + 1];
      const double cd = m.cdiv[c * me + j];
      const Vd<N> h0 = ld<N>(h + (size_t)a0 * L + k), h1 = ld<N>(h + (size_t)a1 * L + k);
      const Vd<N> ve = ld<N>(vn + (size_t)e * L + k);
#pragma unroll
      for (int i = 0; i < N; ++i) acc[i] = acc[i] - cd * (0.5 * (h0[i] + h1[i])) * ve[i];
    }
#pragma unroll
    for (int i = 0; i < N; ++i) T[i] = T[i] + acc[i];
    epilogue<N>(a, a.h, hidx, T);
  } else if (a.epi != EPI_STORE) {
    epilogue<N>(a, a.h, hidx, tend_base<N>(a, a.h, hidx));
  }

  const int ntr = a.ntr;
  if (ntr <= 0) return;
  const bool td2 = a.terms & T_TDEL2, td4 = a.terms & T_TDEL4;
  const double *__restrict__ tr = a.tr.in;
  const Vd<N> hs = ld<N>(h + hidx);
  Vd<N> inv_s;
#pragma unroll
  for (int i = 0; i < N; ++i) inv_s[i] = 1.0 / hs[i];
  for (int l0 = 0; l0 < ntr; l0 += kTrGroup) {
    Vd<N> adv[kTrGroup], dif[kTrGroup], hyp[kTrGroup];
#pragma unroll
    for (int g = 0; g < kTrGroup; ++g) {
      adv[g] = zero<N>();
      dif[g] = zero<N>();
      hyp[g] = zero<N>();
    }
    for (int j = 0; j < n; ++j) {
      const int e = m.eoc[c * me + j];
      const int a0 = m.coe[2 * e], a1 = m.coe[2 * e + 1];
      const double cd = m.cdiv[c * me + j];
      const double idc = m.inv_dc[e];
      const Vd<N> h0 = ld<N>(h + (size_t)a0 * L + k), h1 = ld<N>(h + (size_t)a1 * L + k);
      const Vd<N> ve = ld<N>(vn + (size_t)e * L + k);
      Vd<N> inv0, inv1, fh;
#pragma unroll
      for (int i = 0; i < N; ++i) {
        inv0[i] = (a0 == c) ? inv_s[i] : 1.0 / h0[i];
        inv1[i] = (a1 == c) ? inv_s[i] : 1.0 / h1[i];
        fh[i] = 0.5 * (h0[i] + h1[i]);
      }
      const double c2 = td2 ? mc.ct2[c * me + j] : 0.0;
      const double sc2 = td2 ? m.s2[e] : 0.0;
      const double c4 = td4 ? mc.ct4[c * me + j] : 0.0;
#pragma unroll
      for (int g = 0; g < kTrGroup; ++g) {
        const int l = l0 + g;
        if (l < ntr) {
          const double *trl = tr + (size_t)l * NC * L;
          const Vd<N> t0 = ld<N>(trl + (size_t)a0 * L + k), t1 = ld<N>(trl + (size_t)a1 * L + k);
          Vd<N> n0, n1;
#pragma unroll
          for (int i = 0; i < N; ++i) {
            n0[i] = t0[i] * inv0[i];
            n1[i] = t1[i] * inv1[i];
            const double ntr_e = (n0[i] + n1[i]) * 0.5;
            adv[g][i] = adv[g][i] - cd * fh[i] * ntr_e * ve[i];
          }
          if (td2) {
#pragma unroll
            for (int i = 0; i < N; ++i) {
              const double grad = (n1[i] - n0[i]) * idc;
              dif[g][i] = dif[g][i] + c2 * fh[i] * sc2 * grad;
            }
          }
          if (td4) {
            const double *dl = d.trdel2_c + (size_t)l * NC * L;
            const Vd<N> q0 = ld<N>(dl + (size_t)a0 * L + k), q1 = ld<N>(dl + (size_t)a1 * L + k);
#pragma unroll
            for (int i = 0; i < N; ++i) {
              const double grad = (q1[i] - q0[i]) * idc;
              hyp[g][i] = hyp[g][i] - c4 * grad;
            }
          }
        }
      }
    }
#pragma unroll
    for (int g = 0; g < kTrGroup; ++g) {
      const int l = l0 + g;
      if (l < ntr) {
        const size_t tidx = ((size_t)l * NC + c) * L + k;
        Vd<N> T = tend_base<N>(a, a.tr, tidx);
#pragma unroll
        for (int i = 0; i < N; ++i) {
          T[i] = T[i] + adv[g][i];
          if (td2) T[i] = T[i] + dif[g][i];
          if (td4) T[i] = T[i] + hyp[g][i];
        }
        epilogue<N>(a, a.tr, tidx, T);
      }
    }
  }
}

} // namespace fwd
} // namespace omega_hip
