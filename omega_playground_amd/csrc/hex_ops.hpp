// Per-entity device operators of the pattern-specialised forward path
// (regular hexagonal mesh, layer count a multiple of 64), shared by the
// single-model kernels (kernels_hex.hip) and the ensemble kernels
// (kernels_batch.hip).  Cell c owns edges 3c..3c+2 and vertices 2c, 2c+1
// in device numbering.
#pragma once

#include "device_common.hpp"
#include "hex_pattern.hpp"

namespace omega_hip {
namespace hexops {

using namespace dev;

// Entity index, made wavefront-uniform when one entity fills the wavefront.
template <int N> __device__ __forceinline__ int ent(int x) {
  if constexpr (N == 1) return uniform(x);
  else return x;
}

// Row e of a [rows][L] array at levels k..k+N-1.  W = false when every
// array the kernel reads has fewer than 2^24 rows and is below 4 GiB
// (MeshView::off32, checked on the host): the address is then the array's
// uniform base plus one 32-bit lane offset (one 24-bit multiply-add instead
// of 64-bit index arithmetic).
template <int N, bool W = true>
__device__ __forceinline__ Vd<N> rowv(const double *__restrict__ p, int e, int L, int k) {
  if (W) return ld<N>(p + (size_t)e * L + k);
  const unsigned off = __umul24((unsigned)e, (unsigned)L * 8u) + (unsigned)k * 8u;
  return ld<N>(reinterpret_cast<const double *>(reinterpret_cast<const char *>(p) + off));
}

// ---------------------------------------------------------------------------
// Diagnostics
// ---------------------------------------------------------------------------

// VorticityAuxVars::compute_vort_vertex (vorticity_aux.hpp:85-116)
template <int N>
__device__ __forceinline__ void vertex_diag_u(const MeshView &m, int v, int k,
                                              const double *__restrict__ h,
                                              const double *__restrict__ vn, const DiagOut &d) {
  const int L = m.L;
  int e[3], c[3];
  double crv[3], cth[3];
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    e[j] = m.eov[3 * v + j];
    c[j] = m.cov[3 * v + j];
    crv[j] = m.crv[3 * v + j];
    cth[j] = m.cth[3 * v + j];
  }
  Vd<N> ve[3], hc[3];
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    ve[j] = rowv<N>(vn, e[j], L, k);
    hc[j] = rowv<N>(h, c[j], L, k);
  }
  Vd<N> rv = zero<N>(), th = zero<N>();
#pragma unroll
  for (int j = 0; j < 3; ++j)
#pragma unroll
    for (int i = 0; i < N; ++i) {
      rv[i] = rv[i] + crv[j] * ve[j][i];
      th[i] = th[i] + cth[j] * hc[j][i];
    }
  Vd<N> inv;
#pragma unroll
  for (int i = 0; i < N; ++i) inv[i] = 1.0 / th[i];
  st<N>(d.rvort_v + (size_t)v * L + k, rv);
  st<N>(d.inv_thick_v + (size_t)v * L + k, inv);
}

// KineticAuxVars::compute_kinetic_cell (kinetic_aux.hpp:44-73) and
// TracerHyperDiffusionOnCell::compute_tracer_del2
// (tracer_hyperdiff_cell.hpp:94-129), norm_tr = tr*(1/h) formed on the fly.
template <int N>
__device__ __forceinline__ void cell_diag_u(const MeshView &m, int terms, int ntr, int c, int k,
                                            const double *__restrict__ h,
                                            const double *__restrict__ vn,
                                            const double *__restrict__ tr, const DiagOut &d) {
  const int L = m.L, me = m.me;
  int E[6];
#pragma unroll
  for (int j = 0; j < 6; ++j) E[j] = m.eoc[c * me + j];
  if (terms & (T_KE | T_VDEL2 | T_VDEL4)) {
    Vd<N> ve[6];
#pragma unroll
    for (int j = 0; j < 6; ++j) ve[j] = rowv<N>(vn, E[j], L, k);
    Vd<N> ke = zero<N>(), div = zero<N>();
#pragma unroll
    for (int j = 0; j < 6; ++j) {
      const double cke = m.cke[c * me + j];
      const double cd = m.cdiv[c * me + j];
#pragma unroll
      for (int i = 0; i < N; ++i) {
        ke[i] = ke[i] + cke * ve[j][i] * ve[j][i] * 0.25;
        div[i] = div[i] + cd * ve[j][i];
      }
    }
    st<N>(d.ke_c + (size_t)c * L + k, ke);
    st<N>(d.div_c + (size_t)c * L + k, div);
  }
  if ((terms & T_TDEL4) && ntr > 0) {
    int NB[6];
#pragma unroll
    for (int j = 0; j < 6; ++j) NB[j] = m.coc[c * 6 + j];
    const Vd<N> hC = rowv<N>(h, c, L, k);
    Vd<N> hN[6], invN[6], cm[6], invC;
    double idc[6];
#pragma unroll
    for (int j = 0; j < 6; ++j) hN[j] = rowv<N>(h, NB[j], L, k);
#pragma unroll
    for (int i = 0; i < N; ++i) invC[i] = 1.0 / hC[i];
#pragma unroll
    for (int j = 0; j < 6; ++j) {
      const double cd = m.cdiv[c * me + j];
      idc[j] = m.inv_dc[E[j]];
#pragma unroll
      for (int i = 0; i < N; ++i) {
        invN[j][i] = 1.0 / hN[j][i];
        // h_mean = 0.5*(h[cell0] + h[cell1]); a two-term sum is the same in
        // either order
        cm[j][i] = cd * (0.5 * (hN[j][i] + hC[i]));
      }
    }
    for (int l = 0; l < ntr; ++l) {
      const double *trl = tr + (size_t)l * m.nc * L;
      const Vd<N> tC = rowv<N>(trl, c, L, k);
      Vd<N> tN[6];
#pragma unroll
      for (int j = 0; j < 6; ++j) tN[j] = rowv<N>(trl, NB[j], L, k);
      Vd<N> d2 = zero<N>();
#pragma unroll
      for (int j = 0; j < 6; ++j)
#pragma unroll
        for (int i = 0; i < N; ++i) {
          const double nC = tC[i] * invC[i];
          const double nN = tN[j][i] * invN[j][i];
          // grad = (norm_tr[cell1] - norm_tr[cell0]) * inv_dc
          const double grad = (hex::cell_is_second(j) ? (nC - nN) : (nN - nC)) * idc[j];
          d2[i] = d2[i] + cm[j][i] * grad;
        }
      st<N>(d.trdel2_c + ((size_t)l * m.nc + c) * L + k, d2);
    }
  }
}

// ---------------------------------------------------------------------------
// Tendencies + stage update
// ---------------------------------------------------------------------------

// compute_vn_tendency for an edge of class K (shallow_water_model.cpp:186-217;
// pv_flux_edge.hpp:56-84, ke_grad_edge.hpp:37-51, ssh_grad_edge.hpp:37-50,
// vel_diff_edge.hpp:61-86, vel_hyperdiff_edge.hpp:212-237)
// EM: also write the edge's PreOut rows (its PV sum and flux-weighted
// edges-on-edge sum with the PV term on, its thickness flux always), formed
// by the same expressions as the tendency VJP's own pre-pass
template <int N, int K, int TM = -1, bool EM = false>
__device__ __forceinline__ void edge_tend_hex(const MeshView &m, const ModelCoef &mc,
                                              const StageArgs &a, const DiagOut &d, int e, int k,
                                              const PreOut &po = PreOut{}) {
  // TM >= 0: the term set, compiled in, with both tendencies on
  const int terms = TM >= 0 ? TM : a.terms;
  const bool do_vn = TM >= 0 || a.do_vn, do_h = TM >= 0 || a.do_h;
  (void)do_vn;
  (void)do_h;
  const int L = m.L;
  const size_t idx = (size_t)e * L + k;
  if (!do_vn) {
    if (a.epi == EPI_STORE) return;
    epilogue<N>(a, a.vn, idx, tend_base<N>(a, a.vn, idx));
    return;
  }
  const double *__restrict__ h = a.h.in;
  const double *__restrict__ vn = a.vn.in;
  const int *rec = m.hexrec + (size_t)e * hex::kRecInts;
  int ej[10], uv[10], uc[10];
#pragma unroll
  for (int j = 0; j < 10; ++j) {
    ej[j] = rec[j];
    uv[j] = rec[10 + j];
    uc[j] = rec[20 + j];
  }
  constexpr int o0 = hex::OWNV[K][0], o1 = hex::OWNV[K][1];
  constexpr int q0 = hex::OWNC[K][0], q1 = hex::OWNC[K][1];
  const bool pv = terms & T_PV;
  Vd<N> hU[10], rvU[10], itU[10], vJ[10];
#pragma unroll
  for (int u = 0; u < 10; ++u) hU[u] = rowv<N>(h, uc[u], L, k);
  if (pv) {
#pragma unroll
    for (int j = 0; j < 10; ++j) vJ[j] = rowv<N>(vn, ej[j], L, k);
#pragma unroll
    for (int u = 0; u < 10; ++u) {
      rvU[u] = rowv<N>(d.rvort_v, uv[u], L, k);
      itU[u] = rowv<N>(d.inv_thick_v, uv[u], L, k);
    }
  } else if (terms & T_VDEL2) {
    rvU[o0] = rowv<N>(d.rvort_v, uv[o0], L, k);
    rvU[o1] = rowv<N>(d.rvort_v, uv[o1], L, k);
  }
  const int c0 = uc[q0], c1 = uc[q1];
  Vd<N> T = tend_base<N>(a, a.vn, idx);
  const double idc = m.inv_dc[e];
  if (pv) {
    Vd<N> nrvU[10], npvU[10];
#pragma unroll
    for (int u = 0; u < 10; ++u) {
      const double f = mc.f_v[uv[u]];
#pragma unroll
      for (int i = 0; i < N; ++i) {
        nrvU[u][i] = itU[u][i] * rvU[u][i];
        npvU[u][i] = itU[u][i] * f;
      }
    }
    double w[10];
#pragma unroll
    for (int j = 0; j < 10; ++j) w[j] = m.w[e * 10 + j];
    Vd<N> S, O;
#pragma unroll
    for (int i = 0; i < N; ++i) {
      const double qs = 0.5 * (nrvU[o0][i] + nrvU[o1][i]) + 0.5 * (npvU[o0][i] + npvU[o1][i]);
      double qt = 0.0, ow = 0.0;
#pragma unroll
      for (int j = 0; j < 10; ++j) {
        const double fh = 0.5 * (hU[hex::LC0[K][j]][i] + hU[hex::LC1[K][j]][i]);
        const double nrv = 0.5 * (nrvU[hex::LV0[K][j]][i] + nrvU[hex::LV1[K][j]][i]);
        const double npv = 0.5 * (npvU[hex::LV0[K][j]][i] + npvU[hex::LV1[K][j]][i]);
        const double nv = (qs + nrv + npv) * 0.5;
        qt = qt + w[j] * fh * vJ[j][i] * nv;
        if constexpr (EM) ow = ow + w[j] * (fh * vJ[j][i]);
      }
      T[i] = T[i] + qt;
      S[i] = qs;
      O[i] = ow;
    }
    if constexpr (EM) {
      st<N>(po.se + idx, S);
      st<N>(po.own + idx, O);
    }
  }
  if constexpr (EM) {
    const Vd<N> ue = rowv<N>(vn, e, L, k);
    Vd<N> FU;
#pragma unroll
    for (int i = 0; i < N; ++i) FU[i] = 0.5 * (hU[q0][i] + hU[q1][i]) * ue[i];
    st<N>(po.fu + idx, FU);
  }
  if (terms & T_KE) {
    const Vd<N> k0 = rowv<N>(d.ke_c, c0, L, k), k1 = rowv<N>(d.ke_c, c1, L, k);
#pragma unroll
    for (int i = 0; i < N; ++i) T[i] = T[i] - (k1[i] - k0[i]) * idc;
  }
  if (terms & T_SSH) {
#pragma unroll
    for (int i = 0; i < N; ++i) T[i] = T[i] - mc.grav * (hU[q1][i] - hU[q0][i]) * idc;
  }
  if (terms & (T_VDEL2 | T_VDEL4)) {
    const int ml = m.maxlev_e[e];
    const double idv = m.inv_dv[e];
    if (terms & T_VDEL2) {
      const Vd<N> d0 = rowv<N>(d.div_c, c0, L, k), d1 = rowv<N>(d.div_c, c1, L, k);
      const double cv = mc.cv2[e];
#pragma unroll
      for (int i = 0; i < N; ++i) {
        const double del2u = (d1[i] - d0[i]) * idc - (rvU[o1][i] - rvU[o0][i]) * idv;
        T[i] = T[i] + ((k + i < ml) ? cv : 0.0) * del2u;
      }
    }
    if (terms & T_VDEL4) {
      const int b0 = uv[o0], b1 = uv[o1];
      const Vd<N> d0 = rowv<N>(d.vdel2_div_c, c0, L, k), d1 = rowv<N>(d.vdel2_div_c, c1, L, k);
      const Vd<N> r0 = rowv<N>(d.vdel2_rv_v, b0, L, k), r1 = rowv<N>(d.vdel2_rv_v, b1, L, k);
      const double cv = mc.cv4[e];
#pragma unroll
      for (int i = 0; i < N; ++i) {
        const double del2u = (d1[i] - d0[i]) * idc - (r1[i] - r0[i]) * idv;
        T[i] = T[i] - ((k + i < ml) ? cv : 0.0) * del2u;
      }
    }
  }
  epilogue<N>(a, a.vn, idx, T);
}

// compute_h_tendency / compute_tr_tendency for one cell
// (thickness_hadv_cell.hpp:48-72, tracer_hadv_cell.hpp:60-90,
// tracer_diff_cell.hpp:61-92, tracer_hyperdiff_cell.hpp:131-161)
template <int N, int TM = -1>
__device__ __forceinline__ void cell_tend_hex(const MeshView &m, const ModelCoef &mc,
                                              const StageArgs &a, const DiagOut &d, int c, int k) {
  // TM >= 0: the term set, compiled in, with both tendencies on
  const int terms = TM >= 0 ? TM : a.terms;
  const bool do_vn = TM >= 0 || a.do_vn, do_h = TM >= 0 || a.do_h;
  (void)do_vn;
  (void)do_h;
  const int L = m.L, me = m.me, NC = m.nc;
  const double *__restrict__ h = a.h.in;
  const double *__restrict__ vn = a.vn.in;
  int E[6], NB[6];
#pragma unroll
  for (int j = 0; j < 6; ++j) {
    E[j] = m.eoc[c * me + j];
    NB[j] = m.coc[c * 6 + j];
  }
  const Vd<N> hC = rowv<N>(h, c, L, k);
  Vd<N> hN[6], vE[6], fh[6];
  double cd[6];
#pragma unroll
  for (int j = 0; j < 6; ++j) {
    hN[j] = rowv<N>(h, NB[j], L, k);
    vE[j] = rowv<N>(vn, E[j], L, k);
  }
#pragma unroll
  for (int j = 0; j < 6; ++j) {
    cd[j] = m.cdiv[c * me + j];
#pragma unroll
    for (int i = 0; i < N; ++i) fh[j][i] = 0.5 * (hN[j][i] + hC[i]);
  }
  const size_t hidx = (size_t)c * L + k;
  if (do_h) {
    Vd<N> T = tend_base<N>(a, a.h, hidx);
#pragma unroll
    for (int i = 0; i < N; ++i) {
      double acc = 0.0;
#pragma unroll
      for (int j = 0; j < 6; ++j) acc = acc - cd[j] * fh[j][i] * vE[j][i];
      T[i] = T[i] + acc;
    }
    epilogue<N>(a, a.h, hidx, T);
  } else if (a.epi != EPI_STORE) {
    epilogue<N>(a, a.h, hidx, tend_base<N>(a, a.h, hidx));
  }
  const int ntr = a.ntr;
  if (ntr <= 0) return;
  const bool td2 = terms & T_TDEL2, td4 = terms & T_TDEL4;
  Vd<N> invC, invN[6];
  double idc[6], c2[6], s2[6], c4[6];
#pragma unroll
  for (int i = 0; i < N; ++i) invC[i] = 1.0 / hC[i];
#pragma unroll
  for (int j = 0; j < 6; ++j) {
#pragma unroll
    for (int i = 0; i < N; ++i) invN[j][i] = 1.0 / hN[j][i];
    idc[j] = m.inv_dc[E[j]];
    c2[j] = td2 ? mc.ct2[c * me + j] : 0.0;
    s2[j] = td2 ? m.s2[E[j]] : 0.0;
    c4[j] = td4 ? mc.ct4[c * me + j] : 0.0;
  }
  for (int l = 0; l < ntr; ++l) {
    const double *trl = a.tr.in + (size_t)l * NC * L;
    const Vd<N> tC = rowv<N>(trl, c, L, k);
    Vd<N> tN[6], dN[6], dC = zero<N>();
#pragma unroll
    for (int j = 0; j < 6; ++j) tN[j] = rowv<N>(trl, NB[j], L, k);
    if (td4) {
      const double *dl = d.trdel2_c + (size_t)l * NC * L;
      dC = rowv<N>(dl, c, L, k);
#pragma unroll
      for (int j = 0; j < 6; ++j) dN[j] = rowv<N>(dl, NB[j], L, k);
    }
    const size_t tidx = ((size_t)l * NC + c) * L + k;
    Vd<N> T = tend_base<N>(a, a.tr, tidx);
#pragma unroll
    for (int i = 0; i < N; ++i) {
      const double nC = tC[i] * invC[i];
      double adv = 0.0, dif = 0.0, hyp = 0.0;
#pragma unroll
      for (int j = 0; j < 6; ++j) {
        const double nN = tN[j][i] * invN[j][i];
        const bool second = hex::cell_is_second(j);
        adv = adv - cd[j] * fh[j][i] * ((nN + nC) * 0.5) * vE[j][i];
        if (td2) {
          const double grad = (second ? (nC - nN) : (nN - nC)) * idc[j];
          dif = dif + c2[j] * fh[j][i] * s2[j] * grad;
        }
        if (td4) {
          const double grad = (second ? (dC[i] - dN[j][i]) : (dN[j][i] - dC[i])) * idc[j];
          hyp = hyp - c4[j] * grad;
        }
      }
      T[i] = T[i] + adv;
      if (td2) T[i] = T[i] + dif;
      if (td4) T[i] = T[i] + hyp;
    }
    epilogue<N>(a, a.tr, tidx, T);
  }
}

// ---------------------------------------------------------------------------
// Two cells per wavefront with wavefront-uniform tables (N = 2 layout)
// ---------------------------------------------------------------------------
// In the N = 2 layout lanes 0-31 of a wavefront carry cell cA and lanes 32-63
// cell cB = cA + 1 (both at the same 64 levels).  Indices and coefficients
// are then uniform per half-wavefront: the variants below read the tables of
// both cells at wavefront-uniform addresses (scalar loads) and pick per lane,
// instead of issuing one vector load per table entry.  Same expressions as
// edge_tend_hex / cell_tend_hex, so the same bits.

// Read-only table entry through the constant address space: the tables are
// never written while a kernel runs, and loads from that space at uniform
// addresses go through the scalar cache even after the kernel has stored
// to other global memory.
template <class T>
__device__ __forceinline__ T cld(const T *p, long i) {
  return ((const __attribute__((address_space(4))) T *)p)[i];
}

// both entries read at uniform offsets, then the lane's one picked
template <class T>
__device__ __forceinline__ T pick2(const T *p, long iA, long iB, bool hi) {
  const T a = cld(p, iA);
  const T b = cld(p, iB);
  return hi ? b : a;
}

// A coefficient of the pair's two entities.  U: the coefficient tables are
// the same for every cell of the mesh at each stencil slot (checked bit for
// bit on the host, ModelCoef::uniform), so A's entry is B's as well.
template <bool U, class T>
__device__ __forceinline__ T cpick(const T *p, long iA, long iB, bool hi) {
  if (U) return cld(p, iA);
  return pick2(p, iA, iB, hi);
}

template <int K, bool U, bool W>
__device__ __forceinline__ void edge_tend_hex2(const MeshView &m, const ModelCoef &mc,
                                               const StageArgs &a, const DiagOut &d, int eA,
                                               int eB, bool hi, int k) {
  constexpr int N = 2;
  const int L = m.L;
  const int e = hi ? eB : eA;
  const size_t idx = (size_t)e * L + k;
  if (!a.do_vn) {
    if (a.epi == EPI_STORE) return;
    epilogue<N>(a, a.vn, idx, tend_base<N>(a, a.vn, idx));
    return;
  }
  const double *__restrict__ h = a.h.in;
  const double *__restrict__ vn = a.vn.in;
  const int *rA = m.hexrec + (size_t)eA * hex::kRecInts;
  const int *rB = m.hexrec + (size_t)eB * hex::kRecInts;
  int ej[10], uv[10], uc[10], uvA[10], uvB[10];
#pragma unroll
  for (int j = 0; j < 10; ++j) {
    const int ea = cld(rA, j), eb = cld(rB, j), ca = cld(rA, 20 + j), cb = cld(rB, 20 + j);
    uvA[j] = cld(rA, 10 + j);
    uvB[j] = cld(rB, 10 + j);
    ej[j] = hi ? eb : ea;
    uv[j] = hi ? uvB[j] : uvA[j];
    uc[j] = hi ? cb : ca;
  }
  constexpr int o0 = hex::OWNV[K][0], o1 = hex::OWNV[K][1];
  constexpr int q0 = hex::OWNC[K][0], q1 = hex::OWNC[K][1];
  const bool pv = a.terms & T_PV;
  Vd<N> hU[10], rvU[10], itU[10], vJ[10];
#pragma unroll
  for (int u = 0; u < 10; ++u) hU[u] = rowv<N, W>(h, uc[u], L, k);
  if (pv) {
#pragma unroll
    for (int j = 0; j < 10; ++j) vJ[j] = rowv<N, W>(vn, ej[j], L, k);
#pragma unroll
    for (int u = 0; u < 10; ++u) {
      rvU[u] = rowv<N, W>(d.rvort_v, uv[u], L, k);
      itU[u] = rowv<N, W>(d.inv_thick_v, uv[u], L, k);
    }
  } else if (a.terms & T_VDEL2) {
    rvU[o0] = rowv<N, W>(d.rvort_v, uv[o0], L, k);
    rvU[o1] = rowv<N, W>(d.rvort_v, uv[o1], L, k);
  }
  const int c0 = uc[q0], c1 = uc[q1];
  Vd<N> T = tend_base<N>(a, a.vn, idx);
  const double idc = cpick<U>(m.inv_dc, eA, eB, hi);
  if (pv) {
    Vd<N> nrvU[10], npvU[10];
#pragma unroll
    for (int u = 0; u < 10; ++u) {
      const double fa = cld(mc.f_v, uvA[u]);
      const double fb = cld(mc.f_v, uvB[u]);
      const double f = U ? fa : (hi ? fb : fa);
#pragma unroll
      for (int i = 0; i < N; ++i) {
        nrvU[u][i] = itU[u][i] * rvU[u][i];
        npvU[u][i] = itU[u][i] * f;
      }
    }
    double w[10];
#pragma unroll
    for (int j = 0; j < 10; ++j) w[j] = cpick<U>(m.w, (long)eA * 10 + j, (long)eB * 10 + j, hi);
#pragma unroll
    for (int i = 0; i < N; ++i) {
      const double qs = 0.5 * (nrvU[o0][i] + nrvU[o1][i]) + 0.5 * (npvU[o0][i] + npvU[o1][i]);
      double qt = 0.0;
#pragma unroll
      for (int j = 0; j < 10; ++j) {
        const double fh = 0.5 * (hU[hex::LC0[K][j]][i] + hU[hex::LC1[K][j]][i]);
        const double nrv = 0.5 * (nrvU[hex::LV0[K][j]][i] + nrvU[hex::LV1[K][j]][i]);
        const double npv = 0.5 * (npvU[hex::LV0[K][j]][i] + npvU[hex::LV1[K][j]][i]);
        const double nv = (qs + nrv + npv) * 0.5;
        qt = qt + w[j] * fh * vJ[j][i] * nv;
      }
      T[i] = T[i] + qt;
    }
  }
  if (a.terms & T_KE) {
    const Vd<N> k0 = rowv<N, W>(d.ke_c, c0, L, k), k1 = rowv<N, W>(d.ke_c, c1, L, k);
#pragma unroll
    for (int i = 0; i < N; ++i) T[i] = T[i] - (k1[i] - k0[i]) * idc;
  }
  if (a.terms & T_SSH) {
#pragma unroll
    for (int i = 0; i < N; ++i) T[i] = T[i] - mc.grav * (hU[q1][i] - hU[q0][i]) * idc;
  }
  if (a.terms & (T_VDEL2 | T_VDEL4)) {
    const int ml = cpick<U>(m.maxlev_e, eA, eB, hi);
    const double idv = cpick<U>(m.inv_dv, eA, eB, hi);
    if (a.terms & T_VDEL2) {
      const Vd<N> d0 = rowv<N, W>(d.div_c, c0, L, k), d1 = rowv<N, W>(d.div_c, c1, L, k);
      const double cv = cpick<U>(mc.cv2, eA, eB, hi);
#pragma unroll
      for (int i = 0; i < N; ++i) {
        const double del2u = (d1[i] - d0[i]) * idc - (rvU[o1][i] - rvU[o0][i]) * idv;
        T[i] = T[i] + ((k + i < ml) ? cv : 0.0) * del2u;
      }
    }
    if (a.terms & T_VDEL4) {
      const int b0 = uv[o0], b1 = uv[o1];
      const Vd<N> d0 = rowv<N, W>(d.vdel2_div_c, c0, L, k), d1 = rowv<N, W>(d.vdel2_div_c, c1, L, k);
      const Vd<N> r0 = rowv<N, W>(d.vdel2_rv_v, b0, L, k), r1 = rowv<N, W>(d.vdel2_rv_v, b1, L, k);
      const double cv = cpick<U>(mc.cv4, eA, eB, hi);
#pragma unroll
      for (int i = 0; i < N; ++i) {
        const double del2u = (d1[i] - d0[i]) * idc - (r1[i] - r0[i]) * idv;
        T[i] = T[i] - ((k + i < ml) ? cv : 0.0) * del2u;
      }
    }
  }
  epilogue<N>(a, a.vn, idx, T);
}

// Pair forms of vertex_diag_u / cell_diag_u (N = 2, two cells per
// wavefront, tables through the scalar cache); same expressions.
template <bool U, bool W>
__device__ __forceinline__ void vertex_diag_p(const MeshView &m, int vA, int vB, bool hi, int k,
                                              const double *__restrict__ h,
                                              const double *__restrict__ vn, const DiagOut &d) {
  constexpr int N = 2;
  const int L = m.L;
  const int v = hi ? vB : vA;
  int e[3], c[3];
  double crv[3], cth[3];
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    e[j] = pick2(m.eov, 3L * vA + j, 3L * vB + j, hi);
    c[j] = pick2(m.cov, 3L * vA + j, 3L * vB + j, hi);
    crv[j] = cpick<U>(m.crv, 3L * vA + j, 3L * vB + j, hi);
    cth[j] = cpick<U>(m.cth, 3L * vA + j, 3L * vB + j, hi);
  }
  Vd<N> ve[3], hc[3];
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    ve[j] = rowv<N, W>(vn, e[j], L, k);
    hc[j] = rowv<N, W>(h, c[j], L, k);
  }
  Vd<N> rv = zero<N>(), th = zero<N>();
#pragma unroll
  for (int j = 0; j < 3; ++j)
#pragma unroll
    for (int i = 0; i < N; ++i) {
      rv[i] = rv[i] + crv[j] * ve[j][i];
      th[i] = th[i] + cth[j] * hc[j][i];
    }
  Vd<N> inv;
#pragma unroll
  for (int i = 0; i < N; ++i) inv[i] = 1.0 / th[i];
  st<N>(d.rvort_v + (size_t)v * L + k, rv);
  st<N>(d.inv_thick_v + (size_t)v * L + k, inv);
}

template <bool U, bool W>
__device__ __forceinline__ void cell_diag_p(const MeshView &m, int terms, int ntr, int cA, int cB,
                                            bool hi, int k, const double *__restrict__ h,
                                            const double *__restrict__ vn,
                                            const double *__restrict__ tr, const DiagOut &d) {
  constexpr int N = 2;
  const int L = m.L, me = m.me;
  const int c = hi ? cB : cA;
  int E[6], EA[6], EB[6];
#pragma unroll
  for (int j = 0; j < 6; ++j) {
    EA[j] = cld(m.eoc, (long)cA * me + j);
    EB[j] = cld(m.eoc, (long)cB * me + j);
    E[j] = hi ? EB[j] : EA[j];
  }
  if (terms & (T_KE | T_VDEL2 | T_VDEL4)) {
    Vd<N> ve[6];
#pragma unroll
    for (int j = 0; j < 6; ++j) ve[j] = rowv<N, W>(vn, E[j], L, k);
    Vd<N> ke = zero<N>(), div = zero<N>();
#pragma unroll
    for (int j = 0; j < 6; ++j) {
      const double cke = cpick<U>(m.cke, (long)cA * me + j, (long)cB * me + j, hi);
      const double cd = cpick<U>(m.cdiv, (long)cA * me + j, (long)cB * me + j, hi);
#pragma unroll
      for (int i = 0; i < N; ++i) {
        ke[i] = ke[i] + cke * ve[j][i] * ve[j][i] * 0.25;
        div[i] = div[i] + cd * ve[j][i];
      }
    }
    st<N>(d.ke_c + (size_t)c * L + k, ke);
    st<N>(d.div_c + (size_t)c * L + k, div);
  }
  if ((terms & T_TDEL4) && ntr > 0) {
    int NB[6];
#pragma unroll
    for (int j = 0; j < 6; ++j) NB[j] = pick2(m.coc, (long)cA * 6 + j, (long)cB * 6 + j, hi);
    const Vd<N> hC = rowv<N, W>(h, c, L, k);
    Vd<N> hN[6], invN[6], cm[6], invC;
    double idc[6];
#pragma unroll
    for (int j = 0; j < 6; ++j) hN[j] = rowv<N, W>(h, NB[j], L, k);
#pragma unroll
    for (int i = 0; i < N; ++i) invC[i] = 1.0 / hC[i];
#pragma unroll
    for (int j = 0; j < 6; ++j) {
      const double cd = cpick<U>(m.cdiv, (long)cA * me + j, (long)cB * me + j, hi);
      idc[j] = cpick<U>(m.inv_dc, EA[j], EB[j], hi);
#pragma unroll
      for (int i = 0; i < N; ++i) {
        invN[j][i] = 1.0 / hN[j][i];
        cm[j][i] = cd * (0.5 * (hN[j][i] + hC[i]));
      }
    }
    for (int l = 0; l < ntr; ++l) {
      const double *trl = tr + (size_t)l * m.nc * L;
      const Vd<N> tC = rowv<N, W>(trl, c, L, k);
      Vd<N> tN[6];
#pragma unroll
      for (int j = 0; j < 6; ++j) tN[j] = rowv<N, W>(trl, NB[j], L, k);
      Vd<N> d2 = zero<N>();
#pragma unroll
      for (int j = 0; j < 6; ++j)
#pragma unroll
        for (int i = 0; i < N; ++i) {
          const double nC = tC[i] * invC[i];
          const double nN = tN[j][i] * invN[j][i];
          const double grad = (hex::cell_is_second(j) ? (nC - nN) : (nN - nC)) * idc[j];
          d2[i] = d2[i] + cm[j][i] * grad;
        }
      st<N>(d.trdel2_c + ((size_t)l * m.nc + c) * L + k, d2);
    }
  }
}

// Momentum del4 pre-pass in the pair layout (vel_hyperdiff_edge.hpp:140-210;
// the general kernels' vdel4_edge / vdel4_vertex / vdel4_cell): del2 of the
// velocity on an owned edge, then its curl at a vertex and divergence at a
// cell, with the tables read at wavefront-uniform addresses.  Same
// expressions and summation order, so the same bits.
template <bool U, bool W>
__device__ __forceinline__ void vdel4_edge_p(const MeshView &m, const DiagOut &d, int eA, int eB,
                                             bool hi, int k) {
  constexpr int N = 2;
  const int L = m.L;
  const int e = hi ? eB : eA;
  const int a0 = pick2(m.coe, 2L * eA, 2L * eB, hi), a1 = pick2(m.coe, 2L * eA + 1, 2L * eB + 1, hi);
  const int b0 = pick2(m.voe, 2L * eA, 2L * eB, hi), b1 = pick2(m.voe, 2L * eA + 1, 2L * eB + 1, hi);
  const double idc = cpick<U>(m.inv_dc, eA, eB, hi), idv = cpick<U>(m.inv_dv4, eA, eB, hi);
  const Vd<N> d0 = rowv<N, W>(d.div_c, a0, L, k), d1 = rowv<N, W>(d.div_c, a1, L, k);
  const Vd<N> r0 = rowv<N, W>(d.rvort_v, b0, L, k), r1 = rowv<N, W>(d.rvort_v, b1, L, k);
  Vd<N> out;
#pragma unroll
  for (int i = 0; i < N; ++i) out[i] = (d1[i] - d0[i]) * idc - (r1[i] - r0[i]) * idv;
  st<N>(d.vdel2_e + (size_t)e * L + k, out);
}

template <bool U, bool W>
__device__ __forceinline__ void vdel4_vertex_p(const MeshView &m, const DiagOut &d, int vA, int vB,
                                               bool hi, int k) {
  constexpr int N = 2;
  const int L = m.L;
  const int v = hi ? vB : vA;
  int e[3];
  double crv[3];
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    e[j] = pick2(m.eov, 3L * vA + j, 3L * vB + j, hi);
    crv[j] = cpick<U>(m.crv, 3L * vA + j, 3L * vB + j, hi);
  }
  Vd<N> x[3];
#pragma unroll
  for (int j = 0; j < 3; ++j) x[j] = rowv<N, W>(d.vdel2_e, e[j], L, k);
  Vd<N> rv = zero<N>();
#pragma unroll
  for (int j = 0; j < 3; ++j)
#pragma unroll
    for (int i = 0; i < N; ++i) rv[i] = rv[i] + crv[j] * x[j][i];
  st<N>(d.vdel2_rv_v + (size_t)v * L + k, rv);
}

template <bool U, bool W>
__device__ __forceinline__ void vdel4_cell_p(const MeshView &m, const DiagOut &d, int cA, int cB,
                                             bool hi, int k) {
  constexpr int N = 2;
  const int L = m.L, me = m.me;
  const int c = hi ? cB : cA;
  int e[6];
  double cd[6];
#pragma unroll
  for (int j = 0; j < 6; ++j) {
    e[j] = pick2(m.eoc, (long)cA * me + j, (long)cB * me + j, hi);
    cd[j] = cpick<U>(m.cdiv, (long)cA * me + j, (long)cB * me + j, hi);
  }
  Vd<N> x[6];
#pragma unroll
  for (int j = 0; j < 6; ++j) x[j] = rowv<N, W>(d.vdel2_e, e[j], L, k);
  Vd<N> acc = zero<N>();
#pragma unroll
  for (int j = 0; j < 6; ++j)
#pragma unroll
    for (int i = 0; i < N; ++i) acc[i] = acc[i] + cd[j] * x[j][i];
  st<N>(d.vdel2_div_c + (size_t)c * L + k, acc);
}

// Union form of vertex_diag_p x 2 + cell_diag_p: the rows of the cell's
// six edges and seven cells (itself and its neighbours) plus the two other
// edges of its owned vertices are loaded once and shared by the vertex,
// kinetic-energy/divergence and tracer-Laplacian parts (hexd maps).  Same
// expressions, same bits.
template <bool U, bool W>
__device__ __forceinline__ void diag_union_p(const MeshView &m, int terms, int ntr, int cA,
                                             int cB, bool hi, int k, const double *__restrict__ h,
                                             const double *__restrict__ vn,
                                             const double *__restrict__ tr, const DiagOut &d) {
  constexpr int N = 2;
  const int L = m.L, me = m.me;
  const int c = hi ? cB : cA;
  const bool vert = terms & (T_PV | T_VDEL2 | T_VDEL4);
  const bool kediv = terms & (T_KE | T_VDEL2 | T_VDEL4);
  const bool del2 = (terms & T_TDEL4) && ntr > 0;
  int EA[6], EB[6], E[8], NB[6];
#pragma unroll
  for (int j = 0; j < 6; ++j) {
    EA[j] = cld(m.eoc, (long)cA * me + j);
    EB[j] = cld(m.eoc, (long)cB * me + j);
    E[j] = hi ? EB[j] : EA[j];
    NB[j] = pick2(m.coc, (long)cA * 6 + j, (long)cB * 6 + j, hi);
  }
  // the two other edges of the owned vertices 2c (slot 2) and 2c+1 (slot 1)
  E[6] = pick2(m.eov, 6L * cA + 2, 6L * cB + 2, hi);
  E[7] = pick2(m.eov, 6L * cA + 4, 6L * cB + 4, hi);
  Vd<N> ve[8], hR[7];
  if (vert || kediv) {
#pragma unroll
    for (int j = 0; j < 6; ++j) ve[j] = rowv<N, W>(vn, E[j], L, k);
  }
  if (vert) {
    ve[6] = rowv<N, W>(vn, E[6], L, k);
    ve[7] = rowv<N, W>(vn, E[7], L, k);
  }
  if (vert || del2) {
    hR[0] = rowv<N, W>(h, c, L, k);
#pragma unroll
    for (int j = 0; j < 6; ++j) hR[1 + j] = rowv<N, W>(h, NB[j], L, k);
  }
  if (vert) {
    // VorticityAuxVars::compute_vort_vertex (vorticity_aux.hpp:85-116)
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      const int vA = 2 * cA + i, vB = 2 * cB + i;
      const int v = hi ? vB : vA;
      double crv[3], cth[3];
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        crv[j] = cpick<U>(m.crv, 3L * vA + j, 3L * vB + j, hi);
        cth[j] = cpick<U>(m.cth, 3L * vA + j, 3L * vB + j, hi);
      }
      Vd<N> rv = zero<N>(), th = zero<N>();
#pragma unroll
      for (int j = 0; j < 3; ++j)
#pragma unroll
        for (int q = 0; q < N; ++q) {
          rv[q] = rv[q] + crv[j] * ve[hexd::VE[i][j]][q];
          th[q] = th[q] + cth[j] * hR[hexd::VC[i][j]][q];
        }
      Vd<N> inv;
#pragma unroll
      for (int q = 0; q < N; ++q) inv[q] = 1.0 / th[q];
      st<N>(d.rvort_v + (size_t)v * L + k, rv);
      st<N>(d.inv_thick_v + (size_t)v * L + k, inv);
    }
  }
  if (kediv) {
    // KineticAuxVars::compute_kinetic_cell (kinetic_aux.hpp:44-73)
    Vd<N> ke = zero<N>(), div = zero<N>();
#pragma unroll
    for (int j = 0; j < 6; ++j) {
      const double cke = cpick<U>(m.cke, (long)cA * me + j, (long)cB * me + j, hi);
      const double cd = cpick<U>(m.cdiv, (long)cA * me + j, (long)cB * me + j, hi);
#pragma unroll
      for (int q = 0; q < N; ++q) {
        ke[q] = ke[q] + cke * ve[j][q] * ve[j][q] * 0.25;
        div[q] = div[q] + cd * ve[j][q];
      }
    }
    st<N>(d.ke_c + (size_t)c * L + k, ke);
    st<N>(d.div_c + (size_t)c * L + k, div);
  }
  if (del2) {
    // TracerHyperDiffusionOnCell::compute_tracer_del2
    // (tracer_hyperdiff_cell.hpp:94-129)
    Vd<N> invN[6], cm[6], invC;
    double idc[6];
#pragma unroll
    for (int q = 0; q < N; ++q) invC[q] = 1.0 / hR[0][q];
#pragma unroll
    for (int j = 0; j < 6; ++j) {
      const double cd = cpick<U>(m.cdiv, (long)cA * me + j, (long)cB * me + j, hi);
      idc[j] = cpick<U>(m.inv_dc, EA[j], EB[j], hi);
#pragma unroll
      for (int q = 0; q < N; ++q) {
        invN[j][q] = 1.0 / hR[1 + j][q];
        cm[j][q] = cd * (0.5 * (hR[1 + j][q] + hR[0][q]));
      }
    }
    for (int l = 0; l < ntr; ++l) {
      const double *trl = tr + (size_t)l * m.nc * L;
      const Vd<N> tC = rowv<N, W>(trl, c, L, k);
      Vd<N> tN[6];
#pragma unroll
      for (int j = 0; j < 6; ++j) tN[j] = rowv<N, W>(trl, NB[j], L, k);
      Vd<N> d2 = zero<N>();
#pragma unroll
      for (int j = 0; j < 6; ++j)
#pragma unroll
        for (int q = 0; q < N; ++q) {
          const double nC = tC[q] * invC[q];
          const double nN = tN[j][q] * invN[j][q];
          const double grad = (hex::cell_is_second(j) ? (nC - nN) : (nN - nC)) * idc[j];
          d2[q] = d2[q] + cm[j][q] * grad;
        }
      st<N>(d.trdel2_c + ((size_t)l * m.nc + c) * L + k, d2);
    }
  }
}

// Vertex-union form of edge_tend_hex2: the PV products of the cell's 16
// union vertices (hex_pattern.hpp, namespace hexu) are formed once per cell
// by the caller and shared by its three edges; the rows per edge are then
// 10 thicknesses and 10 velocities.  Same values, same expressions.  EM: also
// write the edge's PreOut rows, as edge_tend_hex does.
template <int K, bool U, bool W, int TM = -1, int EK = -1, bool EM = false>
__device__ __forceinline__ void edge_tend_hex2v(const MeshView &m, const ModelCoef &mc,
                                                const StageArgs &a, const DiagOut &d, int eA,
                                                int eB, bool hi, int k, const Vd<2> (&P)[hexu::kNV],
                                                const Vd<2> (&Q)[hexu::kNV],
                                                const PreOut &po = PreOut{}) {
  constexpr int N = 2;
  const int L = m.L;
  // TM >= 0: the term set is TM and both tendencies are on (compiled in)
  const int terms = TM >= 0 ? TM : a.terms;
  const bool do_vn = TM >= 0 ? true : a.do_vn != 0;
  const int e = hi ? eB : eA;
  const size_t idx = (size_t)e * L + k;
  if (!do_vn) {
    if (a.epi == EPI_STORE) return;
    epilogue<N, EK>(a, a.vn, idx, tend_base<N>(a, a.vn, idx));
    return;
  }
  const double *__restrict__ h = a.h.in;
  const double *__restrict__ vn = a.vn.in;
  const int *rA = m.hexrec + (size_t)eA * hex::kRecInts;
  const int *rB = m.hexrec + (size_t)eB * hex::kRecInts;
  int ej[10], uc[10];
#pragma unroll
  for (int j = 0; j < 10; ++j) {
    const int ea = cld(rA, j), eb = cld(rB, j), ca = cld(rA, 20 + j), cb = cld(rB, 20 + j);
    ej[j] = hi ? eb : ea;
    uc[j] = hi ? cb : ca;
  }
  constexpr int o0 = hex::OWNV[K][0], o1 = hex::OWNV[K][1];
  constexpr int q0 = hex::OWNC[K][0], q1 = hex::OWNC[K][1];
  constexpr int v0 = hexu::OV[K][0], v1 = hexu::OV[K][1];
  const bool pv = terms & T_PV;
  Vd<N> hU[10], vJ[10];
#pragma unroll
  for (int u = 0; u < 10; ++u) hU[u] = rowv<N, W>(h, uc[u], L, k);
  if (pv) {
#pragma unroll
    for (int j = 0; j < 10; ++j) vJ[j] = rowv<N, W>(vn, ej[j], L, k);
  }
  const int c0 = uc[q0], c1 = uc[q1];
  Vd<N> T = tend_base<N>(a, a.vn, idx);
  const double idc = cpick<U>(m.inv_dc, eA, eB, hi);
  if (pv) {
    double w[10];
#pragma unroll
    for (int j = 0; j < 10; ++j) w[j] = cpick<U>(m.w, (long)eA * 10 + j, (long)eB * 10 + j, hi);
    Vd<N> S, O;
#pragma unroll
    for (int i = 0; i < N; ++i) {
      const double qs = 0.5 * (P[v0][i] + P[v1][i]) + 0.5 * (Q[v0][i] + Q[v1][i]);
      double qt = 0.0, ow = 0.0;
#pragma unroll
      for (int j = 0; j < 10; ++j) {
        const double fh = 0.5 * (hU[hex::LC0[K][j]][i] + hU[hex::LC1[K][j]][i]);
        const double nrv = 0.5 * (P[hexu::TV0[K][j]][i] + P[hexu::TV1[K][j]][i]);
        const double npv = 0.5 * (Q[hexu::TV0[K][j]][i] + Q[hexu::TV1[K][j]][i]);
        const double nv = (qs + nrv + npv) * 0.5;
        qt = qt + w[j] * fh * vJ[j][i] * nv;
        if constexpr (EM) ow = ow + w[j] * (fh * vJ[j][i]);
      }
      T[i] = T[i] + qt;
      S[i] = qs;
      O[i] = ow;
    }
    if constexpr (EM) {
      st<N>(po.se + idx, S);
      st<N>(po.own + idx, O);
    }
  }
  if constexpr (EM) {
    const Vd<N> ue = rowv<N, W>(vn, e, L, k);
    Vd<N> FU;
#pragma unroll
    for (int i = 0; i < N; ++i) FU[i] = 0.5 * (hU[q0][i] + hU[q1][i]) * ue[i];
    st<N>(po.fu + idx, FU);
  }
  if (terms & T_KE) {
    const Vd<N> k0 = rowv<N, W>(d.ke_c, c0, L, k), k1 = rowv<N, W>(d.ke_c, c1, L, k);
#pragma unroll
    for (int i = 0; i < N; ++i) T[i] = T[i] - (k1[i] - k0[i]) * idc;
  }
  if (terms & T_SSH) {
#pragma unroll
    for (int i = 0; i < N; ++i) T[i] = T[i] - mc.grav * (hU[q1][i] - hU[q0][i]) * idc;
  }
  if (terms & (T_VDEL2 | T_VDEL4)) {
    const int ml = cpick<U>(m.maxlev_e, eA, eB, hi);
    const double idv = cpick<U>(m.inv_dv, eA, eB, hi);
    const int b0 = pick2(rA, 10 + o0, (long)(rB - rA) + 10 + o0, hi);
    const int b1 = pick2(rA, 10 + o1, (long)(rB - rA) + 10 + o1, hi);
    if (terms & T_VDEL2) {
      const Vd<N> d0 = rowv<N, W>(d.div_c, c0, L, k), d1 = rowv<N, W>(d.div_c, c1, L, k);
      const Vd<N> r0 = rowv<N, W>(d.rvort_v, b0, L, k), r1 = rowv<N, W>(d.rvort_v, b1, L, k);
      const double cv = cpick<U>(mc.cv2, eA, eB, hi);
#pragma unroll
      for (int i = 0; i < N; ++i) {
        const double del2u = (d1[i] - d0[i]) * idc - (r1[i] - r0[i]) * idv;
        T[i] = T[i] + ((k + i < ml) ? cv : 0.0) * del2u;
      }
    }
    if (terms & T_VDEL4) {
      const Vd<N> d0 = rowv<N, W>(d.vdel2_div_c, c0, L, k), d1 = rowv<N, W>(d.vdel2_div_c, c1, L, k);
      const Vd<N> r0 = rowv<N, W>(d.vdel2_rv_v, b0, L, k), r1 = rowv<N, W>(d.vdel2_rv_v, b1, L, k);
      const double cv = cpick<U>(mc.cv4, eA, eB, hi);
#pragma unroll
      for (int i = 0; i < N; ++i) {
        const double del2u = (d1[i] - d0[i]) * idc - (r1[i] - r0[i]) * idv;
        T[i] = T[i] - ((k + i < ml) ? cv : 0.0) * del2u;
      }
    }
  }
  epilogue<N, EK>(a, a.vn, idx, T);
}

// PV products inv_thick*rel_vort and inv_thick*f of the 16 union vertices
// of the wavefront's two cells (lane's cell picked per half-wavefront).
template <bool U, bool W>
__device__ __forceinline__ void union_pv2(const MeshView &m, const ModelCoef &mc,
                                          const DiagOut &d, int cA, int cB, bool hi, int k,
                                          Vd<2> (&P)[hexu::kNV], Vd<2> (&Q)[hexu::kNV]) {
  const int L = m.L;
  const int *uA = m.hexu + (size_t)cA * hexu::kRec + hexu::kNE;
  const int *uB = m.hexu + (size_t)cB * hexu::kRec + hexu::kNE;
#pragma unroll
  for (int x = 0; x < hexu::kNV; ++x) {
    const int va = cld(uA, x), vb = cld(uB, x);
    const int v = hi ? vb : va;
    // U: f is one value for the whole mesh
    const double fa = cld(mc.f_v, U ? 0 : va), fb = cld(mc.f_v, vb);
    const double f = U ? fa : (hi ? fb : fa);
    const Vd<2> rv = rowv<2, W>(d.rvort_v, v, L, k), it = rowv<2, W>(d.inv_thick_v, v, L, k);
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      P[x][i] = it[i] * rv[i];
      Q[x][i] = it[i] * f;
    }
  }
}

template <bool U, bool W, int TM = -1, int EK = -1>
__device__ __forceinline__ void cell_tend_hex2(const MeshView &m, const ModelCoef &mc,
                                               const StageArgs &a, const DiagOut &d, int cA,
                                               int cB, bool hi, int k) {
  constexpr int N = 2;
  const int L = m.L, me = m.me, NC = m.nc;
  const int c = hi ? cB : cA;
  const double *__restrict__ h = a.h.in;
  const double *__restrict__ vn = a.vn.in;
  int E[6], NB[6], EA[6], EB[6];
#pragma unroll
  for (int j = 0; j < 6; ++j) {
    EA[j] = cld(m.eoc, (long)cA * me + j);
    EB[j] = cld(m.eoc, (long)cB * me + j);
    E[j] = hi ? EB[j] : EA[j];
    NB[j] = pick2(m.coc, (long)cA * 6 + j, (long)cB * 6 + j, hi);
  }
  const Vd<N> hC = rowv<N, W>(h, c, L, k);
  Vd<N> hN[6], vE[6], fh[6];
  double cd[6];
#pragma unroll
  for (int j = 0; j < 6; ++j) {
    hN[j] = rowv<N, W>(h, NB[j], L, k);
    vE[j] = rowv<N, W>(vn, E[j], L, k);
  }
#pragma unroll
  for (int j = 0; j < 6; ++j) {
    cd[j] = cpick<U>(m.cdiv, (long)cA * me + j, (long)cB * me + j, hi);
#pragma unroll
    for (int i = 0; i < N; ++i) fh[j][i] = 0.5 * (hN[j][i] + hC[i]);
  }
  const size_t hidx = (size_t)c * L + k;
  if (TM >= 0 || a.do_h) {
    Vd<N> T = tend_base<N>(a, a.h, hidx);
#pragma unroll
    for (int i = 0; i < N; ++i) {
      double acc = 0.0;
#pragma unroll
      for (int j = 0; j < 6; ++j) acc = acc - cd[j] * fh[j][i] * vE[j][i];
      T[i] = T[i] + acc;
    }
    epilogue<N, EK>(a, a.h, hidx, T);
  } else if (a.epi != EPI_STORE) {
    epilogue<N, EK>(a, a.h, hidx, tend_base<N>(a, a.h, hidx));
  }
  const int ntr = a.ntr;
  if (ntr <= 0) return;
  const int terms = TM >= 0 ? TM : a.terms;
  const bool td2 = terms & T_TDEL2, td4 = terms & T_TDEL4;
  Vd<N> invC, invN[6];
  double idc[6], c2[6], s2[6], c4[6];
#pragma unroll
  for (int i = 0; i < N; ++i) invC[i] = 1.0 / hC[i];
#pragma unroll
  for (int j = 0; j < 6; ++j) {
#pragma unroll
    for (int i = 0; i < N; ++i) invN[j][i] = 1.0 / hN[j][i];
    idc[j] = cpick<U>(m.inv_dc, EA[j], EB[j], hi);
    c2[j] = td2 ? cpick<U>(mc.ct2, (long)cA * me + j, (long)cB * me + j, hi) : 0.0;
    s2[j] = td2 ? cpick<U>(m.s2, EA[j], EB[j], hi) : 0.0;
    c4[j] = td4 ? cpick<U>(mc.ct4, (long)cA * me + j, (long)cB * me + j, hi) : 0.0;
  }
  for (int l = 0; l < ntr; ++l) {
    const double *trl = a.tr.in + (size_t)l * NC * L;
    const Vd<N> tC = rowv<N, W>(trl, c, L, k);
    Vd<N> tN[6], dN[6], dC = zero<N>();
#pragma unroll
    for (int j = 0; j < 6; ++j) tN[j] = rowv<N, W>(trl, NB[j], L, k);
    if (td4) {
      const double *dl = d.trdel2_c + (size_t)l * NC * L;
      dC = rowv<N, W>(dl, c, L, k);
#pragma unroll
      for (int j = 0; j < 6; ++j) dN[j] = rowv<N, W>(dl, NB[j], L, k);
    }
    const size_t tidx = ((size_t)l * NC + c) * L + k;
    Vd<N> T = tend_base<N>(a, a.tr, tidx);
#pragma unroll
    for (int i = 0; i < N; ++i) {
      const double nC = tC[i] * invC[i];
      double adv = 0.0, dif = 0.0, hyp = 0.0;
#pragma unroll
      for (int j = 0; j < 6; ++j) {
        const double nN = tN[j][i] * invN[j][i];
        const bool second = hex::cell_is_second(j);
        adv = adv - cd[j] * fh[j][i] * ((nN + nC) * 0.5) * vE[j][i];
        if (td2) {
          const double grad = (second ? (nC - nN) : (nN - nC)) * idc[j];
          dif = dif + c2[j] * fh[j][i] * s2[j] * grad;
        }
        if (td4) {
          const double grad = (second ? (dC[i] - dN[j][i]) : (dN[j][i] - dC[i])) * idc[j];
          hyp = hyp - c4[j] * grad;
        }
      }
      T[i] = T[i] + adv;
      if (td2) T[i] = T[i] + dif;
      if (td4) T[i] = T[i] + hyp;
    }
    epilogue<N, EK>(a, a.tr, tidx, T);
  }
}

} // namespace hexops
} // namespace omega_hip
