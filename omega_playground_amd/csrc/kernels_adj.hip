// Discrete adjoint (vector-Jacobian product) of the tendency, gather-only.
//
// Given a state y and a cotangent k of the tendency F(y), these kernels add
// J(y)^T k to a state cotangent and accumulate the gradient with respect to
// the four mixing coefficients (visc_del2, visc_del4, eddy_diff2,
// eddy_diff4).  Every forward scatter (an edge quantity feeding its two
// cells, a vertex quantity feeding its three edges, an edge feeding the ten
// edges that list it in edges_on_edge) becomes a gather over the transposed
// connectivity, so each output element is written by exactly one thread,
// no atomics are used and the result is bit-reproducible.  The parameter
// gradients are reduced through fixed per-block partials.
//
// Notation (forward, reference file:line):
//   zeta_v = sum crv u            (vorticity_aux.hpp:94-106)
//   theta_v = sum cth h, iota = 1/theta, q = iota zeta, p = iota f
//   K_c = sum cke u u /4, D_c = sum cdiv u   (kinetic_aux.hpp:53-65)
//   F_e = (h_c0 + h_c1)/2, Q_e = (q_v0 + q_v1)/2, P_e likewise
//   nu_lc = tr_lc / h_c                      (tracer_aux.hpp:39-44)
//   G_lc = sum cdiv F (nu_c1 - nu_c0)/dc     (tracer_hyperdiff_cell.hpp:94-129)
//   h_t = -sum cdiv F u                      (thickness_hadv_cell.hpp:48-72)
//   u_t = sum W F_j u_j (Q_e+P_e+Q_j+P_j)/2 - dK/dc - g dh/dc
//         + cv2 [dD/dc - dzeta/dv] - cv4 [dD2/dc - dZ2/dv]
//                                  (pv_flux_edge.hpp, ke_grad_edge.hpp,
//                                   ssh_grad_edge.hpp, vel_diff_edge.hpp,
//                                   vel_hyperdiff_edge.hpp)
//   tr_t = -sum cdiv F nu_e u + sum ct2 F s2 dnu/dc - sum ct4 dG/dc
//                                  (tracer_hadv/diff/hyperdiff_cell.hpp)
#include <algorithm>

#include "adjoint.hpp"
#include "device_common.hpp"

namespace omega_hip {

namespace {

using namespace dev;

constexpr int kAB = 256; // threads per block (lanes x rows)

template <int N>
__device__ __forceinline__ Vd<N> rd(const double *__restrict__ p, int e, int L, int k) {
  return ld<N>(p + (size_t)e * L + k);
}

// A block's sum into its partial slot: stored, or scaled and stored, or
// scaled and added (AdjArgs::part_add / part_keep)
__device__ __forceinline__ void put_partial(const AdjArgs &g, double x, double *out) {
  if (!(g.part_add || g.part_keep)) *out = x;
  else if (g.part_add) *out = *out + g.epi.vs * x;
  else *out = g.epi.vs * x;
}

// Block-wide sum of two values into two partial slots (all threads call it).
__device__ void block_partials2(const AdjArgs &g, double a, double b, double *out_a,
                                double *out_b) {
  __shared__ double sa[kAB], sb[kAB];
  const int t = threadIdx.y * blockDim.x + threadIdx.x;
  const int n = blockDim.x * blockDim.y;
  sa[t] = a;
  sb[t] = b;
  __syncthreads();
  for (int s = kAB / 2; s > 0; s >>= 1) {
    if (t < s && t + s < n) {
      sa[t] = sa[t] + sa[t + s];
      sb[t] = sb[t] + sb[t + s];
    }
    __syncthreads();
  }
  if (t == 0) {
    put_partial(g, sa[0], out_a);
    put_partial(g, sb[0], out_b);
  }
}

// Block-wide sum of one value into one partial slot (all threads call it).
__device__ void block_partial(const AdjArgs &g, double a, double *out) {
  __shared__ double sa[kAB];
  const int t = threadIdx.y * blockDim.x + threadIdx.x;
  const int n = blockDim.x * blockDim.y;
  sa[t] = a;
  __syncthreads();
  for (int s = kAB / 2; s > 0; s >>= 1) {
    if (t < s && t + s < n) sa[t] = sa[t] + sa[t + s];
    __syncthreads();
  }
  if (t == 0) put_partial(g, sa[0], out);
}

struct Seg {
  int c0, c1, k;
  bool active;
};

// The grid is one-dimensional over (cell block, level chunk) pairs, mapped
// as in the forward kernels: with a multiple of 8 chunks the chunk runs
// fastest (XCD x handles chunks x, x + 8, ... of every cell block),
// otherwise slowest with the optional swizzle over cell blocks.
template <int N> __device__ __forceinline__ Seg segment(const MeshView &m, int swz) {
  Seg g;
  const int span = (int)blockDim.x * N;
  const int nch = (m.L + span - 1) / span;
  int s, ch;
  if (nch == 1) {
    s = logical_block(blockIdx.x, gridDim.x, swz);
    ch = 0;
  } else if (nch % 8 == 0) {
    s = __builtin_amdgcn_readfirstlane((int)blockIdx.x / nch);
    ch = (int)blockIdx.x - s * nch;
  } else {
    const int ncb = (int)gridDim.x / nch;
    const int lb = logical_block(blockIdx.x, gridDim.x, swz);
    ch = __builtin_amdgcn_readfirstlane(lb / ncb);
    s = lb - ch * ncb;
  }
  g.c0 = s * (int)blockDim.y;
  g.c1 = min(g.c0 + (int)blockDim.y, m.nc);
  g.k = (ch * (int)blockDim.x + (int)threadIdx.x) * N;
  g.active = g.c0 < m.nc && g.k < m.L;
  return g;
}

// out[idx] = base[idx] + vs*v (base defaults to out); for the RK4 reverse
// sweep also kbo[idx] = db*gy[idx] + da*vs*v, for the low-storage RK reverse
// sweep kbo[idx] = da*k[idx] + db*out[idx] (k: this VJP's input cotangent)
template <int N>
__device__ __forceinline__ void adj_out(const AdjArgs &a, int f, Real *out, size_t idx,
                                        const Vd<N> &v0) {
  Vd<N> v;
#pragma unroll
  for (int i = 0; i < N; ++i) v[i] = a.epi.vs * v0[i];
  Vd<N> acc = ld<N>((a.epi.base[f] ? a.epi.base[f] : out) + idx);
#pragma unroll
  for (int i = 0; i < N; ++i) acc[i] += v[i];
  st<N>(out + idx, acc);
  if (a.epi.rk4) {
    const Vd<N> g = ld<N>(a.epi.gy[f] + idx);
    Vd<N> kn;
#pragma unroll
    for (int i = 0; i < N; ++i) kn[i] = a.epi.db * g[i] + a.epi.da * v[i];
    st<N>(a.epi.kbo[f] + idx, kn);
  }
  if (a.epi.lsrk) {
    const Real *kf = f == 0 ? a.c.h : (f == 1 ? a.c.vn : a.c.tr);
    const Vd<N> kv = ld<N>(kf + idx);
    Vd<N> kn;
#pragma unroll
    for (int i = 0; i < N; ++i) kn[i] = a.epi.da * kv[i] + a.epi.db * acc[i];
    st<N>(a.epi.kbo[f] + idx, kn);
  }
}

template <int N>
__device__ __forceinline__ double mask(int k, int i, int ml, double c) {
  return (k + i < ml) ? c : 0.0;
}

// Entities a block row visits: the owned edges (vertices) of the block's
// cells, strided over the rows, so a wavefront's rows are consecutive edges
// (vertices) that share most of their neighbours' rows.  In the
// slot-uniform form U (regular hexagonal numbering: cell c owns edges
// 3c + K and vertices 2c + x) the callee also gets the edge's class K
// (vertex's x), which selects among per-class coefficients read once per
// wavefront.
template <int U, class F>
__device__ __forceinline__ void for_edges(const MeshView &m, const Seg &g, F &&f) {
  const int e0 = U ? 3 * g.c0 : m.eofs[g.c0], e1 = U ? 3 * g.c1 : m.eofs[g.c1];
  for (int e = e0 + (int)threadIdx.y; e < e1; e += blockDim.y) f(e, U ? e % 3 : 0);
}

template <int U, class F>
__device__ __forceinline__ void for_vertices(const MeshView &m, const Seg &g, F &&f) {
  const int v0 = U ? 2 * g.c0 : m.vofs[g.c0], v1 = U ? 2 * g.c1 : m.vofs[g.c1];
  for (int v = v0 + (int)threadIdx.y; v < v1; v += blockDim.y) f(v, U ? v & 1 : 0);
}

// Read-only table entry at a wavefront-uniform index, through the constant
// address space: the tables are never written while a kernel runs, and such
// loads go through the scalar cache even after the kernel's first stores.
template <class T> __device__ __forceinline__ T uld(const T *p, int i) {
  return ((const __attribute__((address_space(4))) T *)p)[i];
}
// per-edge coefficient (one value for the whole mesh under U)
template <int U, class T> __device__ __forceinline__ T ecoef(const T *t, int e) {
  if constexpr (U) return uld(t, 0);
  else return t[e];
}
// [ne][2] coefficient at slot s of edge e (repeats every 3 edges under U:
// the three classes' values are read once and the lane's one picked)
template <int U> __device__ __forceinline__ double eslot(const double *t, int e, int K, int s) {
  if constexpr (U) {
    const double a = uld(t, s), b = uld(t, 2 + s), c = uld(t, 4 + s);
    return K == 0 ? a : (K == 1 ? b : c);
  } else {
    return t[2 * e + s];
  }
}
// [rows][n] coefficient at slot j of row r (the same for every row under U)
template <int U> __device__ __forceinline__ double rslot(const double *t, int r, int n, int j) {
  if constexpr (U) return uld(t, j);
  else return t[(size_t)r * n + j];
}
// two consecutive ints (cells_on_edge, vertices_on_edge) in one 8-byte load
__device__ __forceinline__ int2 pair_at(const int *t, int e) {
  return *reinterpret_cast<const int2 *>(t + 2 * (size_t)e);
}
// n ints of a row starting at an 8-byte aligned offset, in 8-byte loads
template <int n> __device__ __forceinline__ void row_ints(const int *t, size_t off, int *out) {
  const int2 *p = reinterpret_cast<const int2 *>(t + off);
#pragma unroll
  for (int i = 0; i < n / 2; ++i) {
    const int2 v = p[i];
    out[2 * i] = v.x;
    out[2 * i + 1] = v.y;
  }
}

// ---------------------------------------------------------------------------
// Momentum del4 chain: cotangents of the del2 divergence / vorticity, then
// of the del2 edge field (vel_hyperdiff_edge.hpp:140-237 transposed)
// ---------------------------------------------------------------------------
template <int N, int H>
__global__ void __launch_bounds__(kAB)
    k_adj_v4a(MeshView m, ModelCoef mc, AdjArgs a, AdjScratch s, int swz) {
  const Seg g = segment<N>(m, swz);
  if (!g.active) return;
  const int L = m.L, k = g.k;
  for (int v = m.vofs[g.c0] + (int)threadIdx.y; v < m.vofs[g.c1]; v += blockDim.y) {
    Vd<N> z = zero<N>();
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      const int e = m.eov[3 * v + j];
      const double sg = (m.voe[2 * e + 1] == v) ? 1.0 : -1.0;
      const Vd<N> ut = rd<N>(a.c.vn, e, L, k);
      const int ml = m.maxlev_e[e];
#pragma unroll
      for (int i = 0; i < N; ++i) z[i] += sg * mask<N>(k, i, ml, mc.cv4[e]) * ut[i] * m.inv_dv[e];
    }
    st<N>(s.z2bar_v + (size_t)v * L + k, z);
  }
  for (int c = g.c0 + (int)threadIdx.y; c < g.c1; c += blockDim.y) {
    Vd<N> d2 = zero<N>();
    const int n = H ? 6 : m.noc[c];
#pragma unroll 6
    for (int j = 0; j < n; ++j) {
      const int e = m.eoc[c * m.me + j];
      const double sg = (m.coe[2 * e + 1] == c) ? -1.0 : 1.0;
      const Vd<N> ut = rd<N>(a.c.vn, e, L, k);
      const int ml = m.maxlev_e[e];
#pragma unroll
      for (int i = 0; i < N; ++i) d2[i] += sg * mask<N>(k, i, ml, mc.cv4[e]) * ut[i] * m.inv_dc[e];
    }
    st<N>(s.d2bar_c + (size_t)c * L + k, d2);
  }
}

template <int N, int H>
__global__ void __launch_bounds__(kAB)
    k_adj_v4b(MeshView m, AdjMeshView am, AdjScratch s, int swz) {
  const Seg g = segment<N>(m, swz);
  if (!g.active) return;
  const int L = m.L, k = g.k;
  for (int e = m.eofs[g.c0] + (int)threadIdx.y; e < m.eofs[g.c1]; e += blockDim.y) {
    const int c0 = m.coe[2 * e], c1 = m.coe[2 * e + 1], v0 = m.voe[2 * e], v1 = m.voe[2 * e + 1];
    const Vd<N> z0 = rd<N>(s.z2bar_v, v0, L, k), z1 = rd<N>(s.z2bar_v, v1, L, k);
    const Vd<N> d0 = rd<N>(s.d2bar_c, c0, L, k), d1 = rd<N>(s.d2bar_c, c1, L, k);
    Vd<N> out;
#pragma unroll
    for (int i = 0; i < N; ++i)
      out[i] = am.ecrv[2 * e] * z0[i] + am.ecrv[2 * e + 1] * z1[i] + am.ecd[2 * e] * d0[i] +
               am.ecd[2 * e + 1] * d1[i];
    st<N>(s.e2bar_e + (size_t)e * L + k, out);
  }
}

// ---------------------------------------------------------------------------
// PV flux: cotangent of S_e = Q_e + P_e, from the edge's own sum and from the
// sums of the edges that list it (transposed edges_on_edge)
// ---------------------------------------------------------------------------
// Per edge at the linearisation state: the PV sum S_e = Q_e + P_e (PV term
// only) and the thickness flux F_e u_e, read by the PV pass for every edge
// whose edges-on-edge list holds e, by the edge pass and by the cell pass.
template <int N, int H, int U, int TM = -1>
__global__ void __launch_bounds__(kAB)
    k_adj_pre(MeshView m, ModelCoef mc, AdjArgs a, DiagOut d, AdjScratch s, int swz) {
  // TM >= 0: the term set, compiled in, with both tendencies on
  const int terms = TM >= 0 ? TM : a.terms;
  const bool do_vn = TM >= 0 || a.do_vn, do_h = TM >= 0 || a.do_h;
  (void)do_h;
  (void)do_vn;
  const Seg g = segment<N>(m, swz);
  if (!g.active) return;
  const int L = m.L, k = g.k;
  const bool pv = do_vn && (terms & T_PV);
  for_edges<U>(m, g, [&](int e, int K) {
    if (pv) {
      const int2 vv = pair_at(m.voe, e);
      const int v0 = vv.x, v1 = vv.y;
      const double fv0 = ecoef<U>(mc.f_v, v0), fv1 = ecoef<U>(mc.f_v, v1);
      const Vd<N> i0 = rd<N>(d.inv_thick_v, v0, L, k), i1 = rd<N>(d.inv_thick_v, v1, L, k);
      const Vd<N> z0 = rd<N>(d.rvort_v, v0, L, k), z1 = rd<N>(d.rvort_v, v1, L, k);
      Vd<N> Se;
#pragma unroll
      for (int i = 0; i < N; ++i)
        Se[i] = 0.5 * (i0[i] * z0[i] + i1[i] * z1[i]) + 0.5 * (i0[i] * fv0 + i1[i] * fv1);
      st<N>(s.se_e + (size_t)e * L + k, Se);
    }
    const int2 cc = pair_at(m.coe, e);
    const Vd<N> h0 = rd<N>(a.h, cc.x, L, k), h1 = rd<N>(a.h, cc.y, L, k);
    const Vd<N> ue = rd<N>(a.vn, e, L, k);
    Vd<N> FU;
#pragma unroll
    for (int i = 0; i < N; ++i) FU[i] = 0.5 * (h0[i] + h1[i]) * ue[i];
    st<N>(s.fu_e + (size_t)e * L + k, FU);
  });
}

// PV-sum cotangent of each edge, and the edge pass's PV share
// (A_e + S_e B_e)/2 from the two transposed sums A_e = sum_t w_te kbar_t S_t,
// B_e = sum_t w_te kbar_t over the edges t whose edges-on-edge list holds e
// (kbar = velocity-tendency cotangent).
// This pass keeps its per-edge tables on every mesh: a slot-uniform form
// (weights and lists read per class) measured slower on MI355X.  O: the
// flux-weighted edges-on-edge sum is given (s.own_e, written by the stage
// kernel at the linearisation state) instead of gathered here.
template <int N, int H, int O = 0>
__global__ void __launch_bounds__(kAB)
    k_adj_pv(MeshView m, AdjMeshView am, ModelCoef mc, AdjArgs a, DiagOut d, AdjScratch s,
             int swz) {
  const Seg g = segment<N>(m, swz);
  if (!g.active) return;
  const int L = m.L, k = g.k;
  for (int e = m.eofs[g.c0] + (int)threadIdx.y; e < m.eofs[g.c1]; e += blockDim.y) {
    const Vd<N> fue = rd<N>(s.fu_e, e, L, k);
    const Vd<N> ute = rd<N>(a.c.vn, e, L, k);
    Vd<N> own = zero<N>(), pa = zero<N>(), pb = zero<N>();
    if constexpr (O) {
      own = rd<N>(s.own_e, e, L, k);
    } else {
      const int ne2 = H ? 10 : m.noe[e];
#pragma unroll
      for (int j = 0; j < ne2; ++j) {
        const int ej = m.eoe[e * m.me2 + j];
        const double w = m.w[e * m.me2 + j];
        const Vd<N> fu = rd<N>(s.fu_e, ej, L, k);
#pragma unroll
        for (int i = 0; i < N; ++i) own[i] += w * fu[i];
      }
    }
    const int q0 = am.tofs[e], nq = H ? 10 : am.tofs[e + 1] - q0;
#pragma unroll
    for (int qq = 0; qq < nq; ++qq) {
      const int q = q0 + qq;
      const int t = am.tedge[q];
      const double w = am.tw[q];
      const Vd<N> ut = rd<N>(a.c.vn, t, L, k), St = rd<N>(s.se_e, t, L, k);
#pragma unroll
      for (int i = 0; i < N; ++i) {
        const double wu = ut[i] * w;
        pb[i] += wu;
        pa[i] += wu * St[i];
      }
    }
    const Vd<N> Se = rd<N>(s.se_e, e, L, k);
    Vd<N> out, px;
#pragma unroll
    for (int i = 0; i < N; ++i) {
      out[i] = 0.5 * (ute[i] * own[i] + pb[i] * fue[i]);
      px[i] = 0.5 * (pa[i] + Se[i] * pb[i]);
    }
    st<N>(s.sbar_e + (size_t)e * L + k, out);
    st<N>(s.px_e + (size_t)e * L + k, px);
  }
}

// ---------------------------------------------------------------------------
// Vertex and cell diagnostics cotangents
// ---------------------------------------------------------------------------
template <int N, int H, int U, int TM = -1>
__global__ void __launch_bounds__(kAB)
    k_adj_vc(MeshView m, AdjMeshView am, ModelCoef mc, AdjCoef ac, AdjArgs a, DiagOut d,
             AdjScratch s, int swz) {
  // TM >= 0: the term set, compiled in, with both tendencies on
  const int terms = TM >= 0 ? TM : a.terms;
  const bool do_vn = TM >= 0 || a.do_vn, do_h = TM >= 0 || a.do_h;
  (void)do_h;
  (void)do_vn;
  const Seg g = segment<N>(m, swz);
  const int L = m.L, k = g.k, NC = m.nc;
  const bool pv = do_vn && (terms & T_PV), vd2 = do_vn && (terms & T_VDEL2),
             vd4 = do_vn && (terms & T_VDEL4), ke = do_vn && (terms & T_KE);
  // gradient w.r.t. visc_del2: sum_e kbar_e s2_e [(D_c1 - D_c0)/dc - (Z_v1 - Z_v0)/dv]
  // regrouped by cell and vertex, sum_c D_c ds_c + sum_v Z_v zs_v, with ds and
  // zs formed beside the del2 cotangents (no edge rows re-read)
  double p_v2 = 0.0;
  if (g.active && (pv || vd2 || vd4)) {
    for_vertices<U>(m, g, [&](int v, int x) {
      Vd<N> qb = zero<N>(), zb = zero<N>(), zs = zero<N>();
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        const int e = m.eov[3 * v + j];
        const double sg =
            (U ? (x ? am.uvsg[3 + j] : am.uvsg[j]) != 0 : m.voe[2 * e + 1] == v) ? -1.0 : 1.0;
        if (pv) {
          const Vd<N> sb = rd<N>(s.sbar_e, e, L, k);
#pragma unroll
          for (int i = 0; i < N; ++i) qb[i] += 0.5 * sb[i];
        }
        if (vd2) {
          const Vd<N> ut = rd<N>(a.c.vn, e, L, k);
          const int ml = ecoef<U>(m.maxlev_e, e);
          const double cv2 = ecoef<U>(mc.cv2, e), s2 = ecoef<U>(m.s2, e),
                       idv = ecoef<U>(m.inv_dv, e);
#pragma unroll
          for (int i = 0; i < N; ++i) {
            zb[i] += sg * mask<N>(k, i, ml, cv2) * ut[i] * idv;
            zs[i] += sg * mask<N>(k, i, ml, s2) * ut[i] * idv;
          }
        }
        if (vd4) {
          const Vd<N> eb = rd<N>(s.e2bar_e, e, L, k);
          const double idv4 = ecoef<U>(m.inv_dv4, e);
#pragma unroll
          for (int i = 0; i < N; ++i) zb[i] += sg * eb[i] * idv4;
        }
      }
      Vd<N> tb = zero<N>(), z = zero<N>();
      if (pv || vd2) z = rd<N>(d.rvort_v, v, L, k);
      if (vd2) {
#pragma unroll
        for (int i = 0; i < N; ++i) p_v2 += z[i] * zs[i];
      }
      if (pv) {
        const Vd<N> io = rd<N>(d.inv_thick_v, v, L, k);
        const double f = ecoef<U>(mc.f_v, v);
#pragma unroll
        for (int i = 0; i < N; ++i) {
          zb[i] += io[i] * qb[i];
          tb[i] = -io[i] * io[i] * (z[i] + f) * qb[i];
        }
      }
      st<N>(s.zbar_v + (size_t)v * L + k, zb);
      st<N>(s.tbar_v + (size_t)v * L + k, tb);
    });
  }
  const bool td4 = (terms & T_TDEL4) && a.ntr > 0;
  for (int c = g.c0 + (int)threadIdx.y; g.active && c < g.c1; c += blockDim.y) {
    Vd<N> kb = zero<N>(), db = zero<N>(), ds = zero<N>();
    const int n = H ? 6 : m.noc[c];
    int ec[6];
    if constexpr (U) row_ints<6>(m.eoc, (size_t)c * 6, ec);
#pragma unroll 6
    for (int j = 0; j < n; ++j) {
      const int e = U ? ec[j] : m.eoc[c * m.me + j];
      const bool first = U ? am.ufirst[j] != 0 : m.coe[2 * e] == c;
      const double idc = ecoef<U>(m.inv_dc, e);
      if (ke || vd2) {
        const Vd<N> ut = rd<N>(a.c.vn, e, L, k);
        const int ml = ecoef<U>(m.maxlev_e, e);
        const double cv2 = ecoef<U>(mc.cv2, e), s2 = ecoef<U>(m.s2, e);
#pragma unroll
        for (int i = 0; i < N; ++i) {
          if (ke) kb[i] += (first ? 1.0 : -1.0) * ut[i] * idc;
          if (vd2) {
            db[i] += (first ? -1.0 : 1.0) * mask<N>(k, i, ml, cv2) * ut[i] * idc;
            ds[i] += (first ? -1.0 : 1.0) * mask<N>(k, i, ml, s2) * ut[i] * idc;
          }
        }
      }
      if (vd4) {
        const Vd<N> eb = rd<N>(s.e2bar_e, e, L, k);
#pragma unroll
        for (int i = 0; i < N; ++i) db[i] += (first ? -1.0 : 1.0) * eb[i] * idc;
      }
    }
    if (ke) st<N>(s.kbar_c + (size_t)c * L + k, kb);
    if (vd2 || vd4) st<N>(s.dbar_c + (size_t)c * L + k, db);
    if (vd2) {
      const Vd<N> D = rd<N>(d.div_c, c, L, k);
#pragma unroll
      for (int i = 0; i < N; ++i) p_v2 += D[i] * ds[i];
    }
    if (td4) {
      for (int l = 0; l < a.ntr; ++l) {
        const double *tl = a.c.tr + (size_t)l * NC * L;
        const Vd<N> tc = rd<N>(tl, c, L, k);
        Vd<N> gb = zero<N>();
#pragma unroll 6
        for (int j = 0; j < n; ++j) {
          const int e = U ? ec[j] : m.eoc[c * m.me + j];
          bool second;
          int o;
          if constexpr (U) {
            second = am.ufirst[j] == 0;
            o = m.coc[c * 6 + j];
          } else {
            const int a0 = m.coe[2 * e], a1 = m.coe[2 * e + 1];
            second = a1 == c;
            o = second ? a0 : a1;
          }
          const double sg = second ? -1.0 : 1.0;
          const int Kj = U ? am.ukslot[j] : 0;
          const double c40 = eslot<U>(ac.ect4, e, Kj, 0), c41 = eslot<U>(ac.ect4, e, Kj, 1);
          const double idc = ecoef<U>(m.inv_dc, e);
          const Vd<N> to = rd<N>(tl, o, L, k);
          Vd<N> t0, t1;
#pragma unroll
          for (int i = 0; i < N; ++i) {
            t0[i] = second ? to[i] : tc[i];
            t1[i] = second ? tc[i] : to[i];
          }
#pragma unroll
          for (int i = 0; i < N; ++i) gb[i] += sg * (c40 * t0[i] + c41 * t1[i]) * idc;
        }
        st<N>(s.gbar_c + ((size_t)l * NC + c) * L + k, gb);
      }
    }
  }
  const int b = blockIdx.x;
  block_partial(a, p_v2, s.partial + b);
}

// ---------------------------------------------------------------------------
// Edge pass: velocity cotangent (accumulated) and edge-thickness cotangent
// ---------------------------------------------------------------------------
template <int N, int H, int X, int U, int TM = -1>
// four waves per SIMD (at most 128 VGPRs; fits without spilling)
__global__ void __launch_bounds__(kAB) __attribute__((amdgpu_waves_per_eu(4)))
    k_adj_edge(MeshView m, AdjMeshView am, ModelCoef mc, AdjCoef ac, AdjArgs a, DiagOut d,
               AdjScratch s, int swz) {
  // TM >= 0: the term set, compiled in, with both tendencies on
  const int terms = TM >= 0 ? TM : a.terms;
  const bool do_vn = TM >= 0 || a.do_vn, do_h = TM >= 0 || a.do_h;
  (void)do_h;
  (void)do_vn;
  const Seg g = segment<N>(m, swz);
  const int L = m.L, k = g.k, NC = m.nc;
  const bool pv = do_vn && (terms & T_PV), vd2 = do_vn && (terms & T_VDEL2),
             vd4 = do_vn && (terms & T_VDEL4), ke = do_vn && (terms & T_KE);
  const bool td2 = X && (terms & T_TDEL2), td4 = X && (terms & T_TDEL4);
  double p_v4 = 0.0;
  if (g.active) {
    for_edges<U>(m, g, [&](int e, int K) {
      const int2 cc = pair_at(m.coe, e), vv = pair_at(m.voe, e);
      const int c0 = cc.x, c1 = cc.y, v0 = vv.x, v1 = vv.y;
      const double cd0 = eslot<U>(am.ecd, e, K, 0), cd1 = eslot<U>(am.ecd, e, K, 1),
                   idc = ecoef<U>(m.inv_dc, e);
      const Vd<N> ue = rd<N>(a.vn, e, L, k);
      const Vd<N> h0 = rd<N>(a.h, c0, L, k), h1 = rd<N>(a.h, c1, L, k);
      Vd<N> F, ih0, ih1, ub = zero<N>(), fb = zero<N>();
#pragma unroll
      for (int i = 0; i < N; ++i) F[i] = 0.5 * (h0[i] + h1[i]);
      if (a.ntr > 0) {
#pragma unroll
        for (int i = 0; i < N; ++i) {
          ih0[i] = 1.0 / h0[i];
          ih1[i] = 1.0 / h1[i];
        }
      }
      if (do_h) {
        const Vd<N> g0 = rd<N>(a.c.h, c0, L, k), g1 = rd<N>(a.c.h, c1, L, k);
#pragma unroll
        for (int i = 0; i < N; ++i) {
          const double hc = cd0 * g0[i] + cd1 * g1[i];
          ub[i] -= F[i] * hc;
          fb[i] -= ue[i] * hc;
        }
      }
      const bool mix = td2 || td4;
      for (int l = 0; l < a.ntr; ++l) {
        const size_t off = (size_t)l * NC * L;
        const Vd<N> t0 = rd<N>(a.tr + off, c0, L, k), t1 = rd<N>(a.tr + off, c1, L, k);
        const Vd<N> b0 = rd<N>(a.c.tr + off, c0, L, k), b1 = rd<N>(a.c.tr + off, c1, L, k);
        Vd<N> g0 = zero<N>(), g1 = zero<N>();
        if (td4) {
          g0 = rd<N>(s.gbar_c + off, c0, L, k);
          g1 = rd<N>(s.gbar_c + off, c1, L, k);
        }
#pragma unroll
        for (int i = 0; i < N; ++i) {
          const double n0 = t0[i] * ih0[i], n1 = t1[i] * ih1[i];
          const double ne = (n0 + n1) * 0.5, grad = (n1 - n0) * idc;
          const double w = cd0 * b0[i] + cd1 * b1[i];
          ub[i] -= w * F[i] * ne;
          fb[i] -= w * ue[i] * ne;
          if (td2)
            fb[i] += (eslot<U>(ac.ect2, e, K, 0) * b0[i] + eslot<U>(ac.ect2, e, K, 1) * b1[i]) *
                     ecoef<U>(m.s2, e) * grad;
          if (td4) fb[i] += (cd0 * g0[i] + cd1 * g1[i]) * grad;
        }
        if (!mix) {
          // both cells' advection share; the same value the cell pass would
          // form from the edge's thickness flux and the cotangent rows
          Vd<N> xa;
#pragma unroll
          for (int i = 0; i < N; ++i) {
            const double w = cd0 * b0[i] + cd1 * b1[i];
            xa[i] = 0.5 * (-w * (F[i] * ue[i]));
          }
          st<N>(s.xa_e + ((size_t)l * m.ne + e) * L + k, xa);
        }
      }
      if (do_vn) {
        if (pv) {
          // sum_t kbar_t w_te (S_t + S_e)/2 = (A_e + S_e B_e)/2 (k_adj_pv)
          const Vd<N> px = rd<N>(s.px_e, e, L, k);
#pragma unroll
          for (int i = 0; i < N; ++i) {
            ub[i] += px[i] * F[i];
            fb[i] += px[i] * ue[i];
          }
        }
        if (pv || vd2 || vd4) {
          const Vd<N> zb0 = rd<N>(s.zbar_v, v0, L, k), zb1 = rd<N>(s.zbar_v, v1, L, k);
          const double r0 = eslot<U>(am.ecrv, e, K, 0), r1 = eslot<U>(am.ecrv, e, K, 1);
#pragma unroll
          for (int i = 0; i < N; ++i) ub[i] += r0 * zb0[i] + r1 * zb1[i];
        }
        if (ke) {
          const Vd<N> k0 = rd<N>(s.kbar_c, c0, L, k), k1 = rd<N>(s.kbar_c, c1, L, k);
          const double q0 = eslot<U>(am.ecke, e, K, 0), q1 = eslot<U>(am.ecke, e, K, 1);
#pragma unroll
          for (int i = 0; i < N; ++i) ub[i] += (q0 * k0[i] + q1 * k1[i]) * ue[i] * 0.5;
        }
        if (vd2 || vd4) {
          const Vd<N> d0 = rd<N>(s.dbar_c, c0, L, k), d1 = rd<N>(s.dbar_c, c1, L, k);
#pragma unroll
          for (int i = 0; i < N; ++i) ub[i] += cd0 * d0[i] + cd1 * d1[i];
        }
        if (vd4) {
          const Vd<N> ut = rd<N>(a.c.vn, e, L, k);
          const int ml = ecoef<U>(m.maxlev_e, e);
          const double s4 = ecoef<U>(m.s4, e), idv = ecoef<U>(m.inv_dv, e);
          const Vd<N> D0 = rd<N>(d.vdel2_div_c, c0, L, k), D1 = rd<N>(d.vdel2_div_c, c1, L, k);
          const Vd<N> Z0 = rd<N>(d.vdel2_rv_v, v0, L, k), Z1 = rd<N>(d.vdel2_rv_v, v1, L, k);
#pragma unroll
          for (int i = 0; i < N; ++i)
            p_v4 -= ut[i] * mask<N>(k, i, ml, s4) * ((D1[i] - D0[i]) * idc - (Z1[i] - Z0[i]) * idv);
        }
      }
      adj_out<N>(a, 1, a.c.ovn, (size_t)e * L + k, ub);
      st<N>(s.fbar_e + (size_t)e * L + k, fb);
    });
  }
  const int b = blockIdx.x;
  block_partial(a, p_v4, s.partial + a.nparts + b);
}

// ---------------------------------------------------------------------------
// Cell pass: thickness and tracer cotangents (accumulated)
// ---------------------------------------------------------------------------
template <int N, int H, int X, int U, int TM = -1>
__global__ void __launch_bounds__(kAB)
    k_adj_cell(MeshView m, AdjMeshView am, ModelCoef mc, AdjCoef ac, AdjArgs a, DiagOut d,
               AdjScratch s, int swz) {
  // TM >= 0: the term set, compiled in, with both tendencies on
  const int terms = TM >= 0 ? TM : a.terms;
  const bool do_vn = TM >= 0 || a.do_vn, do_h = TM >= 0 || a.do_h;
  (void)do_h;
  (void)do_vn;
  const Seg g = segment<N>(m, swz);
  const int L = m.L, k = g.k, NC = m.nc, me = m.me;
  const bool pv = do_vn && (terms & T_PV), ssh = do_vn && (terms & T_SSH);
  const bool td2 = X && (terms & T_TDEL2), td4 = X && (terms & T_TDEL4);
  double p_t2 = 0.0, p_t4 = 0.0;
  if (g.active) {
    for (int c = g.c0 + (int)threadIdx.y; c < g.c1; c += blockDim.y) {
      const int n = H ? 6 : m.noc[c];
      const Vd<N> hc = rd<N>(a.h, c, L, k);
      Vd<N> hb = zero<N>();
      // U: the cell's six edges and vertices in 8-byte loads; per-slot
      // coefficients are the same for every cell
      int ec[6], vc[6];
      if constexpr (U) {
        row_ints<6>(m.eoc, (size_t)c * 6, ec);
        if (pv) row_ints<6>(am.voc, (size_t)c * 6, vc);
      }
#pragma unroll 6
      for (int j = 0; j < n; ++j) {
        const int e = U ? ec[j] : m.eoc[c * me + j];
        const Vd<N> fb = rd<N>(s.fbar_e, e, L, k);
#pragma unroll
        for (int i = 0; i < N; ++i) hb[i] += 0.5 * fb[i];
        if (ssh) {
          const double sg = (U ? am.ufirst[j] != 0 : m.coe[2 * e] == c) ? 1.0 : -1.0;
          const double idc = ecoef<U>(m.inv_dc, e);
          const Vd<N> ut = rd<N>(a.c.vn, e, L, k);
#pragma unroll
          for (int i = 0; i < N; ++i) hb[i] += mc.grav * sg * ut[i] * idc;
        }
        if (pv) {
          const int v = U ? vc[j] : am.voc[c * me + j];
          const double th = rslot<U>(am.ccth, c, me, j);
          const Vd<N> tb = rd<N>(s.tbar_v, v, L, k);
#pragma unroll
          for (int i = 0; i < N; ++i) hb[i] += th * tb[i];
        }
      }
      // the cell's own rows are loaded once; each edge then loads only the
      // rows of the cell across it.  The advection term reads the edge's
      // thickness flux F_e u_e (k_adj_pre); thicknesses and tracers across
      // the edge are needed only by the diffusion terms.
      const bool mix = td2 || td4;
      Vd<N> ihc;
#pragma unroll
      for (int i = 0; i < N; ++i) ihc[i] = 1.0 / hc[i];
      // HO (six edges, term set compiled in): the diffusion terms'
      // tracer-independent rows of the six edges (the thickness flux, the mean
      // thickness and 1/h across the edge) are loaded once for all tracers
      // (with run-time terms their registers would cost occupancy)
      constexpr bool HO = H && TM >= 0;
      Vd<N> fuH[6], FH[6], ihoH[6];
      if constexpr (HO) {
        if (mix && a.ntr > 0) {
#pragma unroll
          for (int j = 0; j < 6; ++j) {
            const int e = U ? ec[j] : m.eoc[c * me + j];
            bool second;
            int o;
            if constexpr (U) {
              second = am.ufirst[j] == 0;
              o = m.coc[c * 6 + j];
            } else {
              const int a0 = m.coe[2 * e], a1 = m.coe[2 * e + 1];
              second = a1 == c;
              o = second ? a0 : a1;
            }
            fuH[j] = rd<N>(s.fu_e, e, L, k);
            const Vd<N> ho = rd<N>(a.h, o, L, k);
#pragma unroll
            for (int i = 0; i < N; ++i) {
              const double h0 = second ? ho[i] : hc[i], h1 = second ? hc[i] : ho[i];
              FH[j][i] = 0.5 * (h0 + h1);
              ihoH[j][i] = td2 ? 1.0 / (second ? h0 : h1) : 0.0;
            }
          }
        }
      }
      for (int l = 0; l < a.ntr; ++l) {
        const size_t off = (size_t)l * NC * L;
        const Vd<N> tcl = rd<N>(a.tr + off, c, L, k);
        const Vd<N> bcl = rd<N>(a.c.tr + off, c, L, k);
        Vd<N> gcl = zero<N>(), Gcl = zero<N>();
        if (td4) {
          gcl = rd<N>(s.gbar_c + off, c, L, k);
          Gcl = rd<N>(d.trdel2_c + off, c, L, k);
        }
        Vd<N> nb = zero<N>(), t2 = zero<N>(), t4 = zero<N>();
#pragma unroll 6
        for (int j = 0; j < n; ++j) {
          const int e = U ? ec[j] : m.eoc[c * me + j];
          if (!mix) {
            // advection only: the edge pass stored this edge's share
            const Vd<N> xa = ld<N>(s.xa_e + ((size_t)l * m.ne + e) * L + k);
#pragma unroll
            for (int i = 0; i < N; ++i) nb[i] += xa[i];
            continue;
          }
          bool second;
          int o;
          if constexpr (U) {
            second = am.ufirst[j] == 0;
            o = m.coc[c * 6 + j];
          } else {
            const int a0 = m.coe[2 * e], a1 = m.coe[2 * e + 1];
            second = a1 == c;
            o = second ? a0 : a1;
          }
          const double sg = second ? 1.0 : -1.0;
          const int Kj = U ? am.ukslot[j] : 0;
          const double cd0 = eslot<U>(am.ecd, e, Kj, 0), cd1 = eslot<U>(am.ecd, e, Kj, 1),
                       idc = ecoef<U>(m.inv_dc, e);
          const double c20 = eslot<U>(ac.ect2, e, Kj, 0), c21 = eslot<U>(ac.ect2, e, Kj, 1);
          const double s2 = ecoef<U>(m.s2, e), s4 = ecoef<U>(m.s4, e);
          const Vd<N> fu = HO ? fuH[j] : rd<N>(s.fu_e, e, L, k);
          const Vd<N> bo = rd<N>(a.c.tr + off, o, L, k);
          Vd<N> b0, b1, F, tt0 = zero<N>(), tt1 = zero<N>(), g0 = zero<N>(), g1 = zero<N>(),
                           G0 = zero<N>(), G1 = zero<N>(), ih0 = zero<N>(), ih1 = zero<N>();
#pragma unroll
          for (int i = 0; i < N; ++i) {
            b0[i] = second ? bo[i] : bcl[i];
            b1[i] = second ? bcl[i] : bo[i];
          }
          Vd<N> iho;
          if constexpr (HO) {
            F = FH[j];
            iho = ihoH[j];
          } else {
            const Vd<N> ho = rd<N>(a.h, o, L, k);
#pragma unroll
            for (int i = 0; i < N; ++i) {
              const double h0 = second ? ho[i] : hc[i], h1 = second ? hc[i] : ho[i];
              F[i] = 0.5 * (h0 + h1);
              iho[i] = td2 ? 1.0 / (second ? h0 : h1) : 0.0;
            }
          }
          if (td2) {
            const Vd<N> to = rd<N>(a.tr + off, o, L, k);
#pragma unroll
            for (int i = 0; i < N; ++i) {
              ih0[i] = second ? iho[i] : ihc[i];
              ih1[i] = second ? ihc[i] : iho[i];
              tt0[i] = second ? to[i] : tcl[i];
              tt1[i] = second ? tcl[i] : to[i];
            }
          }
          if (td4) {
            const Vd<N> go = rd<N>(s.gbar_c + off, o, L, k), Go = rd<N>(d.trdel2_c + off, o, L, k);
#pragma unroll
            for (int i = 0; i < N; ++i) {
              g0[i] = second ? go[i] : gcl[i];
              g1[i] = second ? gcl[i] : go[i];
              G0[i] = second ? Go[i] : Gcl[i];
              G1[i] = second ? Gcl[i] : Go[i];
            }
          }
          const double cdc = rslot<U>(m.cdiv, c, me, j);
#pragma unroll
          for (int i = 0; i < N; ++i) {
            const double w = cd0 * b0[i] + cd1 * b1[i];
            nb[i] += 0.5 * (-w * fu[i]);
            if (td2) {
              const double grad = (tt1[i] * ih1[i] - tt0[i] * ih0[i]) * idc;
              const double w2 = c20 * b0[i] + c21 * b1[i];
              nb[i] += sg * w2 * F[i] * s2 * idc;
              t2[i] += cdc * F[i] * s2 * grad;
            }
            if (td4) {
              const double wg = cd0 * g0[i] + cd1 * g1[i];
              nb[i] += sg * wg * F[i] * idc;
              t4[i] -= cdc * s4 * (G1[i] - G0[i]) * idc;
            }
          }
        }
        Vd<N> vt;
#pragma unroll
        for (int i = 0; i < N; ++i) {
          const double ih = ihc[i];
          vt[i] = nb[i] * ih;
          hb[i] -= nb[i] * tcl[i] * ih * ih;
          p_t2 += bcl[i] * t2[i];
          p_t4 += bcl[i] * t4[i];
        }
        adj_out<N>(a, 2, a.c.otr, off + (size_t)c * L + k, vt);
      }
      adj_out<N>(a, 0, a.c.oh, (size_t)c * L + k, hb);
    }
  }
  const int b = blockIdx.x;
  block_partials2(a, p_t2, p_t4, s.partial + 2 * a.nparts + b, s.partial + 3 * a.nparts + b);
}

// pgrad[p] += sum of the block partials of parameter p, in a fixed order
__global__ void __launch_bounds__(kAB) k_adj_param_reduce(double *partial, int nparts,
                                                         int nused, double vs, double *pgrad) {
  __shared__ double sh[kAB];
  const int p = blockIdx.x;
  double acc = 0.0;
  for (int i = threadIdx.x; i < nused; i += kAB) acc += partial[p * nparts + i];
  sh[threadIdx.x] = acc;
  __syncthreads();
  for (int s = kAB / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) pgrad[p] += vs * sh[0];
}

struct AGeo {
  dim3 grid, block;
  int N;
};

AGeo adj_geometry(const MeshView &m, const Launch &lc) {
  const int N = (m.L % 2 == 0) ? 2 : 1;
  int lanes = (m.L + N - 1) / N;
  if (lanes > 64) lanes = 64;
  int lx = 1;
  while (lx < lanes) lx <<= 1;
  const int rows = kAB / lx;
  const int chunks = (m.L + lx * N - 1) / (lx * N);
  return AGeo{dim3((m.nc + rows - 1) / rows * chunks), dim3(lx, rows), N};
}

} // namespace

int adj_partial_slots(const MeshView &m, const Launch &lc) {
  const AGeo g = adj_geometry(m, lc);
  return (int)(g.grid.x * g.grid.y);
}

// ---------------------------------------------------------------------------
// Elementwise helpers and the least-squares misfit
// ---------------------------------------------------------------------------
namespace {

constexpr int kEB = 256;
constexpr int kMisfitMaxBlocks = 1024;

__global__ void __launch_bounds__(kEB) k_fill(double *y, double v, size_t n) {
  for (size_t i = blockIdx.x * (size_t)kEB + threadIdx.x; i < n; i += (size_t)gridDim.x * kEB)
    y[i] = v;
}

__global__ void __launch_bounds__(kEB)
    k_lincomb(double *y, double alpha, const double *x, double beta, const double *z, size_t n) {
  for (size_t i = blockIdx.x * (size_t)kEB + threadIdx.x; i < n; i += (size_t)gridDim.x * kEB)
    y[i] = alpha * x[i] + beta * z[i];
}

__device__ void block_sum_store(double v, double *out) {
  __shared__ double sh[kEB];
  sh[threadIdx.x] = v;
  __syncthreads();
  for (int s = kEB / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = sh[0];
}

__global__ void __launch_bounds__(kEB)
    k_misfit(const double *y, const double *obs, const double *area, int nent, int L, size_t n,
             double w, double *cot, double *partial) {
  double acc = 0.0;
  for (size_t i = blockIdx.x * (size_t)kEB + threadIdx.x; i < n; i += (size_t)gridDim.x * kEB) {
    const int ent = (int)((i / L) % nent);
    const double wa = w * area[ent];
    const double r = y[i] - obs[i];
    cot[i] = wa * r;
    acc += 0.5 * wa * r * r;
  }
  block_sum_store(acc, partial + blockIdx.x);
}

__global__ void __launch_bounds__(kEB)
    k_sum(const double *partial, int n, double *out, int accumulate) {
  double acc = 0.0;
  for (int i = threadIdx.x; i < n; i += kEB) acc += partial[i];
  __shared__ double sh[kEB];
  sh[threadIdx.x] = acc;
  __syncthreads();
  for (int s = kEB / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = accumulate ? *out + sh[0] : sh[0];
}

unsigned ew_blocks(size_t n) {
  const size_t b = (n + kEB - 1) / kEB;
  return (unsigned)std::max<size_t>(1, std::min<size_t>(b, 8192));
}

} // namespace

void launch_fill(Real *y, Real v, size_t n, hipStream_t s) {
  if (!n) return;
  TimerScope ts("update", s);
  hipLaunchKernelGGL(k_fill, dim3(ew_blocks(n)), dim3(kEB), 0, s, y, v, n);
  OMEGA_HIP_CHECK(hipGetLastError());
}

void launch_lincomb(Real *y, Real alpha, const Real *x, Real beta, const Real *z, size_t n,
                    hipStream_t s) {
  if (!n) return;
  TimerScope ts("update", s);
  hipLaunchKernelGGL(k_lincomb, dim3(ew_blocks(n)), dim3(kEB), 0, s, y, alpha, x, beta, z, n);
  OMEGA_HIP_CHECK(hipGetLastError());
}

int misfit_blocks(size_t n) {
  const size_t b = (n + kEB - 1) / kEB;
  return (int)std::max<size_t>(1, std::min<size_t>(b, kMisfitMaxBlocks));
}

void launch_misfit(const Real *y, const Real *obs, const Real *area, int nent, int L, int nrep,
                   Real w, Real *cot, Real *partial, hipStream_t s) {
  const size_t n = (size_t)nrep * nent * L;
  TimerScope ts("reduce", s);
  hipLaunchKernelGGL(k_misfit, dim3(misfit_blocks(n)), dim3(kEB), 0, s, y, obs, area, nent, L,
                     n, w, cot, partial);
  OMEGA_HIP_CHECK(hipGetLastError());
}

void launch_sum(const Real *partial, int n, Real *out, int accumulate, hipStream_t s) {
  TimerScope ts("reduce", s);
  hipLaunchKernelGGL(k_sum, dim3(1), dim3(kEB), 0, s, partial, n, out, accumulate);
  OMEGA_HIP_CHECK(hipGetLastError());
}

void launch_tendency_vjp(const MeshView &m, const AdjMeshView &am, const Launch &lc,
                         const ModelCoef &mc, const AdjCoef &ac, const AdjArgs &a,
                         const DiagOut &d, const AdjScratch &s, hipStream_t st, bool pre) {
  TimerScope ts("adjoint", st);
  const AGeo g = adj_geometry(m, lc);
  const int swz = lc.swizzle;
  const bool vd4 = a.do_vn && (a.terms & T_VDEL4);
  const bool pv = a.do_vn && (a.terms & T_PV);
  // regular hexagonal meshes: constant neighbour counts, unrolled loops;
  // with slot-uniform tables and coefficients also the per-slot reads (U),
  // and for the common term sets with both tendencies on the terms compiled
  // in (tm)
  const bool U = g.N == 2 && am.fixed && am.uni && m.coef_uni && mc.uniform && ac.uniform;
  const bool tset = U && m.term_sets && a.do_h && a.do_vn &&
                    (a.terms == 7 || a.terms == 15 || a.terms == 79 || a.terms == 127);
  const int tm = tset ? a.terms : -1;
#define OMEGA_ADJ_LAUNCH(K, ...)                                                                \
  do {                                                                                         \
    if (tm == 7)                                                                               \
      hipLaunchKernelGGL((K<2, 1, 1, 7>), g.grid, g.block, 0, st, __VA_ARGS__);                 \
    else if (tm == 15)                                                                         \
      hipLaunchKernelGGL((K<2, 1, 1, 15>), g.grid, g.block, 0, st, __VA_ARGS__);                \
    else if (tm == 79)                                                                         \
      hipLaunchKernelGGL((K<2, 1, 1, 79>), g.grid, g.block, 0, st, __VA_ARGS__);                \
    else if (tm == 127)                                                                        \
      hipLaunchKernelGGL((K<2, 1, 1, 127>), g.grid, g.block, 0, st, __VA_ARGS__);               \
    else if (U)                                                                                \
      hipLaunchKernelGGL((K<2, 1, 1>), g.grid, g.block, 0, st, __VA_ARGS__);                    \
    else if (g.N == 2 && am.fixed)                                                             \
      hipLaunchKernelGGL((K<2, 1, 0>), g.grid, g.block, 0, st, __VA_ARGS__);                    \
    else if (g.N == 2)                                                                         \
      hipLaunchKernelGGL((K<2, 0, 0>), g.grid, g.block, 0, st, __VA_ARGS__);                    \
    else if (am.fixed)                                                                         \
      hipLaunchKernelGGL((K<1, 1, 0>), g.grid, g.block, 0, st, __VA_ARGS__);                    \
    else                                                                                       \
      hipLaunchKernelGGL((K<1, 0, 0>), g.grid, g.block, 0, st, __VA_ARGS__);                    \
  } while (0)
#define OMEGA_ADJ_LAUNCH2(K, ...)                                                               \
  do {                                                                                         \
    if (g.N == 2 && am.fixed)                                                                  \
      hipLaunchKernelGGL((K<2, 1>), g.grid, g.block, 0, st, __VA_ARGS__);                       \
    else if (g.N == 2)                                                                         \
      hipLaunchKernelGGL((K<2, 0>), g.grid, g.block, 0, st, __VA_ARGS__);                       \
    else if (am.fixed)                                                                         \
      hipLaunchKernelGGL((K<1, 1>), g.grid, g.block, 0, st, __VA_ARGS__);                       \
    else                                                                                       \
      hipLaunchKernelGGL((K<1, 0>), g.grid, g.block, 0, st, __VA_ARGS__);                       \
  } while (0)
  if (vd4) {
    OMEGA_ADJ_LAUNCH2(k_adj_v4a, m, mc, a, s, swz);
    OMEGA_ADJ_LAUNCH2(k_adj_v4b, m, am, s, swz);
  }
  // edge PV sums (PV term) and thickness fluxes (PV term, tracer advection),
  // unless the stage kernel at this state wrote them
  if (!pre && (pv || a.ntr > 0)) OMEGA_ADJ_LAUNCH(k_adj_pre, m, mc, a, d, s, swz);
  if (pv && pre) {
    OMEGA_REQUIRE(s.own_e != nullptr, "tendency VJP: pre rows without the edges-on-edge sum");
    if (g.N == 2 && am.fixed)
      hipLaunchKernelGGL((k_adj_pv<2, 1, 1>), g.grid, g.block, 0, st, m, am, mc, a, d, s, swz);
    else if (g.N == 2)
      hipLaunchKernelGGL((k_adj_pv<2, 0, 1>), g.grid, g.block, 0, st, m, am, mc, a, d, s, swz);
    else if (am.fixed)
      hipLaunchKernelGGL((k_adj_pv<1, 1, 1>), g.grid, g.block, 0, st, m, am, mc, a, d, s, swz);
    else
      hipLaunchKernelGGL((k_adj_pv<1, 0, 1>), g.grid, g.block, 0, st, m, am, mc, a, d, s, swz);
  } else if (pv) {
    OMEGA_ADJ_LAUNCH2(k_adj_pv, m, am, mc, a, d, s, swz);
  }
  OMEGA_ADJ_LAUNCH(k_adj_vc, m, am, mc, ac, a, d, s, swz);
  // X = 0: no tracer diffusion term (edge and cell passes without its code);
  // the compiled term sets 7 and 15 have none, 79 and 127 have it
#define OMEGA_ADJ_LAUNCH_X(K, X, TA, TB, ...)                                                   \
  do {                                                                                         \
    if (tm == TA)                                                                              \
      hipLaunchKernelGGL((K<2, 1, X, 1, TA>), g.grid, g.block, 0, st, __VA_ARGS__);             \
    else if (tm == TB)                                                                         \
      hipLaunchKernelGGL((K<2, 1, X, 1, TB>), g.grid, g.block, 0, st, __VA_ARGS__);             \
    else if (U)                                                                                \
      hipLaunchKernelGGL((K<2, 1, X, 1>), g.grid, g.block, 0, st, __VA_ARGS__);                 \
    else if (g.N == 2 && am.fixed)                                                             \
      hipLaunchKernelGGL((K<2, 1, X, 0>), g.grid, g.block, 0, st, __VA_ARGS__);                 \
    else if (g.N == 2)                                                                         \
      hipLaunchKernelGGL((K<2, 0, X, 0>), g.grid, g.block, 0, st, __VA_ARGS__);                 \
    else if (am.fixed)                                                                         \
      hipLaunchKernelGGL((K<1, 1, X, 0>), g.grid, g.block, 0, st, __VA_ARGS__);                 \
    else                                                                                       \
      hipLaunchKernelGGL((K<1, 0, X, 0>), g.grid, g.block, 0, st, __VA_ARGS__);                 \
  } while (0)
  if (a.terms & (T_TDEL2 | T_TDEL4)) {
    OMEGA_ADJ_LAUNCH_X(k_adj_edge, 1, 79, 127, m, am, mc, ac, a, d, s, swz);
    OMEGA_ADJ_LAUNCH_X(k_adj_cell, 1, 79, 127, m, am, mc, ac, a, d, s, swz);
  } else {
    OMEGA_ADJ_LAUNCH_X(k_adj_edge, 0, 7, 15, m, am, mc, ac, a, d, s, swz);
    OMEGA_ADJ_LAUNCH_X(k_adj_cell, 0, 7, 15, m, am, mc, ac, a, d, s, swz);
  }
#undef OMEGA_ADJ_LAUNCH_X
#undef OMEGA_ADJ_LAUNCH2
#undef OMEGA_ADJ_LAUNCH
  if (!a.part_keep)
    hipLaunchKernelGGL(k_adj_param_reduce, dim3(4), dim3(kAB), 0, st, s.partial, a.nparts,
                       (int)(g.grid.x * g.grid.y), a.part_add ? 1.0 : a.epi.vs, s.pgrad);
  OMEGA_HIP_CHECK(hipGetLastError());
}

} // namespace omega_hip
