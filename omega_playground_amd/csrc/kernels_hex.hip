// Pattern-specialised forward kernels for the regular hexagonal mesh.
//
// Each lane carries N consecutive levels of one entity (N = 2: one 16-byte
// load per lane and row, 32 lanes per 64-level row, two entities per
// wavefront; N = 1: 64 lanes per row, one entity per wavefront, whose
// indices are then wavefront-uniform and come through the scalar cache).
// One row of lanes covers 64 levels, so the layer count must be a multiple
// of 64.
//
// The edge kernel loads a per-edge record (edges-on-edge, the ten distinct
// vertices and cells they touch) once, issues all row loads of the 10-point
// PV stencil back to back -- 10 distinct cells, 10 vertices, 10 edges
// instead of 20 + 20 + 10 -- and indexes them through the compile-time slot
// tables of hex_pattern.hpp.  The default N = 2 kernel (k_tend_hex2v) puts
// two cells in a wavefront, reads their tables through the scalar cache and
// forms the vertex PV products of each cell's 16 union vertices once for
// its three edges.  Sums and products keep the reference's order,
// so results are the same bits as the general kernels.  The vector-memory
// address path of the CU is what bounds these kernels, so the aim throughout
// is the fewest load instructions per level.
#include "hex_ops.hpp"
#include "tiles.hpp"

namespace omega_hip {

namespace {

using namespace dev;
using namespace hexops;

template <int N>
__global__ void __launch_bounds__(256)
    k_diag_hex(MeshView m, int terms, int ntr, const double *__restrict__ h,
               const double *__restrict__ vn, const double *__restrict__ tr, DiagOut d, int swz) {
  const int s = logical_block(blockIdx.x, gridDim.x, swz);
  const int R = blockDim.y;
  const int c = ent<N>(s * R + (int)threadIdx.y);
  if (c >= m.nc) return;
  const int k = (blockIdx.y * blockDim.x + threadIdx.x) * N;
  if (terms & (T_PV | T_VDEL2 | T_VDEL4)) {
    vertex_diag_u<N>(m, 2 * c, k, h, vn, d);
    vertex_diag_u<N>(m, 2 * c + 1, k, h, vn, d);
  }
  cell_diag_u<N>(m, terms, ntr, c, k, h, vn, tr, d);
}

// N = 2, two cells per wavefront, tables through the scalar cache; U: the
// mesh's coefficients are the same at each stencil slot, W: 64-bit offsets
template <bool U, bool W>
__global__ void __launch_bounds__(256)
    k_diag_hex2(MeshView m, int terms, int ntr, const double *__restrict__ h,
                const double *__restrict__ vn, const double *__restrict__ tr, DiagOut d, int swz) {
  const int s = logical_block(blockIdx.x, gridDim.x, swz);
  const int R = blockDim.y;
  const int cA = uniform(s * R + ((int)threadIdx.y & ~1));
  if (cA >= m.nc) return;
  const int cB = min(cA + 1, m.nc - 1);
  const bool hi = (threadIdx.y & 1) != 0;
  if (s * R + (int)threadIdx.y >= m.nc) return;
  const int k = (blockIdx.y * blockDim.x + threadIdx.x) * 2;
  if (m.hexd_ok) {
    diag_union_p<U, W>(m, terms, ntr, cA, cB, hi, k, h, vn, tr, d);
    return;
  }
  if (terms & (T_PV | T_VDEL2 | T_VDEL4)) {
    vertex_diag_p<U, W>(m, 2 * cA, 2 * cB, hi, k, h, vn, d);
    vertex_diag_p<U, W>(m, 2 * cA + 1, 2 * cB + 1, hi, k, h, vn, d);
  }
  cell_diag_p<U, W>(m, terms, ntr, cA, cB, hi, k, h, vn, tr, d);
}

template <int N>
__global__ void __launch_bounds__(256)
    k_tend_hex(MeshView m, ModelCoef mc, StageArgs a, DiagOut d, int swz) {
  const int s = logical_block(blockIdx.x, gridDim.x, swz);
  const int R = blockDim.y;
  const int c = ent<N>(s * R + (int)threadIdx.y);
  if (c >= m.nc) return;
  const int k = (blockIdx.y * blockDim.x + threadIdx.x) * N;
  const int e0 = 3 * c;
  edge_tend_hex<N, 0>(m, mc, a, d, e0, k);
  edge_tend_hex<N, 1>(m, mc, a, d, e0 + 1, k);
  edge_tend_hex<N, 2>(m, mc, a, d, e0 + 2, k);
  cell_tend_hex<N>(m, mc, a, d, c, k);
}

// N = 2 with two cells per wavefront and their tables read through the
// scalar path (hex_ops.hpp: edge_tend_hex2 / cell_tend_hex2); the block's
// row count is even, so the cells of one wavefront are cA and cA + 1
template <bool U, bool W>
__global__ void __launch_bounds__(256)
    k_tend_hex2(MeshView m, ModelCoef mc, StageArgs a, DiagOut d, int swz) {
  const int s = logical_block(blockIdx.x, gridDim.x, swz);
  const int R = blockDim.y;
  const int cA = uniform(s * R + ((int)threadIdx.y & ~1));
  if (cA >= m.nc) return;
  const int cB = min(cA + 1, m.nc - 1);
  const bool hi = (threadIdx.y & 1) != 0;
  if (s * R + (int)threadIdx.y >= m.nc) return;
  const int k = (blockIdx.y * blockDim.x + threadIdx.x) * 2;
  edge_tend_hex2<0, U, W>(m, mc, a, d, 3 * cA, 3 * cB, hi, k);
  edge_tend_hex2<1, U, W>(m, mc, a, d, 3 * cA + 1, 3 * cB + 1, hi, k);
  edge_tend_hex2<2, U, W>(m, mc, a, d, 3 * cA + 2, 3 * cB + 2, hi, k);
  cell_tend_hex2<U, W>(m, mc, a, d, cA, cB, hi, k);
}

// vertex-union form: the PV products of the cell's 16 union vertices are
// formed once and shared by its three edges (needs the per-cell union
// records, hexu)
template <bool U, bool W>
__global__ void __launch_bounds__(256)
    k_tend_hex2v(MeshView m, ModelCoef mc, StageArgs a, DiagOut d, int swz) {
  const int s = logical_block(blockIdx.x, gridDim.x, swz);
  const int R = blockDim.y;
  const int cA = uniform(s * R + ((int)threadIdx.y & ~1));
  if (cA >= m.nc) return;
  const int cB = min(cA + 1, m.nc - 1);
  const bool hi = (threadIdx.y & 1) != 0;
  if (s * R + (int)threadIdx.y >= m.nc) return;
  const int k = (blockIdx.y * blockDim.x + threadIdx.x) * 2;
  Vd<2> P[hexu::kNV], Q[hexu::kNV];
  if (a.do_vn && (a.terms & T_PV)) union_pv2<U, W>(m, mc, d, cA, cB, hi, k, P, Q);
  edge_tend_hex2v<0, U, W>(m, mc, a, d, 3 * cA, 3 * cB, hi, k, P, Q);
  edge_tend_hex2v<1, U, W>(m, mc, a, d, 3 * cA + 1, 3 * cB + 1, hi, k, P, Q);
  edge_tend_hex2v<2, U, W>(m, mc, a, d, 3 * cA + 2, 3 * cB + 2, hi, k, P, Q);
  cell_tend_hex2<U, W>(m, mc, a, d, cA, cB, hi, k);
}

// the cell part alone (thickness and tracer tendencies), for launchers that
// evaluate the edges elsewhere
template <int N>
__global__ void __launch_bounds__(256)
    k_tend_hex_cells(MeshView m, ModelCoef mc, StageArgs a, DiagOut d, int swz) {
  const int s = logical_block(blockIdx.x, gridDim.x, swz);
  const int R = blockDim.y;
  const int c = ent<N>(s * R + (int)threadIdx.y);
  if (c >= m.nc) return;
  const int k = (blockIdx.y * blockDim.x + threadIdx.x) * N;
  cell_tend_hex<N>(m, mc, a, d, c, k);
}

// ---------------------------------------------------------------------------
// Union-stencil variant: one wavefront per cell and 64-level chunk (lane =
// level), all three owned edges and the cell itself.  The rows the four
// stencils share are loaded once -- 19 edges, 16 vertices, 14 cells
// (hexu tables) instead of 3 x 10 + 3 x 10 + 3 x 10 + 7 -- and every index
// and coefficient is wavefront-uniform (scalar loads).  Same expression
// sequence as the kernels above, so the same bits.
// ---------------------------------------------------------------------------
// compute_vn_tendency of owned edge K of a cell from its union rows
template <int K>
__device__ __forceinline__ void hexu_edge(const MeshView &m, const ModelCoef &mc,
                                          const StageArgs &a, const DiagOut &d,
                                          const int (&UE)[hexu::kNE], const int (&UV)[hexu::kNV],
                                          const int (&UC)[hexu::kNC],
                                          const double (&hU)[hexu::kNC],
                                          const double (&vU)[hexu::kNE],
                                          const double (&P)[hexu::kNV],
                                          const double (&Q)[hexu::kNV],
                                          const double (&rvU)[hexu::kNV],
                                          const double (&keU)[hexu::kNC],
                                          const double (&dvU)[hexu::kNC], bool pv, bool ke,
                                          bool ssh, bool vd2, bool vd4, int k) {
  using namespace hexu;
  const int L = m.L;
  const int e = UE[K];
  const size_t idx = (size_t)e * L + k;
  if (!a.do_vn) {
    if (a.epi != EPI_STORE) epilogue<1>(a, a.vn, idx, tend_base<1>(a, a.vn, idx));
    return;
  }
  Vd<1> T = tend_base<1>(a, a.vn, idx);
  const double idc = m.inv_dc[e];
  constexpr int o0 = OV[K][0], o1 = OV[K][1];
  constexpr int q0 = OC[K][0], q1 = OC[K][1];
  if (pv) {
    const double qs = 0.5 * (P[o0] + P[o1]) + 0.5 * (Q[o0] + Q[o1]);
    double qt = 0.0;
#pragma unroll
    for (int j = 0; j < 10; ++j) {
      const double wj = m.w[e * 10 + j];
      const double fh = 0.5 * (hU[TC0[K][j]] + hU[TC1[K][j]]);
      const double nrv = 0.5 * (P[TV0[K][j]] + P[TV1[K][j]]);
      const double npv = 0.5 * (Q[TV0[K][j]] + Q[TV1[K][j]]);
      const double nv = (qs + nrv + npv) * 0.5;
      qt = qt + wj * fh * vU[TE[K][j]] * nv;
    }
    T[0] = T[0] + qt;
  }
  if (ke) T[0] = T[0] - (keU[q1] - keU[q0]) * idc;
  if (ssh) T[0] = T[0] - mc.grav * (hU[q1] - hU[q0]) * idc;
  if (vd2 || vd4) {
    const int ml = m.maxlev_e[e];
    const double idv = m.inv_dv[e];
    if (vd2) {
      const double del2u = (dvU[q1] - dvU[q0]) * idc - (rvU[o1] - rvU[o0]) * idv;
      T[0] = T[0] + ((k < ml) ? mc.cv2[e] : 0.0) * del2u;
    }
    if (vd4) {
      const double d0 = d.vdel2_div_c[(size_t)UC[q0] * L + k];
      const double d1 = d.vdel2_div_c[(size_t)UC[q1] * L + k];
      const double r0 = d.vdel2_rv_v[(size_t)UV[o0] * L + k];
      const double r1 = d.vdel2_rv_v[(size_t)UV[o1] * L + k];
      const double del2u = (d1 - d0) * idc - (r1 - r0) * idv;
      T[0] = T[0] - ((k < ml) ? mc.cv4[e] : 0.0) * del2u;
    }
  }
  epilogue<1>(a, a.vn, idx, T);
}

__global__ void __launch_bounds__(256) k_tend_hexu(MeshView m, ModelCoef mc, StageArgs a,
                                                   DiagOut d, int swz) {
  using namespace hexu;
  const int s = logical_block(blockIdx.x, gridDim.x, swz);
  const int c = uniform(s * 4 + (int)(threadIdx.x >> 6));
  if (c >= m.nc) return;
  const int L = m.L, me = m.me, NC = m.nc;
  const int k = blockIdx.y * 64 + (int)(threadIdx.x & 63);
  const int *rec = m.hexu + (size_t)c * kRec;
  int UE[kNE], UV[kNV], UC[kNC];
#pragma unroll
  for (int x = 0; x < kNE; ++x) UE[x] = rec[x];
#pragma unroll
  for (int x = 0; x < kNV; ++x) UV[x] = rec[kNE + x];
#pragma unroll
  for (int x = 0; x < kNC; ++x) UC[x] = rec[kNE + kNV + x];
  const double *__restrict__ h = a.h.in;
  const double *__restrict__ vn = a.vn.in;
  const bool pv = a.do_vn && (a.terms & T_PV), ke = a.do_vn && (a.terms & T_KE);
  const bool ssh = a.do_vn && (a.terms & T_SSH), vd2 = a.do_vn && (a.terms & T_VDEL2);
  const bool vd4 = a.do_vn && (a.terms & T_VDEL4);
  const bool cellp = a.do_h || a.ntr > 0;

  // ---- rows (each once)
  double hU[kNC], vU[kNE], P[kNV], Q[kNV], rvU[kNV], keU[kNC], dvU[kNC];
#pragma unroll
  for (int x = 0; x < kNC; ++x) hU[x] = h[(size_t)UC[x] * L + k];
  if (pv) {
#pragma unroll
    for (int x = 0; x < kNE; ++x) vU[x] = vn[(size_t)UE[x] * L + k];
#pragma unroll
    for (int x = 0; x < kNV; ++x) {
      rvU[x] = d.rvort_v[(size_t)UV[x] * L + k];
      const double it = d.inv_thick_v[(size_t)UV[x] * L + k];
      P[x] = it * rvU[x];
      Q[x] = it * mc.f_v[UV[x]];
    }
  } else {
    if (cellp) {
#pragma unroll
      for (int j = 0; j < 6; ++j) vU[CE[j]] = vn[(size_t)UE[CE[j]] * L + k];
    }
    if (vd2) {
#pragma unroll
      for (int K = 0; K < 3; ++K)
#pragma unroll
        for (int q = 0; q < 2; ++q) rvU[OV[K][q]] = d.rvort_v[(size_t)UV[OV[K][q]] * L + k];
    }
  }
  if (ke || vd2) {
#pragma unroll
    for (int K = 0; K < 3; ++K)
#pragma unroll
      for (int q = 0; q < 2; ++q) {
        const int u = OC[K][q];
        if (ke) keU[u] = d.ke_c[(size_t)UC[u] * L + k];
        if (vd2) dvU[u] = d.div_c[(size_t)UC[u] * L + k];
      }
  }

  // ---- the three owned edges (compute_vn_tendency)
  hexu_edge<0>(m, mc, a, d, UE, UV, UC, hU, vU, P, Q, rvU, keU, dvU, pv, ke, ssh, vd2, vd4, k);
  hexu_edge<1>(m, mc, a, d, UE, UV, UC, hU, vU, P, Q, rvU, keU, dvU, pv, ke, ssh, vd2, vd4, k);
  hexu_edge<2>(m, mc, a, d, UE, UV, UC, hU, vU, P, Q, rvU, keU, dvU, pv, ke, ssh, vd2, vd4, k);

  // ---- the cell (compute_h_tendency, compute_tr_tendency)
  if (!cellp) {
    if (a.epi != EPI_STORE) {
      const size_t hidx = (size_t)c * L + k;
      epilogue<1>(a, a.h, hidx, tend_base<1>(a, a.h, hidx));
    }
    return;
  }
  double cd[6], fh[6];
#pragma unroll
  for (int j = 0; j < 6; ++j) {
    cd[j] = m.cdiv[c * me + j];
    fh[j] = 0.5 * (hU[CC0[j]] + hU[CC1[j]]);
  }
  const size_t hidx = (size_t)c * L + k;
  if (a.do_h) {
    Vd<1> T = tend_base<1>(a, a.h, hidx);
    double acc = 0.0;
#pragma unroll
    for (int j = 0; j < 6; ++j) acc = acc - cd[j] * fh[j] * vU[CE[j]];
    T[0] = T[0] + acc;
    epilogue<1>(a, a.h, hidx, T);
  } else if (a.epi != EPI_STORE) {
    epilogue<1>(a, a.h, hidx, tend_base<1>(a, a.h, hidx));
  }
  const int ntr = a.ntr;
  if (ntr <= 0) return;
  const bool td2 = a.terms & T_TDEL2, td4 = a.terms & T_TDEL4;
  double inv[kNT], idc[6], c2[6], s2[6], c4[6];
#pragma unroll
  for (int p = 0; p < kNT; ++p) inv[p] = 1.0 / hU[TCELL[p]];
#pragma unroll
  for (int j = 0; j < 6; ++j) {
    const int e = UE[CE[j]];
    idc[j] = m.inv_dc[e];
    c2[j] = td2 ? mc.ct2[c * me + j] : 0.0;
    s2[j] = td2 ? m.s2[e] : 0.0;
    c4[j] = td4 ? mc.ct4[c * me + j] : 0.0;
  }
  for (int l = 0; l < ntr; ++l) {
    const double *trl = a.tr.in + (size_t)l * NC * L;
    double tP[kNT], qP[kNT];
#pragma unroll
    for (int p = 0; p < kNT; ++p) tP[p] = trl[(size_t)UC[TCELL[p]] * L + k];
    if (td4) {
      const double *dl = d.trdel2_c + (size_t)l * NC * L;
#pragma unroll
      for (int p = 0; p < kNT; ++p) qP[p] = dl[(size_t)UC[TCELL[p]] * L + k];
    }
    double adv = 0.0, dif = 0.0, hyp = 0.0;
#pragma unroll
    for (int j = 0; j < 6; ++j) {
      const int p0 = tpos(CC0[j]), p1 = tpos(CC1[j]);
      const double n0 = tP[p0] * inv[p0], n1 = tP[p1] * inv[p1];
      const double ntr_e = (n0 + n1) * 0.5;
      adv = adv - cd[j] * fh[j] * ntr_e * vU[CE[j]];
      if (td2) {
        const double grad = (n1 - n0) * idc[j];
        dif = dif + c2[j] * fh[j] * s2[j] * grad;
      }
      if (td4) {
        const double grad = (qP[p1] - qP[p0]) * idc[j];
        hyp = hyp - c4[j] * grad;
      }
    }
    const size_t tidx = ((size_t)l * NC + c) * L + k;
    Vd<1> T = tend_base<1>(a, a.tr, tidx);
    T[0] = T[0] + adv;
    if (td2) T[0] = T[0] + dif;
    if (td4) T[0] = T[0] + hyp;
    epilogue<1>(a, a.tr, tidx, T);
  }
}

// lanes per row = 64 / N (one row covers 64 levels); blockIdx.y walks the
// 64-level chunks
// the two-cells-per-wavefront kernels need rows of 64 levels
bool hex_pair(const MeshView &m, int N) { return N == 2 && m.L % 64 == 0; }

// One cell per row of lanes covering 64 levels (L a multiple of 64) or all
// L levels (L = 16 or 32: 4 or 8 cells per wavefront, N = 2).
dim3 hex_grid(const MeshView &m, const Launch &lc, int N, dim3 &block) {
  const int span = m.L % 64 == 0 ? 64 : m.L;
  const int lanes = span / N;
  int rows = lc.rows > 0 ? lc.rows : 256 / lanes;
  if (rows * lanes > 256) rows = 256 / lanes;
  // pair kernels: an even row count keeps the two cells of a wavefront adjacent
  if (hex_pair(m, N) && rows % 2) rows = rows > 1 ? rows - 1 : 2;
  block = dim3(lanes, rows);
  return dim3((m.nc + rows - 1) / rows, m.L / span);
}

int hex_n(const MeshView &m, const Launch &lc) { return lc.lpt == 1 ? 1 : 2; }

} // namespace

bool hex_usable(const MeshView &m, const Launch &lc) {
  // 16 or 32 layers: several cells per wavefront, two levels per lane
  const bool rows_ok = m.L % 64 == 0 || ((m.L == 16 || m.L == 32) && lc.lpt != 1);
  return m.hex_ok && rows_ok && lc.variant != 1;
}

bool hexu_usable(const MeshView &m, const Launch &lc) {
  return m.hexu_ok && (m.L % 64 == 0) && lc.variant == 3;
}

void launch_tend_hexu(const MeshView &m, const Launch &lc, const ModelCoef &mc,
                      const StageArgs &a, const DiagOut &d, hipStream_t s) {
  TimerScope ts("tend", s);
  const dim3 grid((m.nc + 3) / 4, m.L / 64);
  hipLaunchKernelGGL(k_tend_hexu, grid, dim3(256), 0, s, m, mc, a, d, lc.swizzle);
  OMEGA_HIP_CHECK(hipGetLastError());
}

void launch_diag_hex(const MeshView &m, const Launch &lc, int terms, int ntr, const Real *h,
                     const Real *vn, const Real *tr, const DiagOut &d, hipStream_t s) {
  TimerScope ts("diag", s);
  dim3 block;
  const int N = hex_n(m, lc);
  const dim3 grid = hex_grid(m, lc, N, block);
  if (hex_pair(m, N)) {
    // the uniform-coefficient and 32-bit-offset forms give the same bits
    auto kern = m.coef_uni ? (m.off32 ? k_diag_hex2<true, false> : k_diag_hex2<true, true>)
                           : (m.off32 ? k_diag_hex2<false, false> : k_diag_hex2<false, true>);
    hipLaunchKernelGGL(kern, grid, block, 0, s, m, terms, ntr, h, vn, tr, d, lc.swizzle);
  } else if (N == 2)
    hipLaunchKernelGGL(k_diag_hex<2>, grid, block, 0, s, m, terms, ntr, h, vn, tr, d, lc.swizzle);
  else
    hipLaunchKernelGGL(k_diag_hex<1>, grid, block, 0, s, m, terms, ntr, h, vn, tr, d, lc.swizzle);
  OMEGA_HIP_CHECK(hipGetLastError());
}

void launch_tend_hex_cells(const MeshView &m, const Launch &lc, const ModelCoef &mc,
                           const StageArgs &a, const DiagOut &d, hipStream_t s) {
  dim3 block;
  const int N = hex_n(m, lc);
  const dim3 grid = hex_grid(m, lc, N, block);
  if (N == 2)
    hipLaunchKernelGGL(k_tend_hex_cells<2>, grid, block, 0, s, m, mc, a, d, lc.swizzle);
  else
    hipLaunchKernelGGL(k_tend_hex_cells<1>, grid, block, 0, s, m, mc, a, d, lc.swizzle);
  OMEGA_HIP_CHECK(hipGetLastError());
}

void launch_tend_hex(const MeshView &m, const Launch &lc, const ModelCoef &mc, const StageArgs &a,
                     const DiagOut &d, hipStream_t s) {
  TimerScope ts("tend", s);
  dim3 block;
  const int N = hex_n(m, lc);
  const dim3 grid = hex_grid(m, lc, N, block);
  const bool U = m.coef_uni && mc.uniform;
  if (hex_pair(m, N) && m.hexu_ok) {
    auto kern = U ? (m.off32 ? k_tend_hex2v<true, false> : k_tend_hex2v<true, true>)
                  : (m.off32 ? k_tend_hex2v<false, false> : k_tend_hex2v<false, true>);
    hipLaunchKernelGGL(kern, grid, block, 0, s, m, mc, a, d, lc.swizzle);
  } else if (hex_pair(m, N)) {
    auto kern = U ? (m.off32 ? k_tend_hex2<true, false> : k_tend_hex2<true, true>)
                  : (m.off32 ? k_tend_hex2<false, false> : k_tend_hex2<false, true>);
    hipLaunchKernelGGL(kern, grid, block, 0, s, m, mc, a, d, lc.swizzle);
  } else if (N == 2)
    hipLaunchKernelGGL(k_tend_hex<2>, grid, block, 0, s, m, mc, a, d, lc.swizzle);
  else
    hipLaunchKernelGGL(k_tend_hex<1>, grid, block, 0, s, m, mc, a, d, lc.swizzle);
  OMEGA_HIP_CHECK(hipGetLastError());
}

} // namespace omega_hip
