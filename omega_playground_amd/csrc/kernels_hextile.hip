// LDS-tiled vn tendency for the regular hexagonal mesh.
//
// Block (tile, level chunk): a tile is `tcells` consecutive device cells
// (a compact patch along the space-filling curve) and the 3*tcells edges
// they own; a level chunk is K consecutive levels.  The block copies into
// LDS, once each and K levels wide, every row the owned edges' PV stencils
// read: the velocities of the tile's union edges, the thicknesses of its
// union cells and, per union vertex, the two PV products
// inv_thick*rel_vort and inv_thick*f.  Each thread then evaluates the three
// owned edges of one cell from LDS, with the stencil slots resolved by the
// compile-time pattern of hex_pattern.hpp and the row positions by one-byte
// per-edge records.  Gradient and del2 terms read their two rows from
// global memory.  Same expression sequence as hexops::edge_tend_hex, so the
// same bits.
//
// Staged rows replace ~40 cache-hierarchy row loads per edge with ~4; the
// stencil reads go through LDS instead of the vector-memory address path,
// which is what bounds the register-gather kernels.
#include "hex_ops.hpp"
#include "tiles.hpp"

namespace omega_hip {

namespace {

using namespace dev;

constexpr int kHT = 256; // threads per block

// one byte of the 32-byte record held in w[0..7]
__device__ __forceinline__ int rbyte(const uint32_t (&w)[8], int b) {
  return (int)((w[b >> 2] >> (8 * (b & 3))) & 0xffu);
}

// Per-tile edge metadata staged in LDS next to the rows (owned edge o of
// the tile): stencil record, weights, 1/dc, 1/dv, del2 coefficient,
// max level.
struct EdgeMeta {
  uint4 *rec;  // [3T][2]
  double *w;   // [3T][10]
  double *idc, *idv, *cv2; // [3T]
  int *ml;     // [3T]
};

template <int N, int K, int KC>
__device__ __forceinline__ void hextile_edge(const MeshView &m, const ModelCoef &mc,
                                             const StageArgs &a, const DiagOut &d,
                                             const EdgeMeta &em, const double *sv,
                                             const double *sh, const double *sp,
                                             const double *sq, int e, int o, int k, int lo,
                                             bool pv) {
  const int L = m.L;
  const size_t idx = (size_t)e * L + k;
  Vd<N> T = tend_base<N>(a, a.vn, idx);
  const double idc = em.idc[o];
  constexpr int o0 = hex::OWNV[KC][0], o1 = hex::OWNV[KC][1];
  constexpr int q0 = hex::OWNC[KC][0], q1 = hex::OWNC[KC][1];
  const uint4 ra = em.rec[2 * o], rb = em.rec[2 * o + 1];
  const uint32_t rw[8] = {ra.x, ra.y, ra.z, ra.w, rb.x, rb.y, rb.z, rb.w};
  // cells of the edge itself (thickness rows for the SSH gradient)
  const Vd<N> hq0 = ld<N>(sh + rbyte(rw, 20 + q0) * K + lo);
  const Vd<N> hq1 = ld<N>(sh + rbyte(rw, 20 + q1) * K + lo);
  if (pv) {
    Vd<N> hU[10], vJ[10], P[10], Q[10];
#pragma unroll
    for (int u = 0; u < 10; ++u) {
      hU[u] = ld<N>(sh + rbyte(rw, 20 + u) * K + lo);
      P[u] = ld<N>(sp + rbyte(rw, 10 + u) * K + lo);
      Q[u] = ld<N>(sq + rbyte(rw, 10 + u) * K + lo);
    }
#pragma unroll
    for (int j = 0; j < 10; ++j) vJ[j] = ld<N>(sv + rbyte(rw, j) * K + lo);
    double w[10];
#pragma unroll
    for (int j = 0; j < 10; ++j) w[j] = em.w[o * 10 + j];
#pragma unroll
    for (int i = 0; i < N; ++i) {
      const double qs = 0.5 * (P[o0][i] + P[o1][i]) + 0.5 * (Q[o0][i] + Q[o1][i]);
      double qt = 0.0;
#pragma unroll
      for (int j = 0; j < 10; ++j) {
        const double fh = 0.5 * (hU[hex::LC0[KC][j]][i] + hU[hex::LC1[KC][j]][i]);
        const double nrv = 0.5 * (P[hex::LV0[KC][j]][i] + P[hex::LV1[KC][j]][i]);
        const double npv = 0.5 * (Q[hex::LV0[KC][j]][i] + Q[hex::LV1[KC][j]][i]);
        const double nv = (qs + nrv + npv) * 0.5;
        qt = qt + w[j] * fh * vJ[j][i] * nv;
      }
      T[i] = T[i] + qt;
    }
  }
  const int2 cc = reinterpret_cast<const int2 *>(m.coe)[e];
  if (a.terms & T_KE) {
    const Vd<N> k0 = ld<N>(d.ke_c + (size_t)cc.x * L + k), k1 = ld<N>(d.ke_c + (size_t)cc.y * L + k);
#pragma unroll
    for (int i = 0; i < N; ++i) T[i] = T[i] - (k1[i] - k0[i]) * idc;
  }
  if (a.terms & T_SSH) {
#pragma unroll
    for (int i = 0; i < N; ++i) T[i] = T[i] - mc.grav * (hq1[i] - hq0[i]) * idc;
  }
  if (a.terms & (T_VDEL2 | T_VDEL4)) {
    const int ml = em.ml[o];
    const double idv = em.idv[o];
    const int2 vv = reinterpret_cast<const int2 *>(m.voe)[e];
    if (a.terms & T_VDEL2) {
      const Vd<N> d0 = ld<N>(d.div_c + (size_t)cc.x * L + k), d1 = ld<N>(d.div_c + (size_t)cc.y * L + k);
      const Vd<N> r0 = ld<N>(d.rvort_v + (size_t)vv.x * L + k);
      const Vd<N> r1 = ld<N>(d.rvort_v + (size_t)vv.y * L + k);
      const double cv = em.cv2[o];
#pragma unroll
      for (int i = 0; i < N; ++i) {
        const double del2u = (d1[i] - d0[i]) * idc - (r1[i] - r0[i]) * idv;
        T[i] = T[i] + ((k + i < ml) ? cv : 0.0) * del2u;
      }
    }
    if (a.terms & T_VDEL4) {
      const Vd<N> d0 = ld<N>(d.vdel2_div_c + (size_t)cc.x * L + k);
      const Vd<N> d1 = ld<N>(d.vdel2_div_c + (size_t)cc.y * L + k);
      const Vd<N> r0 = ld<N>(d.vdel2_rv_v + (size_t)vv.x * L + k);
      const Vd<N> r1 = ld<N>(d.vdel2_rv_v + (size_t)vv.y * L + k);
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

template <int N, int K>
__global__ void __launch_bounds__(kHT)
    k_hextile_edges(MeshView m, TileView t, ModelCoef mc, StageArgs a, DiagOut d, int swz) {
  extern __shared__ double lds[];
  constexpr int LR = K / N;      // lanes per row segment
  constexpr int RB = kHT / LR;   // row groups per block
  const int tile = logical_block(blockIdx.x, gridDim.x, swz);
  const int L = m.L;
  const int lane = (int)threadIdx.x % LR, grp = (int)threadIdx.x / LR;
  const int lo = lane * N, k = blockIdx.y * K + lo;
  const int c0 = tile * t.tcells, nco = min(t.tcells, m.nc - c0);
  const int e0 = 3 * c0;

  if (!a.do_vn) {
    if (a.epi == EPI_STORE) return;
    for (int i = grp; i < 3 * nco; i += RB) {
      const size_t idx = (size_t)(e0 + i) * L + k;
      epilogue<N>(a, a.vn, idx, tend_base<N>(a, a.vn, idx));
    }
    return;
  }
  const bool pv = a.terms & T_PV;
  const int ue0 = t.ue_ofs[tile], nue = t.ue_ofs[tile + 1] - ue0;
  const int uc0 = t.uc_ofs[tile], nuc = t.uc_ofs[tile + 1] - uc0;
  const int uv0 = t.uv_ofs[tile], nuv = t.uv_ofs[tile + 1] - uv0;
  double *sv = lds;                       // [max_ue][K] edge velocities
  double *sh = sv + t.max_ue * K;         // [max_uc][K] cell thicknesses
  double *sp = sh + t.max_uc * K;         // [max_uv][K] inv_thick * rel_vort
  double *sq = sp + t.max_uv * K;         // [max_uv][K] inv_thick * f
  const int T3 = 3 * t.tcells, neo = 3 * nco;
  EdgeMeta em; // records first: the row block is a multiple of 16 bytes
  em.rec = reinterpret_cast<uint4 *>(sq + t.max_uv * K);
  em.w = reinterpret_cast<double *>(em.rec + 2 * T3);
  em.idc = em.w + 10 * T3;
  em.idv = em.idc + T3;
  em.cv2 = em.idv + T3;
  em.ml = reinterpret_cast<int *>(em.cv2 + T3);

  // ---- staging: edge metadata (contiguous per tile)
  for (int i = (int)threadIdx.x; i < 2 * neo; i += kHT) em.rec[i] = t.hexl[2 * (size_t)e0 + i];
  if (pv)
    for (int i = (int)threadIdx.x; i < 10 * neo; i += kHT) em.w[i] = m.w[10 * (size_t)e0 + i];
  for (int i = (int)threadIdx.x; i < neo; i += kHT) {
    em.idc[i] = m.inv_dc[e0 + i];
    em.idv[i] = m.inv_dv[e0 + i];
    em.cv2[i] = mc.cv2[e0 + i];
    em.ml[i] = m.maxlev_e[e0 + i];
  }
  // ---- staging: velocity and thickness rows (loads first, then stores)
  {
    const int nrow = (pv ? nue : 0) + nuc, skip = pv ? 0 : nue;
    constexpr int U = 8;
    for (int r0 = grp; r0 < nrow; r0 += U * RB) {
      Vd<N> v[U];
      int dst[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int r = min(r0 + u * RB, nrow - 1) + skip;
        const bool isv = r < nue;
        const int id = isv ? t.ue[ue0 + r] : t.uc[uc0 + r - nue];
        const double *src = isv ? a.vn.in : a.h.in;
        v[u] = ld<N>(src + (size_t)id * L + k);
        dst[u] = (isv ? r : t.max_ue + (r - nue)) * K + lo;
      }
#pragma unroll
      for (int u = 0; u < U; ++u)
        if (r0 + u * RB < nrow) st<N>(lds + dst[u], v[u]);
    }
  }
  // ---- staging: PV products per union vertex (vertex_diag results)
  if (pv) {
    constexpr int U = 4;
    for (int r0 = grp; r0 < nuv; r0 += U * RB) {
      Vd<N> it[U], rv[U];
      double f[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int r = min(r0 + u * RB, nuv - 1);
        const int id = t.uv[uv0 + r];
        it[u] = ld<N>(d.inv_thick_v + (size_t)id * L + k);
        rv[u] = ld<N>(d.rvort_v + (size_t)id * L + k);
        f[u] = mc.f_v[id];
      }
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int r = r0 + u * RB;
        if (r < nuv) {
          Vd<N> P, Q;
#pragma unroll
          for (int i = 0; i < N; ++i) {
            P[i] = it[u][i] * rv[u][i];
            Q[i] = it[u][i] * f[u];
          }
          st<N>(sp + r * K + lo, P);
          st<N>(sq + r * K + lo, Q);
        }
      }
    }
  }
  __syncthreads();

  // ---- the three owned edges of each cell
  for (int i = grp; i < nco; i += RB) {
    const int e = e0 + 3 * i;
    hextile_edge<N, K, 0>(m, mc, a, d, em, sv, sh, sp, sq, e, 3 * i, k, lo, pv);
    hextile_edge<N, K, 1>(m, mc, a, d, em, sv, sh, sp, sq, e + 1, 3 * i + 1, k, lo, pv);
    hextile_edge<N, K, 2>(m, mc, a, d, em, sv, sh, sp, sq, e + 2, 3 * i + 2, k, lo, pv);
  }
}

size_t hextile_lds_bytes(const TileView &t, int K) {
  const size_t T3 = 3 * (size_t)t.tcells;
  return (size_t)(t.max_ue + t.max_uc + 2 * t.max_uv) * K * sizeof(double) +
         T3 * (13 * sizeof(double) + 2 * sizeof(uint4) + sizeof(int));
}

size_t lds_limit_bytes() {
  int dev = 0, bytes = 0;
  OMEGA_HIP_CHECK(hipGetDevice(&dev));
  OMEGA_HIP_CHECK(hipDeviceGetAttribute(&bytes, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
  return (size_t)bytes;
}

template <int N, int K>
void launch_k(const MeshView &m, const TileView &t, const Launch &lc, const ModelCoef &mc,
              const StageArgs &a, const DiagOut &d, hipStream_t s) {
  const size_t lds = hextile_lds_bytes(t, K);
  OMEGA_REQUIRE(lds <= lds_limit_bytes(),
                "hexagonal tiled kernel: tile too large for LDS (use fewer cells per tile)");
  OMEGA_REQUIRE(t.tcells <= kHT / (K / N),
                "hexagonal tiled kernel: at most 256/(levels_per_block/levels_per_thread) cells "
                "per tile");
  const dim3 grid(t.ntiles, m.L / K);
  hipLaunchKernelGGL((k_hextile_edges<N, K>), grid, dim3(kHT), lds, s, m, t, mc, a, d,
                     lc.swizzle);
  OMEGA_HIP_CHECK(hipGetLastError());
}

} // namespace

bool hextile_usable(const MeshView &m, const TileTables &t, const Launch &lc) {
  return lc.variant == 4 && m.hex_ok && t.hexl_ok && m.L % 64 == 0;
}

void launch_tend_hextile(const MeshView &m, const TileView &t, const Launch &lc,
                         const ModelCoef &mc, const StageArgs &a, const DiagOut &d,
                         hipStream_t s) {
  TimerScope ts("tend", s);
  OMEGA_REQUIRE(t.hexl != nullptr, "hexagonal tiled kernel: tile records missing");
  {
    TimerScope te("tend_edges", s);
    if (lc.tile_k == 32) launch_k<2, 32>(m, t, lc, mc, a, d, s);
    else if (lc.tile_k == 8) launch_k<1, 8>(m, t, lc, mc, a, d, s);
    else launch_k<2, 16>(m, t, lc, mc, a, d, s);
  }
  TimerScope tc("tend_cells", s);
  launch_tend_hex_cells(m, lc, mc, a, d, s);
}

} // namespace omega_hip
