// LDS-tiled tendency + stage-update kernels (any MPAS mesh).
//
// Block (tile, level chunk): a tile is `tcells` consecutive device cells
// and the edges they own; a level chunk is K consecutive levels.  The block
// first copies into LDS every state/diagnostic row its stencils read --
// each once, K levels wide -- and its connectivity, edge weights and metric
// factors (contiguous per tile, so those copies are coalesced), then
// evaluates the tendencies of its owned entities from LDS only and applies
// the stage epilogue in global memory.  The stencil arithmetic is the same
// expression sequence as the general kernels (fwd_ops.hpp), so the results
// are the same bits; only where the operands come from differs.
//
// Rows are stored with a stride of K+2 doubles so that the rows read by
// one wavefront start in different LDS banks.
//
// Two kernels per stage: the edge kernel (vn tendency: PV flux, KE and SSH
// gradients, del2/del4) and the cell kernel (thickness and tracer
// tendencies, tracers staged four at a time).  The diagnostics the edge
// kernel reads (vertex vorticity and 1/thickness, cell KE and divergence,
// tracer Laplacians) come from the diagnostics kernel of the stage.
#include "device_common.hpp"
#include "tiles.hpp"

namespace omega_hip {

namespace {

using namespace dev;

constexpr int kThreads = 256;

// Row staging: up to six row lists (src[ids[r]][k..k+N) -> LDS row r of
// the list) handled as one concatenated list, eight rows per lane in
// flight, so a block's whole staging costs about two memory round trips.
// The list table lives in LDS so that rows can pick their list without
// per-lane pointer selects.
struct RowSeg {
  const double *src;
  const int *ids;
  int doff; // LDS offset of the list's first row, in doubles
  int n;
};

struct SegTable {
  const double *src[6];
  const int *ids[6];
  int doff[6];
  int cum[7];
};

template <int N, int KS>
__device__ __forceinline__ void stage_rows(SegTable &tab, double *lds, const RowSeg s0,
                                           const RowSeg s1, const RowSeg s2, const RowSeg s3,
                                           const RowSeg s4, const RowSeg s5, int L, int k, int lo,
                                           int row0, int rstep) {
  if (threadIdx.x == 0) {
    const RowSeg sg[6] = {s0, s1, s2, s3, s4, s5};
    int acc = 0;
#pragma unroll
    for (int q = 0; q < 6; ++q) {
      tab.src[q] = sg[q].src;
      tab.ids[q] = sg[q].ids;
      tab.doff[q] = sg[q].doff;
      tab.cum[q] = acc;
      acc += sg[q].n;
    }
    tab.cum[6] = acc;
  }
  __syncthreads();
  const int tot = tab.cum[6];
  if (k < L && tot > 0) {
    // loads are unconditional (rows past the end re-read the last row) so
    // that all of them are issued before the first wait
    constexpr int U = 8;
    for (int r0 = row0; r0 < tot; r0 += U * rstep) {
      int off[U];
      Vd<N> v[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int r = min(r0 + u * rstep, tot - 1);
        int q = 0;
#pragma unroll
        for (int z = 1; z < 6; ++z) q += r >= tab.cum[z];
        const int rr = r - tab.cum[q];
        off[u] = tab.doff[q] + rr * KS + lo;
        v[u] = ld<N>(tab.src[q] + (size_t)tab.ids[q][rr] * L + k);
      }
#pragma unroll
      for (int u = 0; u < U; ++u)
        if (r0 + u * rstep < tot) st<N>(lds + off[u], v[u]);
    }
  }
  __syncthreads(); // the table is reused by the next staging call
}

constexpr RowSeg kNoRows{nullptr, nullptr, 0, 0};

// Flat copies, four elements per lane in flight.
template <class T>
__device__ __forceinline__ void copy_flat(T *__restrict__ dst, const T *__restrict__ src, int n) {
  for (int i0 = (int)threadIdx.x; i0 < n; i0 += 4 * kThreads) {
    T v[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) v[u] = src[min(i0 + u * kThreads, n - 1)];
#pragma unroll
    for (int u = 0; u < 4; ++u)
      if (i0 + u * kThreads < n) dst[i0 + u * kThreads] = v[u];
  }
}

template <int N, int KS>
__device__ __forceinline__ Vd<N> lrow(const double *__restrict__ s, int r, int lo) {
  return ld<N>(s + r * KS + lo);
}

// LDS carve-up of the edge kernel (doubles first, then int4, then ints)
struct EdgeLds {
  double *vn, *h, *rv, *inv, *ke, *div, *f, *w, *idc, *idv, *cv2, *cv4;
  int4 *cv;
  int *eoe, *noe, *ml;
};

template <int KS>
__device__ __forceinline__ EdgeLds edge_lds(double *base, const TileView &t, int me2) {
  EdgeLds s;
  double *p = base;
  s.vn = p;  p += t.max_ue * KS;
  s.h = p;   p += t.max_uc * KS;
  s.rv = p;  p += t.max_uv * KS;
  s.inv = p; p += t.max_uv * KS;
  s.ke = p;  p += t.max_nkc * KS;
  s.div = p; p += t.max_nkc * KS;
  s.f = p;   p += t.max_uv;
  s.w = p;   p += t.max_eo * me2;
  s.idc = p; p += t.max_eo;
  s.idv = p; p += t.max_eo;
  s.cv2 = p; p += t.max_eo;
  s.cv4 = p; p += t.max_eo;
  p += (p - base) & 1; // 16-byte alignment of the int4 table
  s.cv = reinterpret_cast<int4 *>(p);
  int *q = reinterpret_cast<int *>(s.cv + t.max_ue);
  s.eoe = q; q += t.max_eo * me2;
  s.noe = q; q += t.max_eo;
  s.ml = q;
  return s;
}

size_t edge_lds_bytes(const TileView &t, int KS, int me2) {
  const size_t d = (size_t)(t.max_ue + t.max_uc + 2 * t.max_uv + 2 * t.max_nkc) * KS + t.max_uv +
                   (size_t)t.max_eo * me2 + 4 * (size_t)t.max_eo;
  return (d + 1) * sizeof(double) + (size_t)t.max_ue * sizeof(int4) +
         ((size_t)t.max_eo * me2 + 2 * (size_t)t.max_eo) * sizeof(int);
}

// ---------------------------------------------------------------------------
// Edge kernel
// ---------------------------------------------------------------------------
template <int N, int K>
__global__ void __launch_bounds__(kThreads)
    k_tile_edges(MeshView m, TileView t, ModelCoef mc, StageArgs a, DiagOut d, int swz) {
  extern __shared__ double lds[];
  constexpr int LR = K / N;         // lanes per row
  constexpr int RB = kThreads / LR; // rows per pass
  constexpr int KS = K + 2;         // padded row stride
  const int tile = logical_block(blockIdx.x, gridDim.x, swz);
  const int L = m.L, me2 = m.me2;
  const int lane = (int)threadIdx.x % LR, row0 = (int)threadIdx.x / LR;
  const int lo = lane * N, k = blockIdx.y * K + lo;
  const int c0 = tile * t.tcells, c1 = min(c0 + t.tcells, m.nc);
  const int e0 = m.eofs[c0], neo = m.eofs[c1] - e0;

  if (!a.do_vn) {
    if (a.epi == EPI_STORE || k >= L) return;
    for (int i = row0; i < neo; i += RB) {
      const size_t idx = (size_t)(e0 + i) * L + k;
      epilogue<N>(a, a.vn, idx, tend_base<N>(a, a.vn, idx));
    }
    return;
  }

  const int ue0 = t.ue_ofs[tile], nue = t.ue_ofs[tile + 1] - ue0;
  const int uc0 = t.uc_ofs[tile], nuc = t.uc_ofs[tile + 1] - uc0;
  const int uv0 = t.uv_ofs[tile], nuv = t.uv_ofs[tile + 1] - uv0;
  const int nkc = t.nkc[tile];
  const EdgeLds s = edge_lds<KS>(lds, t, me2);
  const bool pv = a.terms & T_PV, ke = a.terms & T_KE, ssh = a.terms & T_SSH;
  const bool vd2 = a.terms & T_VDEL2, vd4 = a.terms & T_VDEL4;

  // ---- stage rows, connectivity and metric factors
  __shared__ SegTable tab;
  auto off = [&](const double *p) { return (int)(p - lds); };
  stage_rows<N, KS>(tab, lds, pv ? RowSeg{a.vn.in, t.ue + ue0, off(s.vn), nue} : kNoRows,
                    pv ? RowSeg{d.inv_thick_v, t.uv + uv0, off(s.inv), nuv} : kNoRows,
                    (pv || ssh) ? RowSeg{a.h.in, t.uc + uc0, off(s.h), pv ? nuc : nkc} : kNoRows,
                    (pv || vd2) ? RowSeg{d.rvort_v, t.uv + uv0, off(s.rv), nuv} : kNoRows,
                    ke ? RowSeg{d.ke_c, t.uc + uc0, off(s.ke), nkc} : kNoRows,
                    vd2 ? RowSeg{d.div_c, t.uc + uc0, off(s.div), nkc} : kNoRows, L, k, lo, row0,
                    RB);
  if (pv) {
    for (int r = (int)threadIdx.x; r < nuv; r += kThreads) s.f[r] = mc.f_v[t.uv[uv0 + r]];
    copy_flat(s.w, m.w + (size_t)e0 * me2, neo * me2);
    copy_flat(s.eoe, t.eoe_l + (size_t)e0 * me2, neo * me2);
    copy_flat(s.noe, m.noe + e0, neo);
  }
  copy_flat(s.cv, t.ue_cv + ue0, nue);
  copy_flat(s.idc, m.inv_dc + e0, neo);
  if (vd2 || vd4) {
    copy_flat(s.idv, m.inv_dv + e0, neo);
    copy_flat(s.ml, m.maxlev_e + e0, neo);
    if (vd2) copy_flat(s.cv2, mc.cv2 + e0, neo);
    if (vd4) copy_flat(s.cv4, mc.cv4 + e0, neo);
  }
  __syncthreads();
  if (k >= L) return;

  for (int i = row0; i < neo; i += RB) {
    const int e = e0 + i;
    const size_t idx = (size_t)e * L + k;
    Vd<N> T = tend_base<N>(a, a.vn, idx);
    const int4 own = s.cv[i];
    const double idc = s.idc[i];
    if (pv) {
      // edge_vort (vorticity_aux.hpp:108-131) of the edge itself
      Vd<N> qs;
      {
        const double f0 = s.f[own.z], f1 = s.f[own.w];
        const Vd<N> r0 = lrow<N, KS>(s.rv, own.z, lo), r1 = lrow<N, KS>(s.rv, own.w, lo);
        const Vd<N> i0 = lrow<N, KS>(s.inv, own.z, lo), i1 = lrow<N, KS>(s.inv, own.w, lo);
#pragma unroll
        for (int q = 0; q < N; ++q) {
          const double nrv = 0.5 * (i0[q] * r0[q] + i1[q] * r1[q]);
          const double npv = 0.5 * (i0[q] * f0 + i1[q] * f1);
          qs[q] = nrv + npv;
        }
      }
      Vd<N> qt = zero<N>();
      const int n = s.noe[i];
      for (int j = 0; j < n; ++j) {
        const int lj = s.eoe[i * me2 + j];
        const double wj = s.w[i * me2 + j];
        const int4 q4 = s.cv[lj];
        const Vd<N> h0 = lrow<N, KS>(s.h, q4.x, lo), h1 = lrow<N, KS>(s.h, q4.y, lo);
        const Vd<N> vj = lrow<N, KS>(s.vn, lj, lo);
        const double f0 = s.f[q4.z], f1 = s.f[q4.w];
        const Vd<N> r0 = lrow<N, KS>(s.rv, q4.z, lo), r1 = lrow<N, KS>(s.rv, q4.w, lo);
        const Vd<N> i0 = lrow<N, KS>(s.inv, q4.z, lo), i1 = lrow<N, KS>(s.inv, q4.w, lo);
#pragma unroll
        for (int q = 0; q < N; ++q) {
          const double nrv_j = 0.5 * (i0[q] * r0[q] + i1[q] * r1[q]);
          const double npv_j = 0.5 * (i0[q] * f0 + i1[q] * f1);
          const double fh = 0.5 * (h0[q] + h1[q]);
          const double nv = (qs[q] + nrv_j + npv_j) * 0.5;
          qt[q] = qt[q] + wj * fh * vj[q] * nv;
        }
      }
#pragma unroll
      for (int q = 0; q < N; ++q) T[q] = T[q] + qt[q];
    }
    if (ke) {
      const Vd<N> k0 = lrow<N, KS>(s.ke, own.x, lo), k1 = lrow<N, KS>(s.ke, own.y, lo);
#pragma unroll
      for (int q = 0; q < N; ++q) T[q] = T[q] - (k1[q] - k0[q]) * idc;
    }
    if (ssh) {
      const Vd<N> h0 = lrow<N, KS>(s.h, own.x, lo), h1 = lrow<N, KS>(s.h, own.y, lo);
#pragma unroll
      for (int q = 0; q < N; ++q) T[q] = T[q] - mc.grav * (h1[q] - h0[q]) * idc;
    }
    if (vd2 || vd4) {
      const double idv = s.idv[i];
      const int ml = s.ml[i];
      if (vd2) {
        const Vd<N> d0 = lrow<N, KS>(s.div, own.x, lo), d1 = lrow<N, KS>(s.div, own.y, lo);
        const Vd<N> r0 = lrow<N, KS>(s.rv, own.z, lo), r1 = lrow<N, KS>(s.rv, own.w, lo);
        const double cvv = s.cv2[i];
#pragma unroll
        for (int q = 0; q < N; ++q) {
          const double del2u = (d1[q] - d0[q]) * idc - (r1[q] - r0[q]) * idv;
          const double coef = (k + q < ml) ? cvv : 0.0;
          T[q] = T[q] + coef * del2u;
        }
      }
      if (vd4) {
        // the del4 pre-pass fields are read in place
        const int gc0 = m.coe[2 * e], gc1 = m.coe[2 * e + 1];
        const int gb0 = m.voe[2 * e], gb1 = m.voe[2 * e + 1];
        const Vd<N> d0 = ld<N>(d.vdel2_div_c + (size_t)gc0 * L + k);
        const Vd<N> d1 = ld<N>(d.vdel2_div_c + (size_t)gc1 * L + k);
        const Vd<N> r0 = ld<N>(d.vdel2_rv_v + (size_t)gb0 * L + k);
        const Vd<N> r1 = ld<N>(d.vdel2_rv_v + (size_t)gb1 * L + k);
        const double cvv = s.cv4[i];
#pragma unroll
        for (int q = 0; q < N; ++q) {
          const double del2u = (d1[q] - d0[q]) * idc - (r1[q] - r0[q]) * idv;
          const double coef = (k + q < ml) ? cvv : 0.0;
          T[q] = T[q] - coef * del2u;
        }
      }
    }
    epilogue<N>(a, a.vn, idx, T);
  }
}

// ---------------------------------------------------------------------------
// Cell kernel
// ---------------------------------------------------------------------------
struct CellLds {
  double *h, *vn, *tr, *q, *cdiv, *ct2, *ct4, *idc, *s2;
  int4 *eoc;
  int *noc;
};

template <int KS>
__device__ __forceinline__ CellLds cell_lds(double *base, const TileView &t, int me) {
  CellLds s;
  double *p = base;
  s.h = p;    p += t.max_cc * KS;
  s.vn = p;   p += t.max_ce * KS;
  s.tr = p;   p += kTrGroup * t.max_cc * KS;
  s.q = p;    p += kTrGroup * t.max_cc * KS;
  s.cdiv = p; p += t.tcells * me;
  s.ct2 = p;  p += t.tcells * me;
  s.ct4 = p;  p += t.tcells * me;
  s.idc = p;  p += t.max_ce;
  s.s2 = p;   p += t.max_ce;
  p += (p - base) & 1; // 16-byte alignment of the int4 table
  s.eoc = reinterpret_cast<int4 *>(p);
  s.noc = reinterpret_cast<int *>(s.eoc + t.tcells * me);
  return s;
}

size_t cell_lds_bytes(const TileView &t, int KS, int me) {
  const size_t d = (size_t)(t.max_cc + t.max_ce + 2 * kTrGroup * t.max_cc) * KS +
                   3 * (size_t)t.tcells * me + 2 * (size_t)t.max_ce;
  return (d + 1) * sizeof(double) + (size_t)t.tcells * me * sizeof(int4) +
         (size_t)t.tcells * sizeof(int);
}

template <int N, int K>
__global__ void __launch_bounds__(kThreads)
    k_tile_cells(MeshView m, TileView t, ModelCoef mc, StageArgs a, DiagOut d, int swz) {
  extern __shared__ double lds[];
  constexpr int LR = K / N;
  constexpr int RB = kThreads / LR;
  constexpr int KS = K + 2;
  const int tile = logical_block(blockIdx.x, gridDim.x, swz);
  const int L = m.L, NC = m.nc, me = m.me;
  const int lane = (int)threadIdx.x % LR, row0 = (int)threadIdx.x / LR;
  const int lo = lane * N, k = blockIdx.y * K + lo;
  const int c0 = tile * t.tcells, nco = min(c0 + t.tcells, NC) - c0;
  const int ce0 = t.ce_ofs[tile], nce = t.ce_ofs[tile + 1] - ce0;
  const int cc0 = t.cc_ofs[tile], ncc = t.cc_ofs[tile + 1] - cc0;
  const CellLds s = cell_lds<KS>(lds, t, me);
  const int ntr = a.ntr;
  const bool td2 = a.terms & T_TDEL2, td4 = a.terms & T_TDEL4;
  __shared__ SegTable tab;
  auto off = [&](const double *p) { return (int)(p - lds); };
  if (a.do_h || ntr > 0) {
    // tracers 0 and 1 are staged with the state rows
    const size_t tl = (size_t)NC * L;
    stage_rows<N, KS>(tab, lds, RowSeg{a.h.in, t.cc + cc0, off(s.h), ncc},
                      RowSeg{a.vn.in, t.ce + ce0, off(s.vn), nce},
                      ntr > 0 ? RowSeg{a.tr.in, t.cc + cc0, off(s.tr), ncc} : kNoRows,
                      ntr > 1 ? RowSeg{a.tr.in + tl, t.cc + cc0, off(s.tr) + t.max_cc * KS, ncc}
                              : kNoRows,
                      kNoRows, kNoRows, L, k, lo, row0, RB);
    copy_flat(s.eoc, t.eoc_l + (size_t)c0 * me, nco * me);
    copy_flat(s.cdiv, m.cdiv + (size_t)c0 * me, nco * me);
    copy_flat(s.noc, m.noc + c0, nco);
    if (ntr > 0) {
      for (int r = (int)threadIdx.x; r < nce; r += kThreads) s.idc[r] = m.inv_dc[t.ce[ce0 + r]];
      if (td2) {
        copy_flat(s.ct2, mc.ct2 + (size_t)c0 * me, nco * me);
        for (int r = (int)threadIdx.x; r < nce; r += kThreads) s.s2[r] = m.s2[t.ce[ce0 + r]];
      }
      if (td4) copy_flat(s.ct4, mc.ct4 + (size_t)c0 * me, nco * me);
    }
  }
  __syncthreads();
  const bool kin = k < L;
  // thickness (thickness_hadv_cell.hpp:48-72)
  if (kin) {
    for (int i = row0; i < nco; i += RB) {
      const int c = c0 + i;
      const size_t hidx = (size_t)c * L + k;
      if (a.do_h) {
        Vd<N> T = tend_base<N>(a, a.h, hidx);
        Vd<N> acc = zero<N>();
        const int n = s.noc[i];
        for (int j = 0; j < n; ++j) {
          const int4 q4 = s.eoc[i * me + j];
          const double cd = s.cdiv[i * me + j];
          const Vd<N> h0 = lrow<N, KS>(s.h, q4.y, lo), h1 = lrow<N, KS>(s.h, q4.z, lo);
          const Vd<N> ve = lrow<N, KS>(s.vn, q4.x, lo);
#pragma unroll
          for (int q = 0; q < N; ++q) acc[q] = acc[q] - cd * (0.5 * (h0[q] + h1[q])) * ve[q];
        }
#pragma unroll
        for (int q = 0; q < N; ++q) T[q] = T[q] + acc[q];
        epilogue<N>(a, a.h, hidx, T);
      } else if (a.epi != EPI_STORE) {
        epilogue<N>(a, a.h, hidx, tend_base<N>(a, a.h, hidx));
      }
    }
  }
  if (ntr <= 0) return;
  for (int l0 = 0; l0 < ntr; l0 += kTrGroup) {
    const int ng = min(kTrGroup, ntr - l0);
    __syncthreads(); // the previous group is done with the tracer rows
    {
      // tracer g of the group at s.tr[g], its Laplacian at s.q[g]; tracers 0
      // and 1 of the first group came with the state rows
      const size_t tl = (size_t)NC * L;
      const int gfirst = l0 == 0 ? min(ntr, 2) : 0;
      auto trs = [&](int g) {
        return (g >= gfirst && g < ng) ? RowSeg{a.tr.in + (size_t)(l0 + g) * tl, t.cc + cc0,
                                                off(s.tr) + g * t.max_cc * KS, ncc}
                                       : kNoRows;
      };
      auto qs = [&](int g) {
        return (td4 && g < ng) ? RowSeg{d.trdel2_c + (size_t)(l0 + g) * tl, t.cc + cc0,
                                        off(s.q) + g * t.max_cc * KS, ncc}
                               : kNoRows;
      };
      stage_rows<N, KS>(tab, lds, trs(2), trs(3), qs(0), qs(1), qs(2), qs(3), L, k, lo, row0, RB);
      if (gfirst == 0)
        stage_rows<N, KS>(tab, lds, trs(0), trs(1), kNoRows, kNoRows, kNoRows, kNoRows, L, k, lo,
                          row0, RB);
    }
    __syncthreads();
    if (!kin) continue;
    // tracer_hadv_cell.hpp:60-90, tracer_diff_cell.hpp:61-92,
    // tracer_hyperdiff_cell.hpp:131-161; one task per (cell, tracer)
    for (int task = row0; task < nco * ng; task += RB) {
      const int i = task / ng, g = task % ng, l = l0 + g;
      const int c = c0 + i;
      const double *str = s.tr + g * t.max_cc * KS;
      const double *sq = s.q + g * t.max_cc * KS;
      const Vd<N> hs = lrow<N, KS>(s.h, i, lo);
      Vd<N> inv_s;
#pragma unroll
      for (int q = 0; q < N; ++q) inv_s[q] = 1.0 / hs[q];
      Vd<N> adv = zero<N>(), dif = zero<N>(), hyp = zero<N>();
      const int n = s.noc[i];
      for (int j = 0; j < n; ++j) {
        const int4 q4 = s.eoc[i * me + j];
        const double cd = s.cdiv[i * me + j];
        const double idc = s.idc[q4.x];
        const Vd<N> h0 = lrow<N, KS>(s.h, q4.y, lo), h1 = lrow<N, KS>(s.h, q4.z, lo);
        const Vd<N> ve = lrow<N, KS>(s.vn, q4.x, lo);
        const Vd<N> t0 = lrow<N, KS>(str, q4.y, lo), t1 = lrow<N, KS>(str, q4.z, lo);
        const double c2 = td2 ? s.ct2[i * me + j] : 0.0;
        const double sc2 = td2 ? s.s2[q4.x] : 0.0;
        const double c4 = td4 ? s.ct4[i * me + j] : 0.0;
        Vd<N> q0 = zero<N>(), q1 = zero<N>();
        if (td4) {
          q0 = lrow<N, KS>(sq, q4.y, lo);
          q1 = lrow<N, KS>(sq, q4.z, lo);
        }
#pragma unroll
        for (int q = 0; q < N; ++q) {
          const double inv0 = (q4.y == i) ? inv_s[q] : 1.0 / h0[q];
          const double inv1 = (q4.z == i) ? inv_s[q] : 1.0 / h1[q];
          const double fh = 0.5 * (h0[q] + h1[q]);
          const double n0 = t0[q] * inv0, n1 = t1[q] * inv1;
          const double ntr_e = (n0 + n1) * 0.5;
          adv[q] = adv[q] - cd * fh * ntr_e * ve[q];
          if (td2) {
            const double grad = (n1 - n0) * idc;
            dif[q] = dif[q] + c2 * fh * sc2 * grad;
          }
          if (td4) {
            const double grad = (q1[q] - q0[q]) * idc;
            hyp[q] = hyp[q] - c4 * grad;
          }
        }
      }
      const size_t tidx = ((size_t)l * NC + c) * L + k;
      Vd<N> T = tend_base<N>(a, a.tr, tidx);
#pragma unroll
      for (int q = 0; q < N; ++q) {
        T[q] = T[q] + adv[q];
        if (td2) T[q] = T[q] + dif[q];
        if (td4) T[q] = T[q] + hyp[q];
      }
      epilogue<N>(a, a.tr, tidx, T);
    }
  }
}

// Largest dynamic LDS allocation of one block on the current device.
size_t lds_limit() {
  int dev = 0, bytes = 0;
  OMEGA_HIP_CHECK(hipGetDevice(&dev));
  OMEGA_HIP_CHECK(hipDeviceGetAttribute(&bytes, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
  return (size_t)bytes;
}

template <int N, int K>
void launch_nk(const MeshView &m, const TileView &t, const Launch &lc, const ModelCoef &mc,
               const StageArgs &a, const DiagOut &d, hipStream_t s) {
  const dim3 grid(t.ntiles, (m.L + K - 1) / K);
  const size_t lds_e = edge_lds_bytes(t, K + 2, m.me2);
  const size_t lds_c = cell_lds_bytes(t, K + 2, m.me);
  const size_t limit = lds_limit();
  OMEGA_REQUIRE(lds_e <= limit && lds_c <= limit,
                "tiled kernels: tile too large for LDS (use fewer cells per tile or fewer levels)");
  hipLaunchKernelGGL((k_tile_edges<N, K>), grid, dim3(kThreads), lds_e, s, m, t, mc, a, d,
                     lc.swizzle);
  OMEGA_HIP_CHECK(hipGetLastError());
  hipLaunchKernelGGL((k_tile_cells<N, K>), grid, dim3(kThreads), lds_c, s, m, t, mc, a, d,
                     lc.swizzle);
  OMEGA_HIP_CHECK(hipGetLastError());
}

} // namespace

void launch_tend_tile(const MeshView &m, const TileView &t, const Launch &lc,
                      const ModelCoef &mc, const StageArgs &a, const DiagOut &d, hipStream_t s) {
  TimerScope ts("tend", s);
  const int N = (lc.lpt == 2 && m.L % 2 == 0) ? 2 : 1;
  const int K = lc.tile_k == 8 ? 8 : 16;
  if (N == 2 && K == 16) launch_nk<2, 16>(m, t, lc, mc, a, d, s);
  else if (N == 2) launch_nk<2, 8>(m, t, lc, mc, a, d, s);
  else if (K == 16) launch_nk<1, 16>(m, t, lc, mc, a, d, s);
  else launch_nk<1, 8>(m, t, lc, mc, a, d, s);
}

} // namespace omega_hip
