// Ensemble kernels: the forward operators of fwd_ops.hpp and hex_ops.hpp
// applied to every member of an ensemble in one launch per stage phase.
// Table entry b belongs to member blk[b] (wavefront-uniform, so the
// member's descriptor comes through scalar loads); inside the member the
// entry covers `rows` cells of its space-filling-curve order and `lx`
// lanes x N levels of one level chunk, exactly as the single-model kernels
// do, so each member's results are the same bits as stepping it alone.
// Members on the regular hexagonal mesh run the pattern-specialised
// operators, the others the general ones, in the same launch.  With a
// persistent grid each block walks a contiguous run of table entries.
#include <algorithm>

#include "batch.hpp"
#include "fwd_ops.hpp"
#include "hex_ops.hpp"

namespace omega_hip {

namespace {

using namespace dev;
using namespace fwd;

struct Place {
  int c0, c1, k, ty, R;
  bool active;
};

__device__ __forceinline__ Place place(const BatchMember &B, int lb, int N) {
  Place p;
  const int chunk = lb % B.chunks, seg = lb / B.chunks;
  const int tx = (int)threadIdx.x % B.lx;
  p.ty = (int)threadIdx.x / B.lx;
  p.R = B.rows;
  p.c0 = seg * p.R;
  p.c1 = min(p.c0 + p.R, B.m.nc);
  p.k = (chunk * B.lx + tx) * N;
  p.active = p.ty < p.R && p.c0 < B.m.nc && p.k < B.m.L;
  return p;
}

// Table entries a block handles: one entry per block, or (persistent grid,
// PERS) a contiguous run of entries per block.  Either way the entries are
// dealt so that each XCD (workgroups go to the 8 XCDs round-robin) walks a
// contiguous part of the table, which keeps a member's neighbouring cells,
// and their halo re-reads, in one XCD's L2.  PERS is a template parameter:
// one code path per kernel keeps the register budget of the entry body.
template <bool PERS, class F>
__device__ __forceinline__ void for_entries(int nblocks, F &&f) {
  if constexpr (!PERS) {
    f(logical_block(blockIdx.x, gridDim.x, 1));
  } else {
    const int G = gridDim.x;
    const int lb = logical_block(blockIdx.x, G, 1);
    const int b0 = (int)((long long)lb * nblocks / G);
    const int b1 = (int)((long long)(lb + 1) * nblocks / G);
    for (int b = b0; b < b1; ++b) f(b);
  }
}

// general-kernel phases of one table entry (member B, block lb of it)
template <int N>
__device__ __forceinline__ void gen_diag(const BatchMember &B, int lb) {
  const Place p = place(B, lb, N);
  if (!p.active) return;
  const MeshView &m = B.m;
  const DiagOut &d = B.d;
  const int terms = B.dterms, ntr = B.a.ntr;
  const double *h = B.a.h.in, *vn = B.a.vn.in, *tr = B.a.tr.in;
  if (terms & (T_PV | T_VDEL2 | T_VDEL4)) {
    const int vb = m.vofs[p.c0], ve = m.vofs[p.c1];
    for (int v = vb + p.ty; v < ve; v += p.R) vertex_diag<N>(m, v, p.k, h, vn, d);
  }
  for (int c = p.c0 + p.ty; c < p.c1; c += p.R) cell_diag<N>(m, terms, ntr, c, p.k, h, vn, tr, d);
}

template <int N>
__device__ __forceinline__ void gen_tend(const BatchMember &B, int lb) {
  const Place p = place(B, lb, N);
  if (!p.active) return;
  // the descriptor is read in place (global memory, wavefront-uniform
  // address, scalar loads); copying it out would put it on the stack
  const MeshView &m = B.m;
  const ModelCoef &mc = B.mc;
  const StageArgs &a = B.a;
  const DiagOut &d = B.d;
  const int eb = m.eofs[p.c0], ee = m.eofs[p.c1];
  for (int e = eb + p.ty; e < ee; e += p.R) edge_tend<N>(m, mc, a, d, e, p.k);
  for (int c = p.c0 + p.ty; c < p.c1; c += p.R) cell_tend<N>(m, mc, a, d, c, p.k);
}

template <int N, bool PERS>
__global__ void __launch_bounds__(256)
    k_batch_vdel4_edge(const BatchMember *__restrict__ desc, const int *__restrict__ blk,
                       int nblocks) {
  for_entries<PERS>(nblocks, [&](int b) {
    const int id = uniform(blk[b]);
    const BatchMember &B = desc[id];
    if (!B.vdel4) return;
    const Place p = place(B, b - B.block0, N);
    if (!p.active) return;
    const MeshView &m = B.m;
    const DiagOut &d = B.d;
    const int eb = m.eofs[p.c0], ee = m.eofs[p.c1];
    for (int e = eb + p.ty; e < ee; e += p.R) vdel4_edge<N>(m, d, e, p.k);
  });
}

template <int N, bool PERS>
__global__ void __launch_bounds__(256)
    k_batch_vdel4_vc(const BatchMember *__restrict__ desc, const int *__restrict__ blk,
                     int nblocks) {
  for_entries<PERS>(nblocks, [&](int b) {
    const int id = uniform(blk[b]);
    const BatchMember &B = desc[id];
    if (!B.vdel4) return;
    const Place p = place(B, b - B.block0, N);
    if (!p.active) return;
    const MeshView &m = B.m;
    const DiagOut &d = B.d;
    const int vb = m.vofs[p.c0], ve = m.vofs[p.c1];
    for (int v = vb + p.ty; v < ve; v += p.R) vdel4_vertex<N>(m, d, v, p.k);
    for (int c = p.c0 + p.ty; c < p.c1; c += p.R) vdel4_cell<N>(m, d, c, p.k);
  });
}

// Pattern-specialised phases, N = 1: one cell (its 3 owned edges and 2
// owned vertices) per row of 64 lanes.
__device__ __forceinline__ void hex1_diag(const BatchMember &B, int lb) {
  const int chunk = lb % B.chunks, seg = lb / B.chunks;
  const int c = hexops::ent<1>(seg * B.rows + (int)threadIdx.x / B.lx);
  if (c >= B.m.nc) return;
  const int k = chunk * B.lx + (int)threadIdx.x % B.lx;
  const MeshView &m = B.m;
  const DiagOut &d = B.d;
  const double *h = B.a.h.in, *vn = B.a.vn.in, *tr = B.a.tr.in;
  if (B.dterms & (T_PV | T_VDEL2 | T_VDEL4)) {
    hexops::vertex_diag_u<1>(m, 2 * c, k, h, vn, d);
    hexops::vertex_diag_u<1>(m, 2 * c + 1, k, h, vn, d);
  }
  hexops::cell_diag_u<1>(m, B.dterms, B.a.ntr, c, k, h, vn, tr, d);
}

__device__ __forceinline__ void hex1_tend(const BatchMember &B, int lb) {
  const int chunk = lb % B.chunks, seg = lb / B.chunks;
  const int c = hexops::ent<1>(seg * B.rows + (int)threadIdx.x / B.lx);
  if (c >= B.m.nc) return;
  const int k = chunk * B.lx + (int)threadIdx.x % B.lx;
  const MeshView &m = B.m;
  const ModelCoef &mc = B.mc;
  const StageArgs &a = B.a;
  const DiagOut &d = B.d;
  const int e0 = 3 * c;
  hexops::edge_tend_hex<1, 0>(m, mc, a, d, e0, k);
  hexops::edge_tend_hex<1, 1>(m, mc, a, d, e0 + 1, k);
  hexops::edge_tend_hex<1, 2>(m, mc, a, d, e0 + 2, k);
  hexops::cell_tend_hex<1>(m, mc, a, d, c, k);
}
// N = 2 forms: the two rows of 32 lanes of a wavefront carry consecutive
// cells cA, cA + 1 of the member, whose tables come through the scalar
// cache (hex_ops.hpp pair operators), as in the single-model kernels.
struct Pair {
  int cA, cB, k;
  bool hi, ok;
};

__device__ __forceinline__ Pair pair_of(const BatchMember &B, int lb) {
  Pair q;
  const int chunk = lb % B.chunks, seg = lb / B.chunks;
  const int row = (int)threadIdx.x / B.lx;
  q.cA = uniform(seg * B.rows + (row & ~1));
  q.cB = min(q.cA + 1, B.m.nc - 1);
  q.hi = (row & 1) != 0;
  q.ok = q.cA < B.m.nc && seg * B.rows + row < B.m.nc;
  q.k = (chunk * B.lx + (int)threadIdx.x % B.lx) * 2;
  return q;
}

// the forms of the pair kernels a member's flags allow (same bits)
template <bool U, bool W>
__device__ __forceinline__ void batch_diag_pair(const BatchMember &B, const Pair &q) {
  const MeshView &m = B.m;
  const DiagOut &d = B.d;
  const double *h = B.a.h.in, *vn = B.a.vn.in, *tr = B.a.tr.in;
  if (m.hexd_ok) {
    hexops::diag_union_p<U, W>(m, B.dterms, B.a.ntr, q.cA, q.cB, q.hi, q.k, h, vn, tr, d);
    return;
  }
  if (B.dterms & (T_PV | T_VDEL2 | T_VDEL4)) {
    hexops::vertex_diag_p<U, W>(m, 2 * q.cA, 2 * q.cB, q.hi, q.k, h, vn, d);
    hexops::vertex_diag_p<U, W>(m, 2 * q.cA + 1, 2 * q.cB + 1, q.hi, q.k, h, vn, d);
  }
  hexops::cell_diag_p<U, W>(m, B.dterms, B.a.ntr, q.cA, q.cB, q.hi, q.k, h, vn, tr, d);
}

// members with 16 or 32 layers in the N = 2 launches: one cell per row of
// L/2 lanes (hex_ops.hpp single-cell operators)
struct Row {
  int c, k;
};
__device__ __forceinline__ Row row_of(const BatchMember &B, int lb) {
  const int chunk = lb % B.chunks, seg = lb / B.chunks;
  return Row{seg * B.rows + (int)threadIdx.x / B.lx, (chunk * B.lx + (int)threadIdx.x % B.lx) * 2};
}

// TM >= 0: the member's term set, compiled in (both tendencies on, the
// vertex-union records present)
template <bool U, bool W, int TM = -1>
__device__ __forceinline__ void batch_tend_pair(const BatchMember &B, const Pair &q) {
  const MeshView &m = B.m;
  const ModelCoef &mc = B.mc;
  const StageArgs &a = B.a;
  const DiagOut &d = B.d;
  const int eA = 3 * q.cA, eB = 3 * q.cB;
  if (TM >= 0 || m.hexu_ok) {
    Vd<2> P[hexu::kNV], Q[hexu::kNV];
    if (a.do_vn && (a.terms & T_PV)) hexops::union_pv2<U, W>(m, mc, d, q.cA, q.cB, q.hi, q.k, P, Q);
    hexops::edge_tend_hex2v<0, U, W, TM>(m, mc, a, d, eA, eB, q.hi, q.k, P, Q);
    hexops::edge_tend_hex2v<1, U, W, TM>(m, mc, a, d, eA + 1, eB + 1, q.hi, q.k, P, Q);
    hexops::edge_tend_hex2v<2, U, W, TM>(m, mc, a, d, eA + 2, eB + 2, q.hi, q.k, P, Q);
  } else {
    hexops::edge_tend_hex2<0, U, W>(m, mc, a, d, eA, eB, q.hi, q.k);
    hexops::edge_tend_hex2<1, U, W>(m, mc, a, d, eA + 1, eB + 1, q.hi, q.k);
    hexops::edge_tend_hex2<2, U, W>(m, mc, a, d, eA + 2, eB + 2, q.hi, q.k);
  }
  hexops::cell_tend_hex2<U, W, TM>(m, mc, a, d, q.cA, q.cB, q.hi, q.k);
}

__device__ __forceinline__ void hex2_diag(const BatchMember &B, int lb) {
  if (B.m.L % 64 != 0) {
    const Row r = row_of(B, lb);
    if (r.c >= B.m.nc) return;
    const double *h = B.a.h.in, *vn = B.a.vn.in, *tr = B.a.tr.in;
    if (B.dterms & (T_PV | T_VDEL2 | T_VDEL4)) {
      hexops::vertex_diag_u<2>(B.m, 2 * r.c, r.k, h, vn, B.d);
      hexops::vertex_diag_u<2>(B.m, 2 * r.c + 1, r.k, h, vn, B.d);
    }
    hexops::cell_diag_u<2>(B.m, B.dterms, B.a.ntr, r.c, r.k, h, vn, tr, B.d);
    return;
  }
  const Pair q = pair_of(B, lb);
  if (!q.ok) return;
  if (B.m.coef_uni && B.m.off32) batch_diag_pair<true, false>(B, q);
  else batch_diag_pair<false, true>(B, q);
}

template <int TM = -1> __device__ __forceinline__ void hex2_tend(const BatchMember &B, int lb) {
  if (B.m.L % 64 != 0) {
    const Row r = row_of(B, lb);
    if (r.c >= B.m.nc) return;
    hexops::edge_tend_hex<2, 0, TM>(B.m, B.mc, B.a, B.d, 3 * r.c, r.k);
    hexops::edge_tend_hex<2, 1, TM>(B.m, B.mc, B.a, B.d, 3 * r.c + 1, r.k);
    hexops::edge_tend_hex<2, 2, TM>(B.m, B.mc, B.a, B.d, 3 * r.c + 2, r.k);
    hexops::cell_tend_hex<2, TM>(B.m, B.mc, B.a, B.d, r.c, r.k);
    return;
  }
  const Pair q = pair_of(B, lb);
  if (!q.ok) return;
  if constexpr (TM >= 0) batch_tend_pair<true, false, TM>(B, q);
  else if (B.m.coef_uni && B.mc.uniform && B.m.off32) batch_tend_pair<true, false>(B, q);
  else batch_tend_pair<false, true>(B, q);
}

// One launch per stage phase over every member: each table entry runs the
// pattern-specialised or the general operators its member was built for.
// KINDS (bit 0 general members, bit 1 pattern-specialised members) compiles
// only the paths the ensemble has, so an ensemble of one kind keeps the
// register budget (and occupancy) of that kind's operators alone.
template <int N, int KINDS, bool PERS>
__global__ void __launch_bounds__(256)
    k_batch_diag(const BatchMember *__restrict__ desc, const int *__restrict__ blk, int nblocks) {
  for_entries<PERS>(nblocks, [&](int b) {
    const int id = uniform(blk[b]);
    const BatchMember &B = desc[id];
    if (!B.dterms) return;
    const int lb = b - B.block0;
    if (KINDS == 2 || (KINDS == 3 && B.hex)) {
      if constexpr ((KINDS & 2) && N == 2) hex2_diag(B, lb);
      else if constexpr (KINDS & 2) hex1_diag(B, lb);
    } else if constexpr (KINDS & 1) {
      gen_diag<N>(B, lb);
    }
  });
}

// TM >= 0 (pattern-specialised members only, N = 2): every member has the
// term set TM with both tendencies on, and every 64k-layer member the
// shared-coefficient, 32-bit-offset forms, so their operators get it
// compiled in.
template <int N, int KINDS, bool PERS, int TM = -1>
__global__ void __launch_bounds__(256)
    k_batch_tend(const BatchMember *__restrict__ desc, const int *__restrict__ blk, int nblocks) {
  for_entries<PERS>(nblocks, [&](int b) {
    const int id = uniform(blk[b]);
    const BatchMember &B = desc[id];
    const int lb = b - B.block0;
    if (KINDS == 2 || (KINDS == 3 && B.hex)) {
      if constexpr ((KINDS & 2) && N == 2) hex2_tend<TM>(B, lb);
      else if constexpr (KINDS & 2) hex1_tend(B, lb);
    } else if constexpr (KINDS & 1) {
      gen_tend<N>(B, lb);
    }
  });
}

} // namespace

namespace {
// grid of a launch: one block per table entry, or the persistent grid
void grid_of(int nblocks, int persistent, int &grid, bool &pers) {
  grid = persistent ? std::min(persistent, nblocks) : nblocks;
  pers = persistent && persistent < nblocks;
}

template <int N, int KINDS, bool PERS>
void launch_phase(int phase, const BatchMember *desc, const int *blk, int nblocks, int grid,
                  int threads, hipStream_t s) {
  if (phase == 0)
    hipLaunchKernelGGL((k_batch_diag<N, KINDS, PERS>), dim3(grid), dim3(threads), 0, s, desc,
                       blk, nblocks);
  else
    hipLaunchKernelGGL((k_batch_tend<N, KINDS, PERS>), dim3(grid), dim3(threads), 0, s, desc,
                       blk, nblocks);
}

// the tendency launch of a pattern-specialised ensemble with a compiled-in
// term set (term_set_batch)
template <bool PERS>
bool launch_tend_set(int tset, const BatchMember *desc, const int *blk, int nblocks, int grid,
                     int threads, hipStream_t s) {
  void (*k)(const BatchMember *, const int *, int) = nullptr;
  switch (tset) {
  case 7: k = k_batch_tend<2, 2, PERS, 7>; break;
  case 15: k = k_batch_tend<2, 2, PERS, 15>; break;
  case 79: k = k_batch_tend<2, 2, PERS, 79>; break;
  case 127: k = k_batch_tend<2, 2, PERS, 127>; break;
  default: return false;
  }
  hipLaunchKernelGGL(k, dim3(grid), dim3(threads), 0, s, desc, blk, nblocks);
  return true;
}

template <int KINDS, bool PERS>
void launch_n(int phase, const BatchMember *desc, const int *blk, int nblocks, int grid,
              int threads, int N, hipStream_t s) {
  if (N == 2) launch_phase<2, KINDS, PERS>(phase, desc, blk, nblocks, grid, threads, s);
  else launch_phase<1, KINDS, PERS>(phase, desc, blk, nblocks, grid, threads, s);
}

template <bool PERS>
void launch_k(int phase, const BatchMember *desc, const int *blk, int nblocks, int grid,
              int threads, int N, int kinds, hipStream_t s) {
  if (kinds == 1) launch_n<1, PERS>(phase, desc, blk, nblocks, grid, threads, N, s);
  else if (kinds == 2) launch_n<2, PERS>(phase, desc, blk, nblocks, grid, threads, N, s);
  else launch_n<3, PERS>(phase, desc, blk, nblocks, grid, threads, N, s);
}

void launch_kinds(int phase, const BatchMember *desc, const int *blk, int nblocks, int threads,
                  int N, int persistent, int kinds, hipStream_t s) {
  int grid;
  bool pers;
  grid_of(nblocks, persistent, grid, pers);
  if (pers) launch_k<true>(phase, desc, blk, nblocks, grid, threads, N, kinds, s);
  else launch_k<false>(phase, desc, blk, nblocks, grid, threads, N, kinds, s);
  OMEGA_HIP_CHECK(hipGetLastError());
}

template <int N, bool PERS>
void launch_vdel4_n(const BatchMember *desc, const int *blk, int nblocks, int grid, int threads,
                    hipStream_t s) {
  hipLaunchKernelGGL((k_batch_vdel4_edge<N, PERS>), dim3(grid), dim3(threads), 0, s, desc, blk,
                     nblocks);
  hipLaunchKernelGGL((k_batch_vdel4_vc<N, PERS>), dim3(grid), dim3(threads), 0, s, desc, blk,
                     nblocks);
}
} // namespace

void launch_batch_diag(const BatchMember *desc, const int *blk, int nblocks, int threads, int N,
                       int persistent, int kinds, hipStream_t s) {
  TimerScope ts("batch_diag", s);
  launch_kinds(0, desc, blk, nblocks, threads, N, persistent, kinds, s);
}

void launch_batch_vdel4(const BatchMember *desc, const int *blk, int nblocks, int threads, int N,
                        int persistent, hipStream_t s) {
  TimerScope ts("batch_vdel4", s);
  int grid;
  bool pers;
  grid_of(nblocks, persistent, grid, pers);
  if (N == 2) {
    if (pers) launch_vdel4_n<2, true>(desc, blk, nblocks, grid, threads, s);
    else launch_vdel4_n<2, false>(desc, blk, nblocks, grid, threads, s);
  } else {
    if (pers) launch_vdel4_n<1, true>(desc, blk, nblocks, grid, threads, s);
    else launch_vdel4_n<1, false>(desc, blk, nblocks, grid, threads, s);
  }
  OMEGA_HIP_CHECK(hipGetLastError());
}

void launch_batch_tend(const BatchMember *desc, const int *blk, int nblocks, int threads, int N,
                       int persistent, int kinds, int tset, hipStream_t s) {
  TimerScope ts("batch_tend", s);
  if (tset > 0 && N == 2 && kinds == 2) {
    int grid;
    bool pers;
    grid_of(nblocks, persistent, grid, pers);
    const bool done = pers ? launch_tend_set<true>(tset, desc, blk, nblocks, grid, threads, s)
                           : launch_tend_set<false>(tset, desc, blk, nblocks, grid, threads, s);
    if (done) {
      OMEGA_HIP_CHECK(hipGetLastError());
      return;
    }
  }
  launch_kinds(1, desc, blk, nblocks, threads, N, persistent, kinds, s);
}

bool term_set_batch(int terms) { return terms == 7 || terms == 15 || terms == 79 || terms == 127; }

} // namespace omega_hip
