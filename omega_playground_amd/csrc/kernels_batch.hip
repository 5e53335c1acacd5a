// Ensemble kernels: the forward operators of fwd_ops.hpp applied to every
// member of an ensemble in one launch.  Block b belongs to member blk[b]
// (wavefront-uniform, so the member's descriptor comes through scalar
// loads); inside the member the block covers `rows` cells of its
// space-filling-curve order and `lx` lanes x N levels of one level chunk,
// exactly as the single-model kernels do, so each member's results are the
// same bits as stepping it alone.
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

template <int N>
__global__ void __launch_bounds__(256)
    k_batch_diag(const BatchMember *__restrict__ desc, const int *__restrict__ blk) {
  const int id = uniform(blk[blockIdx.x]);
  const BatchMember &B = desc[id];
  if (!B.dterms) return;
  const Place p = place(B, (int)blockIdx.x - B.block0, N);
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
__global__ void __launch_bounds__(256)
    k_batch_vdel4_edge(const BatchMember *__restrict__ desc, const int *__restrict__ blk) {
  const int id = uniform(blk[blockIdx.x]);
  const BatchMember &B = desc[id];
  if (!B.vdel4) return;
  const Place p = place(B, (int)blockIdx.x - B.block0, N);
  if (!p.active) return;
  const MeshView &m = B.m;
  const DiagOut &d = B.d;
  const int eb = m.eofs[p.c0], ee = m.eofs[p.c1];
  for (int e = eb + p.ty; e < ee; e += p.R) vdel4_edge<N>(m, d, e, p.k);
}

template <int N>
__global__ void __launch_bounds__(256)
    k_batch_vdel4_vc(const BatchMember *__restrict__ desc, const int *__restrict__ blk) {
  const int id = uniform(blk[blockIdx.x]);
  const BatchMember &B = desc[id];
  if (!B.vdel4) return;
  const Place p = place(B, (int)blockIdx.x - B.block0, N);
  if (!p.active) return;
  const MeshView &m = B.m;
  const DiagOut &d = B.d;
  const int vb = m.vofs[p.c0], ve = m.vofs[p.c1];
  for (int v = vb + p.ty; v < ve; v += p.R) vdel4_vertex<N>(m, d, v, p.k);
  for (int c = p.c0 + p.ty; c < p.c1; c += p.R) vdel4_cell<N>(m, d, c, p.k);
}

template <int N>
__global__ void __launch_bounds__(256)
    k_batch_tend(const BatchMember *__restrict__ desc, const int *__restrict__ blk) {
  const int id = uniform(blk[blockIdx.x]);
  const BatchMember &B = desc[id];
  const Place p = place(B, (int)blockIdx.x - B.block0, N);
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

// Pattern-specialised phases: one cell (its 3 owned edges and 2 owned
// vertices) per row of 64/N lanes.
template <int N>
__global__ void __launch_bounds__(256)
    k_batch_diag_hex(const BatchMember *__restrict__ desc, const int *__restrict__ blk) {
  const int id = uniform(blk[blockIdx.x]);
  const BatchMember &B = desc[id];
  if (!B.dterms) return;
  const int lb = (int)blockIdx.x - B.block0;
  const int chunk = lb % B.chunks, seg = lb / B.chunks;
  const int c = hexops::ent<N>(seg * B.rows + (int)threadIdx.x / B.lx);
  if (c >= B.m.nc) return;
  const int k = (chunk * B.lx + (int)threadIdx.x % B.lx) * N;
  const MeshView &m = B.m;
  const DiagOut &d = B.d;
  const double *h = B.a.h.in, *vn = B.a.vn.in, *tr = B.a.tr.in;
  if (B.dterms & (T_PV | T_VDEL2 | T_VDEL4)) {
    hexops::vertex_diag_u<N>(m, 2 * c, k, h, vn, d);
    hexops::vertex_diag_u<N>(m, 2 * c + 1, k, h, vn, d);
  }
  hexops::cell_diag_u<N>(m, B.dterms, B.a.ntr, c, k, h, vn, tr, d);
}

template <int N>
__global__ void __launch_bounds__(256)
    k_batch_tend_hex(const BatchMember *__restrict__ desc, const int *__restrict__ blk) {
  const int id = uniform(blk[blockIdx.x]);
  const BatchMember &B = desc[id];
  const int lb = (int)blockIdx.x - B.block0;
  const int chunk = lb % B.chunks, seg = lb / B.chunks;
  const int c = hexops::ent<N>(seg * B.rows + (int)threadIdx.x / B.lx);
  if (c >= B.m.nc) return;
  const int k = (chunk * B.lx + (int)threadIdx.x % B.lx) * N;
  const MeshView &m = B.m;
  const ModelCoef &mc = B.mc;
  const StageArgs &a = B.a;
  const DiagOut &d = B.d;
  const int e0 = 3 * c;
  hexops::edge_tend_hex<N, 0>(m, mc, a, d, e0, k);
  hexops::edge_tend_hex<N, 1>(m, mc, a, d, e0 + 1, k);
  hexops::edge_tend_hex<N, 2>(m, mc, a, d, e0 + 2, k);
  hexops::cell_tend_hex<N>(m, mc, a, d, c, k);
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

__global__ void __launch_bounds__(256)
    k_batch_diag_hex2(const BatchMember *__restrict__ desc, const int *__restrict__ blk) {
  const int id = uniform(blk[blockIdx.x]);
  const BatchMember &B = desc[id];
  if (!B.dterms) return;
  if (B.m.L % 64 != 0) {
    const Row r = row_of(B, (int)blockIdx.x - B.block0);
    if (r.c >= B.m.nc) return;
    const double *h = B.a.h.in, *vn = B.a.vn.in, *tr = B.a.tr.in;
    if (B.dterms & (T_PV | T_VDEL2 | T_VDEL4)) {
      hexops::vertex_diag_u<2>(B.m, 2 * r.c, r.k, h, vn, B.d);
      hexops::vertex_diag_u<2>(B.m, 2 * r.c + 1, r.k, h, vn, B.d);
    }
    hexops::cell_diag_u<2>(B.m, B.dterms, B.a.ntr, r.c, r.k, h, vn, tr, B.d);
    return;
  }
  const Pair q = pair_of(B, (int)blockIdx.x - B.block0);
  if (!q.ok) return;
  if (B.m.coef_uni && B.m.off32) batch_diag_pair<true, false>(B, q);
  else batch_diag_pair<false, true>(B, q);
}

template <bool U, bool W>
__device__ __forceinline__ void batch_tend_pair(const BatchMember &B, const Pair &q) {
  const MeshView &m = B.m;
  const ModelCoef &mc = B.mc;
  const StageArgs &a = B.a;
  const DiagOut &d = B.d;
  const int eA = 3 * q.cA, eB = 3 * q.cB;
  if (m.hexu_ok) {
    Vd<2> P[hexu::kNV], Q[hexu::kNV];
    if (a.do_vn && (a.terms & T_PV)) hexops::union_pv2<U, W>(m, mc, d, q.cA, q.cB, q.hi, q.k, P, Q);
    hexops::edge_tend_hex2v<0, U, W>(m, mc, a, d, eA, eB, q.hi, q.k, P, Q);
    hexops::edge_tend_hex2v<1, U, W>(m, mc, a, d, eA + 1, eB + 1, q.hi, q.k, P, Q);
    hexops::edge_tend_hex2v<2, U, W>(m, mc, a, d, eA + 2, eB + 2, q.hi, q.k, P, Q);
  } else {
    hexops::edge_tend_hex2<0, U, W>(m, mc, a, d, eA, eB, q.hi, q.k);
    hexops::edge_tend_hex2<1, U, W>(m, mc, a, d, eA + 1, eB + 1, q.hi, q.k);
    hexops::edge_tend_hex2<2, U, W>(m, mc, a, d, eA + 2, eB + 2, q.hi, q.k);
  }
  hexops::cell_tend_hex2<U, W>(m, mc, a, d, q.cA, q.cB, q.hi, q.k);
}

__global__ void __launch_bounds__(256)
    k_batch_tend_hex2(const BatchMember *__restrict__ desc, const int *__restrict__ blk) {
  const int id = uniform(blk[blockIdx.x]);
  const BatchMember &B = desc[id];
  if (B.m.L % 64 != 0) {
    const Row r = row_of(B, (int)blockIdx.x - B.block0);
    if (r.c >= B.m.nc) return;
    hexops::edge_tend_hex<2, 0>(B.m, B.mc, B.a, B.d, 3 * r.c, r.k);
    hexops::edge_tend_hex<2, 1>(B.m, B.mc, B.a, B.d, 3 * r.c + 1, r.k);
    hexops::edge_tend_hex<2, 2>(B.m, B.mc, B.a, B.d, 3 * r.c + 2, r.k);
    hexops::cell_tend_hex<2>(B.m, B.mc, B.a, B.d, r.c, r.k);
    return;
  }
  const Pair q = pair_of(B, (int)blockIdx.x - B.block0);
  if (!q.ok) return;
  if (B.m.coef_uni && B.mc.uniform && B.m.off32) batch_tend_pair<true, false>(B, q);
  else batch_tend_pair<false, true>(B, q);
}

} // namespace

#define OMEGA_BATCH_LAUNCH(K)                                                               \
  do {                                                                                      \
    if (N == 2) hipLaunchKernelGGL(K<2>, dim3(nblocks), dim3(threads), 0, s, desc, blk);    \
    else hipLaunchKernelGGL(K<1>, dim3(nblocks), dim3(threads), 0, s, desc, blk);           \
    OMEGA_HIP_CHECK(hipGetLastError());                                                     \
  } while (0)

void launch_batch_diag(const BatchMember *desc, const int *blk, int nblocks, int threads, int N,
                       hipStream_t s) {
  TimerScope ts("batch_diag", s);
  OMEGA_BATCH_LAUNCH(k_batch_diag);
}

void launch_batch_vdel4(const BatchMember *desc, const int *blk, int nblocks, int threads, int N,
                        hipStream_t s) {
  TimerScope ts("batch_vdel4", s);
  OMEGA_BATCH_LAUNCH(k_batch_vdel4_edge);
  OMEGA_BATCH_LAUNCH(k_batch_vdel4_vc);
}

void launch_batch_tend(const BatchMember *desc, const int *blk, int nblocks, int threads, int N,
                       hipStream_t s) {
  TimerScope ts("batch_tend", s);
  OMEGA_BATCH_LAUNCH(k_batch_tend);
}

void launch_batch_diag_hex(const BatchMember *desc, const int *blk, int nblocks, int N,
                           hipStream_t s) {
  TimerScope ts("batch_diag", s);
  const int threads = 256;
  if (N == 2) {
    hipLaunchKernelGGL(k_batch_diag_hex2, dim3(nblocks), dim3(threads), 0, s, desc, blk);
    OMEGA_HIP_CHECK(hipGetLastError());
  } else {
    OMEGA_BATCH_LAUNCH(k_batch_diag_hex);
  }
}

void launch_batch_tend_hex(const BatchMember *desc, const int *blk, int nblocks, int N,
                           hipStream_t s) {
  TimerScope ts("batch_tend", s);
  const int threads = 256;
  if (N == 2) {
    hipLaunchKernelGGL(k_batch_tend_hex2, dim3(nblocks), dim3(threads), 0, s, desc, blk);
    OMEGA_HIP_CHECK(hipGetLastError());
  } else {
    OMEGA_BATCH_LAUNCH(k_batch_tend_hex);
  }
}

#undef OMEGA_BATCH_LAUNCH

} // namespace omega_hip
