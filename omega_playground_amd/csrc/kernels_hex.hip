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

namespace omega_hip {

namespace {

using namespace dev;
using namespace hexops;

// The grid is one-dimensional over (cell block, 64-level chunk) pairs.
// Workgroups are dealt to the 8 XCDs round-robin; with a multiple of 8
// chunks the chunk runs fastest, so XCD x handles chunks x, x + 8, ... of
// every cell block (per XCD the working set of a model with an eighth of the
// levels; the stacked ensemble adjoint's 1024-level rows ran 5 % faster than
// with the chunk slowest); otherwise the chunk runs slowest and the optional
// swizzle gives each XCD a contiguous range of cell blocks.  s: the cell
// block; returns this thread's first level.
__device__ __forceinline__ int block_rows(const MeshView &m, int swz, int N, int &s) {
  const int nch = m.L % 64 == 0 ? m.L / 64 : 1;
  int ch;
  if (nch == 1) {
    // one chunk: no division in the block's prologue (it measured 0.6 % of
    // the tendency kernel at nx=512, 64 layers)
    s = logical_block(blockIdx.x, gridDim.x, swz);
    ch = 0;
  } else if (nch % 8 == 0) {
    s = uniform((int)blockIdx.x / nch);
    ch = (int)blockIdx.x - s * nch;
  } else {
    const int ncb = (int)gridDim.x / nch;
    const int lb = logical_block(blockIdx.x, gridDim.x, swz);
    ch = uniform(lb / ncb);
    s = lb - ch * ncb;
  }
  return (ch * (int)blockDim.x + (int)threadIdx.x) * N;
}

template <int N>
__global__ void __launch_bounds__(256)
    k_diag_hex(MeshView m, int terms, int ntr, const double *__restrict__ h,
               const double *__restrict__ vn, const double *__restrict__ tr, DiagOut d, int swz) {
  int s;
  const int k = block_rows(m, swz, N, s);
  const int R = blockDim.y;
  const int c = ent<N>(s * R + (int)threadIdx.y);
  if (c >= m.nc) return;
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
  int s;
  const int k = block_rows(m, swz, 2, s);
  const int R = blockDim.y;
  const int cA = uniform(s * R + ((int)threadIdx.y & ~1));
  if (cA >= m.nc) return;
  const int cB = min(cA + 1, m.nc - 1);
  const bool hi = (threadIdx.y & 1) != 0;
  if (s * R + (int)threadIdx.y >= m.nc) return;
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

// EM: the edges also write their PreOut rows (po)
template <int N, bool EM = false>
__global__ void __launch_bounds__(256)
    k_tend_hex(MeshView m, ModelCoef mc, StageArgs a, DiagOut d, int swz, PreOut po) {
  int s;
  const int k = block_rows(m, swz, N, s);
  const int R = blockDim.y;
  const int c = ent<N>(s * R + (int)threadIdx.y);
  if (c >= m.nc) return;
  const int e0 = 3 * c;
  edge_tend_hex<N, 0, -1, EM>(m, mc, a, d, e0, k, po);
  edge_tend_hex<N, 1, -1, EM>(m, mc, a, d, e0 + 1, k, po);
  edge_tend_hex<N, 2, -1, EM>(m, mc, a, d, e0 + 2, k, po);
  cell_tend_hex<N>(m, mc, a, d, c, k);
}

// N = 2 with two cells per wavefront and their tables read through the
// scalar path (hex_ops.hpp: edge_tend_hex2 / cell_tend_hex2); the block's
// row count is even, so the cells of one wavefront are cA and cA + 1
template <bool U, bool W>
__global__ void __launch_bounds__(256)
    k_tend_hex2(MeshView m, ModelCoef mc, StageArgs a, DiagOut d, int swz) {
  int s;
  const int k = block_rows(m, swz, 2, s);
  const int R = blockDim.y;
  const int cA = uniform(s * R + ((int)threadIdx.y & ~1));
  if (cA >= m.nc) return;
  const int cB = min(cA + 1, m.nc - 1);
  const bool hi = (threadIdx.y & 1) != 0;
  if (s * R + (int)threadIdx.y >= m.nc) return;
  edge_tend_hex2<0, U, W>(m, mc, a, d, 3 * cA, 3 * cB, hi, k);
  edge_tend_hex2<1, U, W>(m, mc, a, d, 3 * cA + 1, 3 * cB + 1, hi, k);
  edge_tend_hex2<2, U, W>(m, mc, a, d, 3 * cA + 2, 3 * cB + 2, hi, k);
  cell_tend_hex2<U, W>(m, mc, a, d, cA, cB, hi, k);
}

// vertex-union form: the PV products of the cell's 16 union vertices are
// formed once and shared by its three edges (needs the per-cell union
// records, hexu).  TM >= 0: the term set TM with both tendencies on,
// resolved at compile time (no run-time term branches between the stencil
// phases, so the compiler schedules across them); E >= 0: the stage
// epilogue likewise; -1: read from a.
template <bool U, bool W, int TM, int E, bool EM>
__device__ __forceinline__ void tend_hex2v(const MeshView &m, const ModelCoef &mc,
                                           const StageArgs &a, const DiagOut &d, int swz,
                                           const PreOut &po) {
  int s;
  const int k = block_rows(m, swz, 2, s);
  const int R = blockDim.y;
  const int cA = uniform(s * R + ((int)threadIdx.y & ~1));
  if (cA >= m.nc) return;
  const int cB = min(cA + 1, m.nc - 1);
  const bool hi = (threadIdx.y & 1) != 0;
  if (s * R + (int)threadIdx.y >= m.nc) return;
  Vd<2> P[hexu::kNV], Q[hexu::kNV];
  // a run-time test also for TM >= 0 (where it is always true): the branch
  // keeps the compiler from hoisting the edges' loads above the union PV
  // products, which measured 6 % slower (1.957 against 1.850 ms at nx=512)
  if (a.do_vn && (a.terms & T_PV)) union_pv2<U, W>(m, mc, d, cA, cB, hi, k, P, Q);
  edge_tend_hex2v<0, U, W, TM, E, EM>(m, mc, a, d, 3 * cA, 3 * cB, hi, k, P, Q, po);
  edge_tend_hex2v<1, U, W, TM, E, EM>(m, mc, a, d, 3 * cA + 1, 3 * cB + 1, hi, k, P, Q, po);
  edge_tend_hex2v<2, U, W, TM, E, EM>(m, mc, a, d, 3 * cA + 2, 3 * cB + 2, hi, k, P, Q, po);
  cell_tend_hex2<U, W, TM, E>(m, mc, a, d, cA, cB, hi, k);
}

template <bool U, bool W, int TM = -1, int E = -1>
__global__ void __launch_bounds__(256)
    k_tend_hex2v(MeshView m, ModelCoef mc, StageArgs a, DiagOut d, int swz, PreOut po) {
  tend_hex2v<U, W, TM, E, false>(m, mc, a, d, swz, po);
}

// The same with the VJP rows written (the adjoint's stage recomputation:
// the RK4 provisional-stage or low-storage epilogue, compiled in with the
// term set); held at two waves per SIMD, where the stores' extra registers
// would otherwise cost half the occupancy
template <int TM, int E>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2)))
    k_tend_hex2v_pre(MeshView m, ModelCoef mc, StageArgs a, DiagOut d, int swz, PreOut po) {
  tend_hex2v<true, false, TM, E, true>(m, mc, a, d, swz, po);
}

// Momentum del4 pre-pass, two cells per wavefront: del2 of the velocity on
// each cell's three owned edges, then (after the whole grid) its curl at the
// cell's two vertices and its divergence at the cell
template <bool U, bool W>
__global__ void __launch_bounds__(256) k_vdel4_edge_hex2(MeshView m, DiagOut d, int swz) {
  int s;
  const int k = block_rows(m, swz, 2, s);
  const int R = blockDim.y;
  const int cA = uniform(s * R + ((int)threadIdx.y & ~1));
  if (cA >= m.nc) return;
  const int cB = min(cA + 1, m.nc - 1);
  const bool hi = (threadIdx.y & 1) != 0;
  if (s * R + (int)threadIdx.y >= m.nc) return;
  vdel4_edge_p<U, W>(m, d, 3 * cA, 3 * cB, hi, k);
  vdel4_edge_p<U, W>(m, d, 3 * cA + 1, 3 * cB + 1, hi, k);
  vdel4_edge_p<U, W>(m, d, 3 * cA + 2, 3 * cB + 2, hi, k);
}

template <bool U, bool W>
__global__ void __launch_bounds__(256) k_vdel4_vc_hex2(MeshView m, DiagOut d, int swz) {
  int s;
  const int k = block_rows(m, swz, 2, s);
  const int R = blockDim.y;
  const int cA = uniform(s * R + ((int)threadIdx.y & ~1));
  if (cA >= m.nc) return;
  const int cB = min(cA + 1, m.nc - 1);
  const bool hi = (threadIdx.y & 1) != 0;
  if (s * R + (int)threadIdx.y >= m.nc) return;
  vdel4_vertex_p<U, W>(m, d, 2 * cA, 2 * cB, hi, k);
  vdel4_vertex_p<U, W>(m, d, 2 * cA + 1, 2 * cB + 1, hi, k);
  vdel4_cell_p<U, W>(m, d, cA, cB, hi, k);
}

// Term sets compiled into the shared-coefficient, 32-bit-offset form:
// momentum and thickness only (7), with del2 viscosity (15), with del2
// viscosity and tracer del4 (79), everything (127); each also with the
// epilogue of a time step's stages compiled in (RK4 first, middle and last
// stage, low-storage RK stage)
using TendFn = void (*)(MeshView, ModelCoef, StageArgs, DiagOut, int, PreOut);
template <int TM> TendFn tend_for_epilogue(int epi) {
  switch (epi) {
  case EPI_RK4_FIRST: return k_tend_hex2v<true, false, TM, EPI_RK4_FIRST>;
  case EPI_RK4_MID: return k_tend_hex2v<true, false, TM, EPI_RK4_MID>;
  case EPI_RK4_LAST: return k_tend_hex2v<true, false, TM, EPI_RK4_LAST>;
  case EPI_LSRK: return k_tend_hex2v<true, false, TM, EPI_LSRK>;
  default: return k_tend_hex2v<true, false, TM>;
  }
}
// the forms that also write the VJP rows (the adjoint's stage recomputation)
template <int TM> TendFn tend_writing_pre_epi(int epi) {
  switch (epi) {
  case EPI_PROVIS: return k_tend_hex2v_pre<TM, EPI_PROVIS>;
  case EPI_LSRK: return k_tend_hex2v_pre<TM, EPI_LSRK>;
  default: return k_tend_hex2v_pre<-1, -1>;
  }
}
TendFn tend_writing_pre(int terms, bool term_sets, int epi) {
  switch (term_sets ? terms : -1) {
  case 7: return tend_writing_pre_epi<7>(epi);
  case 15: return tend_writing_pre_epi<15>(epi);
  case 79: return tend_writing_pre_epi<79>(epi);
  case 127: return tend_writing_pre_epi<127>(epi);
  default: return k_tend_hex2v_pre<-1, -1>;
  }
}
TendFn tend_term_set(int terms, int epi) {
  switch (terms) {
  case 7: return tend_for_epilogue<7>(epi);
  case 15: return tend_for_epilogue<15>(epi);
  case 79: return tend_for_epilogue<79>(epi);
  case 127: return tend_for_epilogue<127>(epi);
  default: return nullptr;
  }
}

// lanes per row = 64 / N (one row covers 64 levels); the one-dimensional
// grid walks (cell block, 64-level chunk) pairs (block_rows)
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
  return dim3((m.nc + rows - 1) / rows * (m.L / span));
}

int hex_n(const MeshView &m, const Launch &lc) { return lc.lpt == 1 ? 1 : 2; }

} // namespace

bool hex_term_set(int terms) { return tend_term_set(terms, -1) != nullptr; }

bool hex_writes_pre(const MeshView &m, const Launch &lc, const ModelCoef &mc) {
  // the one-cell-per-row kernels, and the pair kernel's vertex-union form with
  // shared coefficients and 32-bit offsets
  return !hex_pair(m, hex_n(m, lc)) || (m.hexu_ok && m.coef_uni && mc.uniform && m.off32);
}

bool hex_usable(const MeshView &m, const Launch &lc) {
  // 16 or 32 layers: several cells per wavefront, two levels per lane
  const bool rows_ok = m.L % 64 == 0 || ((m.L == 16 || m.L == 32) && lc.lpt != 1);
  return m.hex_ok && rows_ok && lc.variant != 1;
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

bool hex_vdel4(const MeshView &m, const Launch &lc) {
  return hex_usable(m, lc) && hex_pair(m, hex_n(m, lc));
}

void launch_vdel4_hex(const MeshView &m, const Launch &lc, const DiagOut &d, hipStream_t s) {
  TimerScope ts("vdel4", s);
  OMEGA_REQUIRE(hex_vdel4(m, lc), "del4 pre-pass: the pair layout needs rows of 64 levels");
  dim3 block;
  const dim3 grid = hex_grid(m, lc, 2, block);
  auto ke = m.coef_uni ? (m.off32 ? k_vdel4_edge_hex2<true, false> : k_vdel4_edge_hex2<true, true>)
                       : (m.off32 ? k_vdel4_edge_hex2<false, false> : k_vdel4_edge_hex2<false, true>);
  auto kv = m.coef_uni ? (m.off32 ? k_vdel4_vc_hex2<true, false> : k_vdel4_vc_hex2<true, true>)
                       : (m.off32 ? k_vdel4_vc_hex2<false, false> : k_vdel4_vc_hex2<false, true>);
  hipLaunchKernelGGL(ke, grid, block, 0, s, m, d, lc.swizzle);
  hipLaunchKernelGGL(kv, grid, block, 0, s, m, d, lc.swizzle);
  OMEGA_HIP_CHECK(hipGetLastError());
}

void launch_tend_hex(const MeshView &m, const Launch &lc, const ModelCoef &mc, const StageArgs &a,
                     const DiagOut &d, hipStream_t s, const PreOut *pre) {
  TimerScope ts("tend", s);
  dim3 block;
  const int N = hex_n(m, lc);
  const dim3 grid = hex_grid(m, lc, N, block);
  const bool U = m.coef_uni && mc.uniform;
  OMEGA_REQUIRE(!pre || (hex_writes_pre(m, lc, mc) && a.do_vn),
                "tendency launch: this kernel form does not write the VJP rows");
  const PreOut po = pre ? *pre : PreOut{};
  if (hex_pair(m, N) && m.hexu_ok) {
    auto kern = U ? (m.off32 ? k_tend_hex2v<true, false> : k_tend_hex2v<true, true>)
                  : (m.off32 ? k_tend_hex2v<false, false> : k_tend_hex2v<false, true>);
    if (U && m.off32 && m.term_sets && a.do_h && a.do_vn)
      if (auto ks = tend_term_set(a.terms, a.epi)) kern = ks;
    if (pre) kern = tend_writing_pre(a.terms, m.term_sets && a.do_h && a.do_vn, a.epi);
    hipLaunchKernelGGL(kern, grid, block, 0, s, m, mc, a, d, lc.swizzle, po);
  } else if (hex_pair(m, N)) {
    auto kern = U ? (m.off32 ? k_tend_hex2<true, false> : k_tend_hex2<true, true>)
                  : (m.off32 ? k_tend_hex2<false, false> : k_tend_hex2<false, true>);
    hipLaunchKernelGGL(kern, grid, block, 0, s, m, mc, a, d, lc.swizzle);
  } else if (N == 2) {
    // (compiling the term set into this form measured slower: 1.34 against
    // 1.18 ms per RK4 step at nx=256, 32 layers)
    auto kern = pre ? k_tend_hex<2, true> : k_tend_hex<2, false>;
    hipLaunchKernelGGL(kern, grid, block, 0, s, m, mc, a, d, lc.swizzle, po);
  } else {
    auto kern = pre ? k_tend_hex<1, true> : k_tend_hex<1, false>;
    hipLaunchKernelGGL(kern, grid, block, 0, s, m, mc, a, d, lc.swizzle, po);
  }
  OMEGA_HIP_CHECK(hipGetLastError());
}

} // namespace omega_hip
