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
// tables of hex_pattern.hpp.  Sums and products keep the reference's order,
// so results are the same bits as the general kernels.  The vector-memory
// address path of the CU is what bounds these kernels, so the aim throughout
// is the fewest load instructions per level.
#include "hex_ops.hpp"

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

// lanes per row = 64 / N (one row covers 64 levels); blockIdx.y walks the
// 64-level chunks
dim3 hex_grid(const MeshView &m, const Launch &lc, int N, dim3 &block) {
  const int lanes = 64 / N;
  int rows = lc.rows > 0 ? lc.rows : 256 / lanes;
  if (rows * lanes > 256) rows = 256 / lanes;
  block = dim3(lanes, rows);
  return dim3((m.nc + rows - 1) / rows, m.L / 64);
}

int hex_n(const MeshView &m, const Launch &lc) { return lc.lpt == 1 ? 1 : 2; }

} // namespace

bool hex_usable(const MeshView &m, const Launch &lc) {
  return m.hex_ok && (m.L % 64 == 0) && lc.variant != 1;
}

void launch_diag_hex(const MeshView &m, const Launch &lc, int terms, int ntr, const Real *h,
                     const Real *vn, const Real *tr, const DiagOut &d, hipStream_t s) {
  TimerScope ts("diag", s);
  dim3 block;
  const int N = hex_n(m, lc);
  const dim3 grid = hex_grid(m, lc, N, block);
  if (N == 2)
    hipLaunchKernelGGL(k_diag_hex<2>, grid, block, 0, s, m, terms, ntr, h, vn, tr, d, lc.swizzle);
  else
    hipLaunchKernelGGL(k_diag_hex<1>, grid, block, 0, s, m, terms, ntr, h, vn, tr, d, lc.swizzle);
  OMEGA_HIP_CHECK(hipGetLastError());
}

void launch_tend_hex(const MeshView &m, const Launch &lc, const ModelCoef &mc, const StageArgs &a,
                     const DiagOut &d, hipStream_t s) {
  TimerScope ts("tend", s);
  dim3 block;
  const int N = hex_n(m, lc);
  const dim3 grid = hex_grid(m, lc, N, block);
  if (N == 2)
    hipLaunchKernelGGL(k_tend_hex<2>, grid, block, 0, s, m, mc, a, d, lc.swizzle);
  else
    hipLaunchKernelGGL(k_tend_hex<1>, grid, block, 0, s, m, mc, a, d, lc.swizzle);
  OMEGA_HIP_CHECK(hipGetLastError());
}

} // namespace omega_hip
