// Forward kernels of the shallow-water hot path for gfx950 (MI355X).
//
// Layout: every field is [entity][level] with the level fastest, so the
// lanes of a row of threads cover consecutive levels of one entity and each
// neighbour access is one contiguous, coalesced row.  A block owns a stretch
// of cells along the space-filling curve (blockDim.y rows) and the edges and
// vertices those cells own; blockDim.x lanes times N levels per lane cover a
// chunk of levels (blockIdx.y).  Blocks are optionally renumbered so that
// each of the 8 XCDs walks a contiguous part of the curve and its neighbour
// re-reads stay in that XCD's L2.
//
// All operators are gathers: each output element is written by exactly one
// thread, no atomics are used, and every sum runs over the same terms in the
// same order as the reference, so results are bit-reproducible.  The library
// is compiled with -ffp-contract=off; together with host-precomputed leading
// factors (see mesh_host.cpp) the arithmetic is bit-identical to the
// reference's expressions.
#include "fwd_ops.hpp"

namespace omega_hip {

namespace {

using namespace dev;
using namespace fwd;

// ---------------------------------------------------------------------------
// Diagnostics
// ---------------------------------------------------------------------------
template <int N>
__global__ void __launch_bounds__(256)
    k_diag(MeshView m, int terms, int ntr, const double *__restrict__ h,
           const double *__restrict__ vn, const double *__restrict__ tr, DiagOut d, int swz) {
  const int s = logical_block(blockIdx.x, gridDim.x, swz);
  const int R = blockDim.y;
  const int c0 = s * R;
  if (c0 >= m.nc) return;
  const int c1 = min(c0 + R, m.nc);
  const int k = (blockIdx.y * blockDim.x + threadIdx.x) * N;
  if (k >= m.L) return;
  if (terms & (T_PV | T_VDEL2 | T_VDEL4)) {
    const int vb = m.vofs[c0], ve = m.vofs[c1];
    for (int v = vb + (int)threadIdx.y; v < ve; v += R) vertex_diag<N>(m, v, k, h, vn, d);
  }
  for (int c = c0 + (int)threadIdx.y; c < c1; c += R) cell_diag<N>(m, terms, ntr, c, k, h, vn, tr, d);
}

// ---------------------------------------------------------------------------
// Momentum del4 pre-passes (vel_hyperdiff_edge.hpp:140-210)
// ---------------------------------------------------------------------------
template <int N>
__global__ void __launch_bounds__(256) k_vdel4_edge(MeshView m, DiagOut d, int swz) {
  const int s = logical_block(blockIdx.x, gridDim.x, swz);
  const int R = blockDim.y;
  const int c0 = s * R;
  if (c0 >= m.nc) return;
  const int c1 = min(c0 + R, m.nc);
  const int k = (blockIdx.y * blockDim.x + threadIdx.x) * N;
  if (k >= m.L) return;
  const int eb = m.eofs[c0], ee = m.eofs[c1];
  for (int e = eb + (int)threadIdx.y; e < ee; e += R) vdel4_edge<N>(m, d, e, k);
}

template <int N>
__global__ void __launch_bounds__(256) k_vdel4_vc(MeshView m, DiagOut d, int swz) {
  const int s = logical_block(blockIdx.x, gridDim.x, swz);
  const int R = blockDim.y;
  const int c0 = s * R;
  if (c0 >= m.nc) return;
  const int c1 = min(c0 + R, m.nc);
  const int k = (blockIdx.y * blockDim.x + threadIdx.x) * N;
  if (k >= m.L) return;
  const int vb = m.vofs[c0], ve = m.vofs[c1];
  for (int v = vb + (int)threadIdx.y; v < ve; v += R) vdel4_vertex<N>(m, d, v, k);
  for (int c = c0 + (int)threadIdx.y; c < c1; c += R) vdel4_cell<N>(m, d, c, k);
}

// ---------------------------------------------------------------------------
// Tendencies + stage update
// ---------------------------------------------------------------------------

template <int N>
__global__ void __launch_bounds__(256)
    k_tend(MeshView m, ModelCoef mc, StageArgs a, DiagOut d, int swz) {
  const int s = logical_block(blockIdx.x, gridDim.x, swz);
  const int R = blockDim.y;
  const int c0 = s * R;
  if (c0 >= m.nc) return;
  const int c1 = min(c0 + R, m.nc);
  const int k = (blockIdx.y * blockDim.x + threadIdx.x) * N;
  if (k >= m.L) return;
  const int eb = m.eofs[c0], ee = m.eofs[c1];
  for (int e = eb + (int)threadIdx.y; e < ee; e += R) edge_tend<N>(m, mc, a, d, e, k);
  for (int c = c0 + (int)threadIdx.y; c < c1; c += R) cell_tend<N>(m, mc, a, d, c, k);
}

// ---------------------------------------------------------------------------
// Linear model (shallow_water_hierarchical.cpp:680-741) + stage update
// ---------------------------------------------------------------------------

// In replace mode the reference assigns the term itself (a -0 stays -0), in
// increment mode it adds it to the (scaled) tendency.
template <int N>
__device__ __forceinline__ Vd<N> lin_total(const StageArgs &a, const FieldIO &f, size_t idx,
                                           const Vd<N> &x) {
  if (!a.base_from_tend) return x;
  Vd<N> T = tend_base<N>(a, f, idx);
#pragma unroll
  for (int i = 0; i < N; ++i) T[i] = T[i] + x[i];
  return T;
}

template <int N>
__global__ void __launch_bounds__(256)
    k_lin(MeshView m, LinCoef lc, StageArgs a, int swz) {
  const int s = logical_block(blockIdx.x, gridDim.x, swz);
  const int R = blockDim.y;
  const int c0 = s * R;
  if (c0 >= m.nc) return;
  const int c1 = min(c0 + R, m.nc);
  const int k = (blockIdx.y * blockDim.x + threadIdx.x) * N;
  if (k >= m.L) return;
  const int L = m.L;
  const double *__restrict__ h = a.h.in;
  const double *__restrict__ vn = a.vn.in;
  const int eb = m.eofs[c0], ee = m.eofs[c1];
  for (int e = eb + (int)threadIdx.y; e < ee; e += R) {
    const size_t idx = (size_t)e * L + k;
    if (!a.do_vn) {
      if (a.epi != EPI_STORE) epilogue<N>(a, a.vn, idx, tend_base<N>(a, a.vn, idx));
      continue;
    }
    // compute_vn_tendency (:709-741)
    Vd<N> vt = zero<N>();
    const int n = m.noe[e];
    for (int j = 0; j < n; ++j) {
      const int ej = m.eoe[e * m.me2 + j];
      const double wj = m.w[e * m.me2 + j];
      const Vd<N> v = ld<N>(vn + (size_t)ej * L + k);
#pragma unroll
      for (int i = 0; i < N; ++i) vt[i] = vt[i] + wj * v[i];
    }
    const int a0 = m.coe[2 * e], a1 = m.coe[2 * e + 1];
    const Vd<N> h0 = ld<N>(h + (size_t)a0 * L + k), h1 = ld<N>(h + (size_t)a1 * L + k);
    const double fe = lc.f_e[e], dc = lc.dc[e];
    Vd<N> x;
#pragma unroll
    for (int i = 0; i < N; ++i) x[i] = fe * vt[i] - lc.grav * ((h1[i] - h0[i]) / dc);
    epilogue<N>(a, a.vn, idx, lin_total<N>(a, a.vn, idx, x));
  }
  for (int c = c0 + (int)threadIdx.y; c < c1; c += R) {
    const size_t idx = (size_t)c * L + k;
    if (!a.do_h) {
      if (a.epi != EPI_STORE) epilogue<N>(a, a.h, idx, tend_base<N>(a, a.h, idx));
      continue;
    }
    // compute_h_tendency (:680-707)
    Vd<N> acc = zero<N>();
    const int n = m.noc[c];
    for (int j = 0; j < n; ++j) {
      const int e = m.eoc[c * m.me + j];
      const double cd = lc.cdv[c * m.me + j];
      const Vd<N> v = ld<N>(vn + (size_t)e * L + k);
#pragma unroll
      for (int i = 0; i < N; ++i) acc[i] = acc[i] + cd * v[i];
    }
    const double area = lc.area[c];
    Vd<N> x;
#pragma unroll
    for (int i = 0; i < N; ++i) x[i] = -lc.h0 * acc[i] / area;
    epilogue<N>(a, a.h, idx, lin_total<N>(a, a.h, idx, x));
  }
}

// ---------------------------------------------------------------------------
// Element-wise helpers
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_axpy(double *__restrict__ y, double alpha,
                                              const double *__restrict__ x, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    y[i] = y[i] + alpha * x[i];
}

__global__ void __launch_bounds__(256) k_xpay(double *__restrict__ y, const double *__restrict__ x,
                                              double alpha, const double *__restrict__ z,
                                              size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    y[i] = x[i] + alpha * z[i];
}

__global__ void __launch_bounds__(256) k_scale(double *__restrict__ y, double alpha, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    y[i] = y[i] * alpha;
}

__global__ void __launch_bounds__(256) k_add_field(double *__restrict__ dst,
                                                   const double *__restrict__ src, int n, int L,
                                                   int ncols, double alpha) {
  const size_t total = (size_t)n * L;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < total;
       i += (size_t)gridDim.x * blockDim.x) {
    const size_t e = i / L, k = i % L;
    const double v = ncols == 1 ? src[e] : src[e * L + k];
    dst[i] = dst[i] + alpha * v;
  }
}

int grid_for(size_t n) {
  size_t b = (n + 255) / 256;
  if (b > 65536) b = 65536;
  return (int)(b ? b : 1);
}

struct Geo {
  dim3 grid, block;
};

Geo geometry(const MeshView &m, const Launch &lc, int N) {
  // lanes per row: enough to cover L levels with N levels per lane, capped
  // at 64 (one wavefront); further levels go to blockIdx.y
  // levels per block: lc.chunk (0 = up to one wavefront of lanes); smaller
  // chunks make the grid sweep the mesh once per level chunk, which shrinks
  // the working set the neighbour gathers hit in L2
  const int want = lc.chunk > 0 ? std::min(lc.chunk, m.L) : m.L;
  int lanes = (want + N - 1) / N;
  if (lanes > 64) lanes = 64;
  int lx = 1;
  while (lx < lanes) lx <<= 1;
  // rows = 0: fill a 256-thread block
  int rows = lc.rows > 0 ? lc.rows : 256 / lx;
  if (lx * rows > 256) rows = 256 / lx;
  if (rows < 1) rows = 1;
  const int chunks = (m.L + lx * N - 1) / (lx * N);
  const int blocks = (m.nc + rows - 1) / rows;
  return Geo{dim3(blocks, chunks), dim3(lx, rows)};
}

} // namespace

void launch_diag(const MeshView &m, const Launch &lc, const ModelCoef &mc, int terms, int ntr,
                 const Real *h, const Real *vn, const Real *tr, const DiagOut &d, hipStream_t s) {
  (void)mc;
  TimerScope ts("diag", s);
  const int N = (lc.lpt == 2 && m.L % 2 == 0) ? 2 : 1;
  const Geo g = geometry(m, lc, N);
  if (N == 2)
    hipLaunchKernelGGL(k_diag<2>, g.grid, g.block, 0, s, m, terms, ntr, h, vn, tr, d, lc.swizzle);
  else
    hipLaunchKernelGGL(k_diag<1>, g.grid, g.block, 0, s, m, terms, ntr, h, vn, tr, d, lc.swizzle);
  OMEGA_HIP_CHECK(hipGetLastError());
}

void launch_vdel4(const MeshView &m, const Launch &lc, const DiagOut &d, hipStream_t s) {
  TimerScope ts("vdel4", s);
  const int N = (lc.lpt == 2 && m.L % 2 == 0) ? 2 : 1;
  const Geo g = geometry(m, lc, N);
  if (N == 2) {
    hipLaunchKernelGGL(k_vdel4_edge<2>, g.grid, g.block, 0, s, m, d, lc.swizzle);
    hipLaunchKernelGGL(k_vdel4_vc<2>, g.grid, g.block, 0, s, m, d, lc.swizzle);
  } else {
    hipLaunchKernelGGL(k_vdel4_edge<1>, g.grid, g.block, 0, s, m, d, lc.swizzle);
    hipLaunchKernelGGL(k_vdel4_vc<1>, g.grid, g.block, 0, s, m, d, lc.swizzle);
  }
  OMEGA_HIP_CHECK(hipGetLastError());
}

void launch_tend(const MeshView &m, const Launch &lc, const ModelCoef &mc, const StageArgs &a,
                 const DiagOut &d, hipStream_t s) {
  TimerScope ts("tend", s);
  const int N = (lc.lpt == 2 && m.L % 2 == 0) ? 2 : 1;
  const Geo g = geometry(m, lc, N);
  if (N == 2)
    hipLaunchKernelGGL(k_tend<2>, g.grid, g.block, 0, s, m, mc, a, d, lc.swizzle);
  else
    hipLaunchKernelGGL(k_tend<1>, g.grid, g.block, 0, s, m, mc, a, d, lc.swizzle);
  OMEGA_HIP_CHECK(hipGetLastError());
}

void launch_tend_linear(const MeshView &m, const Launch &lc, const LinCoef &lin,
                        const StageArgs &a, hipStream_t s) {
  TimerScope ts("tend", s);
  const int N = (lc.lpt == 2 && m.L % 2 == 0) ? 2 : 1;
  const Geo g = geometry(m, lc, N);
  if (N == 2)
    hipLaunchKernelGGL(k_lin<2>, g.grid, g.block, 0, s, m, lin, a, lc.swizzle);
  else
    hipLaunchKernelGGL(k_lin<1>, g.grid, g.block, 0, s, m, lin, a, lc.swizzle);
  OMEGA_HIP_CHECK(hipGetLastError());
}

void launch_axpy(Real *y, Real alpha, const Real *x, size_t n, hipStream_t s) {
  if (!n) return;
  TimerScope ts("update", s);
  hipLaunchKernelGGL(k_axpy, dim3(grid_for(n)), dim3(256), 0, s, y, alpha, x, n);
  OMEGA_HIP_CHECK(hipGetLastError());
}

void launch_xpay(Real *y, const Real *x, Real alpha, const Real *z, size_t n, hipStream_t s) {
  if (!n) return;
  TimerScope ts("update", s);
  hipLaunchKernelGGL(k_xpay, dim3(grid_for(n)), dim3(256), 0, s, y, x, alpha, z, n);
  OMEGA_HIP_CHECK(hipGetLastError());
}

void launch_scale(Real *y, Real alpha, size_t n, hipStream_t s) {
  if (!n) return;
  TimerScope ts("update", s);
  hipLaunchKernelGGL(k_scale, dim3(grid_for(n)), dim3(256), 0, s, y, alpha, n);
  OMEGA_HIP_CHECK(hipGetLastError());
}

void launch_add_field(Real *dst, const Real *src, int n, int L, int ncols, Real alpha,
                      hipStream_t s) {
  const size_t total = (size_t)n * L;
  if (!total) return;
  hipLaunchKernelGGL(k_add_field, dim3(grid_for(total)), dim3(256), 0, s, dst, src, n, L, ncols,
                     alpha);
  OMEGA_HIP_CHECK(hipGetLastError());
}

} // namespace omega_hip
