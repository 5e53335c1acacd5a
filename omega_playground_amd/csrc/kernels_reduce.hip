// Deterministic global integrals (energy, mass, circulation).
//
// Replaces the Kokkos parallel_reduce kernels of
// src/model/shallow_water_diagnostics.cpp:5-89.  Each block sums a fixed
// range of entities (each entity's levels in order) through a fixed LDS
// tree into one partial; a single block then folds the partials in a fixed
// order.  The result is bit-identical from run to run.
#include "internal.hpp"

namespace omega_hip {

namespace {

constexpr int kRB = 256;

__device__ double block_sum(double v, double *sh) {
  const int t = threadIdx.x;
  sh[t] = v;
  __syncthreads();
  for (int s = kRB / 2; s > 0; s >>= 1) {
    if (t < s) sh[t] = sh[t] + sh[t + s];
    __syncthreads();
  }
  return sh[0];
}

struct DiagGeom {
  int nc, nv, me, L;
  const int *noc, *eoc, *eov;
  const double *ae_c;   // dv*dc of each cell edge [nc][me]
  const double *area_c; // [nc]
  const double *dcs_v;  // dc*sign of each vertex edge [nv][3]
  const double *area_t; // [nv]
  const double *f_v;    // [nv]
  double grav;
  double h0; // >= 0: the linear model's mean thickness weighs K; < 0: h does
};

// which: 0 energy, 1 mass, 2 circulation
__global__ void __launch_bounds__(kRB)
    k_reduce_partial(DiagGeom g, int which, const double *__restrict__ h,
                     const double *__restrict__ vn, double *__restrict__ partial) {
  __shared__ double sh[kRB];
  const int i = blockIdx.x * kRB + threadIdx.x;
  const int L = g.L;
  double acc = 0.0;
  if (which == 0 && i < g.nc) {
    // shallow_water_diagnostics.cpp:22-35
    const double A = g.area_c[i];
    for (int k = 0; k < L; ++k) {
      double K = 0;
      for (int j = 0; j < g.noc[i]; ++j) {
        const int e = g.eoc[i * g.me + j];
        const double v = vn[(size_t)e * L + k];
        K += g.ae_c[i * g.me + j] * v * v / 4;
      }
      K /= A;
      const double hc = h[(size_t)i * L + k];
      // the linear model weighs K with h0 (shallow_water_hierarchical.cpp:767-769)
      acc += A * (g.grav * hc * hc / 2 + (g.h0 >= 0 ? g.h0 : hc) * K);
    }
  } else if (which == 1 && i < g.nc) {
    // shallow_water_diagnostics.cpp:51-55
    const double A = g.area_c[i];
    for (int k = 0; k < L; ++k) acc += A * h[(size_t)i * L + k];
  } else if (which == 2 && i < g.nv) {
    // shallow_water_diagnostics.cpp:75-86
    for (int k = 0; k < L; ++k) {
      double cir = -0.0;
      for (int j = 0; j < 3; ++j) {
        const int e = g.eov[3 * i + j];
        cir += g.dcs_v[3 * i + j] * vn[(size_t)e * L + k];
      }
      acc += cir + g.f_v[i] * g.area_t[i];
    }
  }
  const double s = block_sum(acc, sh);
  if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

__global__ void __launch_bounds__(kRB)
    k_reduce_final(const double *__restrict__ partial, int n, double *__restrict__ out) {
  __shared__ double sh[kRB];
  double acc = 0.0;
  for (int i = threadIdx.x; i < n; i += kRB) acc += partial[i];
  const double s = block_sum(acc, sh);
  if (threadIdx.x == 0) *out = s;
}

} // namespace

int reduce_blocks(int n) { return (n + kRB - 1) / kRB; }

void launch_reduce(int which, int nc, int nv, int me, int L, const int *noc, const int *eoc,
                   const int *eov, const double *ae_c, const double *area_c, const double *dcs_v,
                   const double *area_t, const double *f_v, double grav, const double *h,
                   const double *vn, double *partial, double *out, hipStream_t s,
                   double linear_h0) {
  TimerScope ts("reduce", s);
  DiagGeom g{nc, nv, me, L, noc, eoc, eov, ae_c, area_c, dcs_v, area_t, f_v, grav, linear_h0};
  const int n = which == 2 ? nv : nc;
  const int nb = reduce_blocks(n);
  hipLaunchKernelGGL(k_reduce_partial, dim3(nb), dim3(kRB), 0, s, g, which, h, vn, partial);
  hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(kRB), 0, s, partial, nb, out);
  OMEGA_HIP_CHECK(hipGetLastError());
}

} // namespace omega_hip
