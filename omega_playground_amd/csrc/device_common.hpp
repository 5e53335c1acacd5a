// Device helpers shared by the forward kernels.
#pragma once

#include "internal.hpp"

namespace omega_hip {
namespace dev {

constexpr int kTrGroup = 4; // tracers handled per pass in the cell kernels

template <int N> struct Vd {
  double v[N];
  __device__ double &operator[](int i) { return v[i]; }
  __device__ double operator[](int i) const { return v[i]; }
};

template <int N> __device__ __forceinline__ Vd<N> ld(const double *__restrict__ p) {
  Vd<N> r;
  if constexpr (N == 2) {
    const double2 t = *reinterpret_cast<const double2 *>(p);
    r.v[0] = t.x;
    r.v[1] = t.y;
  } else {
#pragma unroll
    for (int i = 0; i < N; ++i) r.v[i] = p[i];
  }
  return r;
}

template <int N> __device__ __forceinline__ void st(double *__restrict__ p, const Vd<N> &x) {
  if constexpr (N == 2) {
    *reinterpret_cast<double2 *>(p) = make_double2(x.v[0], x.v[1]);
  } else {
#pragma unroll
    for (int i = 0; i < N; ++i) p[i] = x.v[i];
  }
}

template <int N> __device__ __forceinline__ Vd<N> zero() {
  Vd<N> r;
#pragma unroll
  for (int i = 0; i < N; ++i) r.v[i] = 0.0;
  return r;
}

// Block b of G -> position on the curve such that XCD (b mod 8) processes a
// contiguous range (workgroups are dealt to the 8 XCDs round-robin).
__device__ __forceinline__ int logical_block(int b, int G, int swz) {
  if (!swz) return b;
  const int x = b & 7, i = b >> 3;
  const int q = G >> 3, r = G & 7;
  return x * q + min(x, r) + i;
}

// Stage epilogue: consumes the tendency T of one (entity, level-vector).
// E >= 0: the epilogue kind compiled in (it must equal a.epi).
template <int N, int E = -1>
__device__ __forceinline__ void epilogue(const StageArgs &a, const FieldIO &f, size_t idx,
                                         const Vd<N> &T) {
  switch (E >= 0 ? E : a.epi) {
  case EPI_STORE:
    st<N>(f.tend + idx, T);
    break;
  case EPI_RK4_FIRST: {
    // reference rk4.cpp:145-162 (state += dt*b*k) and :168-189 (provis)
    const Vd<N> y = ld<N>(f.in + idx);
    Vd<N> acc, out;
#pragma unroll
    for (int i = 0; i < N; ++i) {
      acc[i] = y[i] + a.db * T[i];
      out[i] = y[i] + a.da * T[i];
    }
    st<N>(f.acc + idx, acc);
    st<N>(f.out + idx, out);
    break;
  }
  case EPI_RK4_MID: {
    Vd<N> acc = ld<N>(f.acc + idx);
    const Vd<N> yo = ld<N>(f.yold + idx);
    Vd<N> out;
#pragma unroll
    for (int i = 0; i < N; ++i) {
      acc[i] = acc[i] + a.db * T[i];
      out[i] = yo[i] + a.da * T[i];
    }
    st<N>(f.acc + idx, acc);
    st<N>(f.out + idx, out);
    break;
  }
  case EPI_RK4_LAST: {
    Vd<N> acc = ld<N>(f.acc + idx);
#pragma unroll
    for (int i = 0; i < N; ++i) acc[i] = acc[i] + a.db * T[i];
    st<N>(f.out + idx, acc);
    break;
  }
  case EPI_LSRK: {
    // reference lsrk.cpp:70-87 (state += dt*b*tend); tend keeps T
    st<N>(f.tend + idx, T);
    const Vd<N> y = ld<N>(f.in + idx);
    Vd<N> out;
#pragma unroll
    for (int i = 0; i < N; ++i) out[i] = y[i] + a.db * T[i];
    st<N>(f.out + idx, out);
    break;
  }
  case EPI_PROVIS: {
    const Vd<N> yo = ld<N>(f.yold + idx);
    Vd<N> out;
#pragma unroll
    for (int i = 0; i < N; ++i) out[i] = yo[i] + a.da * T[i];
    st<N>(f.out + idx, out);
    break;
  }
  }
}

// The stage epilogue split in two so that its loads can be issued together
// with a kernel's stencil loads: epi_prefetch reads the operands (and the
// starting value of the tendency), epi_finish applies the same update as
// epilogue().
template <int N> struct EpiIn {
  Vd<N> base; // starting value of the tendency (tend_base)
  Vd<N> x, y; // FIRST/LSRK: x = y^s; MID: x = acc, y = yold; LAST: x = acc; PROVIS: y = yold
};

template <int N>
__device__ __forceinline__ EpiIn<N> epi_prefetch(const StageArgs &a, const FieldIO &f, size_t idx) {
  EpiIn<N> p;
  p.base = zero<N>();
  p.x = zero<N>();
  p.y = zero<N>();
  if (a.base_from_tend) p.base = ld<N>(f.tend + idx);
  switch (a.epi) {
  case EPI_RK4_FIRST:
  case EPI_LSRK:
    p.x = ld<N>(f.in + idx);
    break;
  case EPI_RK4_MID:
    p.x = ld<N>(f.acc + idx);
    p.y = ld<N>(f.yold + idx);
    break;
  case EPI_RK4_LAST:
    p.x = ld<N>(f.acc + idx);
    break;
  case EPI_PROVIS:
    p.y = ld<N>(f.yold + idx);
    break;
  default:
    break;
  }
  if (a.base_from_tend) {
#pragma unroll
    for (int i = 0; i < N; ++i) p.base[i] = p.base[i] * a.a;
  }
  return p;
}

template <int N>
__device__ __forceinline__ void epi_finish(const StageArgs &a, const FieldIO &f, size_t idx,
                                           const Vd<N> &T, const EpiIn<N> &p) {
  Vd<N> u, w;
  switch (a.epi) {
  case EPI_STORE:
    st<N>(f.tend + idx, T);
    break;
  case EPI_RK4_FIRST:
#pragma unroll
    for (int i = 0; i < N; ++i) {
      u[i] = p.x[i] + a.db * T[i];
      w[i] = p.x[i] + a.da * T[i];
    }
    st<N>(f.acc + idx, u);
    st<N>(f.out + idx, w);
    break;
  case EPI_RK4_MID:
#pragma unroll
    for (int i = 0; i < N; ++i) {
      u[i] = p.x[i] + a.db * T[i];
      w[i] = p.y[i] + a.da * T[i];
    }
    st<N>(f.acc + idx, u);
    st<N>(f.out + idx, w);
    break;
  case EPI_RK4_LAST:
#pragma unroll
    for (int i = 0; i < N; ++i) u[i] = p.x[i] + a.db * T[i];
    st<N>(f.out + idx, u);
    break;
  case EPI_LSRK:
    st<N>(f.tend + idx, T);
#pragma unroll
    for (int i = 0; i < N; ++i) u[i] = p.x[i] + a.db * T[i];
    st<N>(f.out + idx, u);
    break;
  case EPI_PROVIS:
#pragma unroll
    for (int i = 0; i < N; ++i) u[i] = p.y[i] + a.da * T[i];
    st<N>(f.out + idx, u);
    break;
  }
}

// Starting value of a tendency: 0 (replace) or a*tend (increment mode after
// the low-storage scaling tend *= a, lsrk.cpp:47-63).
template <int N>
__device__ __forceinline__ Vd<N> tend_base(const StageArgs &a, const FieldIO &f, size_t idx) {
  if (!a.base_from_tend) return zero<N>();
  Vd<N> t = ld<N>(f.tend + idx);
#pragma unroll
  for (int i = 0; i < N; ++i) t[i] = t[i] * a.a;
  return t;
}


// Value known to be equal in all lanes of the wavefront (one entity per
// wavefront): lets the compiler use scalar loads for its connectivity.
__device__ __forceinline__ int uniform(int x) { return __builtin_amdgcn_readfirstlane(x); }

} // namespace dev
} // namespace omega_hip
