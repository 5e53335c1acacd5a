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
template <int N>
__device__ __forceinline__ void epilogue(const StageArgs &a, const FieldIO &f, size_t idx,
                                         const Vd<N> &T) {
  switch (a.epi) {
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
