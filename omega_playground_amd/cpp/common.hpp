// Common types of the C++ drop-in API (the counterpart of the reference's
// src/common.hpp:1-113 without Kokkos).
//
// Arrays are host arrays with LayoutRight indexing and shared ownership:
// copying one aliases the same storage, as copying a Kokkos::View does.
// Mesh fields and state mirrors use the reference's original numbering;
// device data lives behind the C ABI (include/omega_swe.h).
#pragma once

#include <omega_swe.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstddef>
#include <memory>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

namespace omega {

using Real = double;
using Int = int;

constexpr Real pi = M_PI;

// Throws std::runtime_error carrying the library's message on a non-zero
// status (the reference aborts through Kokkos on failure).
inline void check(int rc) {
  if (rc != OMEGA_OK) throw std::runtime_error(std::string("omega: ") + omega_last_error());
}

template <class T, int R> class HostArray {
public:
  HostArray() = default;
  template <class... I> explicit HostArray(const std::string &label, I... extents)
      : m_label(label) {
    static_assert(sizeof...(I) == R, "wrong number of extents");
    const size_t e[R] = {static_cast<size_t>(extents)...};
    size_t n = 1;
    for (int d = 0; d < R; ++d) {
      m_ext[d] = e[d];
      n *= e[d];
    }
    m_data = std::make_shared<std::vector<T>>(n, T(0));
  }

  template <class... I> T &operator()(I... idx) const {
    static_assert(sizeof...(I) == R, "wrong number of indices");
    const size_t i[R] = {static_cast<size_t>(idx)...};
    size_t off = 0;
    for (int d = 0; d < R; ++d) off = off * m_ext[d] + i[d];
    return (*m_data)[off];
  }

  size_t extent(int d) const { return m_ext[d]; }
  int extent_int(int d) const { return static_cast<int>(m_ext[d]); }
  size_t size() const { return m_data ? m_data->size() : 0; }
  T *data() const { return m_data ? m_data->data() : nullptr; }
  const std::string &label() const { return m_label; }
  bool is_allocated() const { return static_cast<bool>(m_data); }

private:
  std::string m_label;
  std::array<size_t, R> m_ext{};
  std::shared_ptr<std::vector<T>> m_data;
};

using Real1d = HostArray<Real, 1>;
using Real2d = HostArray<Real, 2>;
using Real3d = HostArray<Real, 3>;
using Int1d = HostArray<Int, 1>;
using Int2d = HostArray<Int, 2>;

// deep_copy(array, value)
template <class T, int R> inline void deep_copy(const HostArray<T, R> &a, T v) {
  T *p = a.data();
  for (size_t i = 0; i < a.size(); ++i) p[i] = v;
}

// A device array of one state field in device numbering, as handed to the
// additional-tendency hook (rows = entities in device order, cols = layers).
struct DeviceArray2d {
  double *ptr = nullptr;
  Int rows = 0, cols = 0;
  double *data() const { return ptr; }
  Int extent(int d) const { return d == 0 ? rows : cols; }
};
struct ConstDeviceArray2d {
  const double *ptr = nullptr;
  Int rows = 0, cols = 0;
  ConstDeviceArray2d() = default;
  ConstDeviceArray2d(const double *p, Int r, Int c) : ptr(p), rows(r), cols(c) {}
  ConstDeviceArray2d(const DeviceArray2d &a) : ptr(a.ptr), rows(a.rows), cols(a.cols) {}
  const double *data() const { return ptr; }
  Int extent(int d) const { return d == 0 ? rows : cols; }
};

// A device tracer array [ntracers][cells][layers] in device numbering.
struct DeviceArray3d {
  double *ptr = nullptr;
  Int n0 = 0, n1 = 0, n2 = 0;
  double *data() const { return ptr; }
  Int extent(int d) const { return d == 0 ? n0 : d == 1 ? n1 : n2; }
};
struct ConstDeviceArray3d {
  const double *ptr = nullptr;
  Int n0 = 0, n1 = 0, n2 = 0;
  ConstDeviceArray3d() = default;
  ConstDeviceArray3d(const double *p, Int a, Int b, Int c) : ptr(p), n0(a), n1(b), n2(c) {}
  ConstDeviceArray3d(const DeviceArray3d &a) : ptr(a.ptr), n0(a.n0), n1(a.n1), n2(a.n2) {}
  const double *data() const { return ptr; }
  Int extent(int d) const { return d == 0 ? n0 : d == 1 ? n1 : n2; }
};

// Waits for the library's work on a stream (nullptr: the device), the
// counterpart of Kokkos::fence().
inline void fence(void *stream = nullptr) { check(omega_synchronize(stream)); }

} // namespace omega
