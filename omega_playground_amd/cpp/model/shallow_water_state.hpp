// ShallowWaterState (reference: src/model/shallow_water_state.hpp:9-16,
// shallow_water_state.cpp:7-14).
//
// The prognostic fields live on the device; m_h_cell, m_vn_edge and
// m_tr_cell are their host mirrors in the original numbering.  Fill the
// mirrors and call copy_to_device() before stepping, and copy_to_host()
// to read results -- the deep_copy pair a Kokkos host mirror needs.
// Copies of a state share the same device data and mirrors.
#pragma once

#include "../mesh/mpas_mesh.hpp"
#include "shallow_water_params.hpp"

namespace omega {

struct ShallowWaterState {
  Real2d m_h_cell;
  Real2d m_vn_edge;
  Real3d m_tr_cell;

  ShallowWaterState(const MPASMesh *mesh, Int ntracers)
      : m_h_cell("h_cell", mesh->m_ncells, mesh->m_nlayers),
        m_vn_edge("vn_edge", mesh->m_nedges, mesh->m_nlayers),
        m_tr_cell("tr_cell", ntracers, mesh->m_ncells, mesh->m_nlayers), m_mesh(mesh),
        m_ntracers(ntracers) {
    omega_state_t h = nullptr;
    check(omega_state_create(mesh->handle(), ntracers, &h));
    m_handle = std::shared_ptr<omega_state_s>(h, [](omega_state_t s) { omega_state_destroy(s); });
  }

  ShallowWaterState(const MPASMesh *mesh, const ShallowWaterParams &params)
      : ShallowWaterState(mesh, params.m_ntracers) {}

  // mirrors -> device
  void copy_to_device(void *stream = nullptr) const {
    check(omega_state_upload(handle(), m_h_cell.data(), m_vn_edge.data(),
                             m_ntracers ? m_tr_cell.data() : nullptr, stream));
  }
  // device -> mirrors (synchronises the stream)
  void copy_to_host(void *stream = nullptr) const {
    check(omega_state_download(handle(), m_h_cell.data(), m_vn_edge.data(),
                               m_ntracers ? m_tr_cell.data() : nullptr, stream));
  }
  // deep_copy(*this, other) on the device
  void copy_from(const ShallowWaterState &other, void *stream = nullptr) const {
    check(omega_state_copy(handle(), other.handle(), stream));
  }

  // The current device buffers of the fields (device numbering); low-storage
  // RK alternates two buffers per field, so query again after stepping.
  DeviceArray2d h_cell_device() const {
    double *h = nullptr;
    check(omega_state_device_ptrs(handle(), &h, nullptr, nullptr));
    return DeviceArray2d{h, m_mesh->m_ncells, m_mesh->m_nlayers};
  }
  DeviceArray2d vn_edge_device() const {
    double *vn = nullptr;
    check(omega_state_device_ptrs(handle(), nullptr, &vn, nullptr));
    return DeviceArray2d{vn, m_mesh->m_nedges, m_mesh->m_nlayers};
  }
  DeviceArray3d tr_cell_device() const {
    double *tr = nullptr;
    check(omega_state_device_ptrs(handle(), nullptr, nullptr, &tr));
    return DeviceArray3d{tr, m_ntracers, m_mesh->m_ncells, m_mesh->m_nlayers};
  }

  omega_state_t handle() const { return m_handle.get(); }
  const MPASMesh *mesh() const { return m_mesh; }
  Int ntracers() const { return m_ntracers; }

private:
  const MPASMesh *m_mesh;
  Int m_ntracers;
  std::shared_ptr<omega_state_s> m_handle;
};

} // namespace omega
