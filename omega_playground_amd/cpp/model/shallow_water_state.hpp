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

  omega_state_t handle() const { return m_handle.get(); }
  const MPASMesh *mesh() const { return m_mesh; }
  Int ntracers() const { return m_ntracers; }

private:
  const MPASMesh *m_mesh;
  Int m_ntracers;
  std::shared_ptr<omega_state_s> m_handle;
};

} // namespace omega
