// ShallowWaterModel (reference: src/model/shallow_water_model.hpp:10-51,
// shallow_water_model.cpp:5-347).
//
// The tendency terms run as fused HIP kernels behind the C ABI.  A derived
// class may override additional_tendency() as in the reference; it receives
// the tendency and state arrays on the device (device numbering, see
// MPASMesh permutations) and the stream the library works on, and must only
// enqueue work on that stream (add_host_field() adds a host-computed field).
#pragma once

#include "../mesh/mpas_mesh.hpp"
#include "shallow_water_params.hpp"
#include "shallow_water_state.hpp"

namespace omega {

enum class AddMode { replace, increment };

struct ShallowWaterModel {
  ShallowWaterParams m_params;
  MPASMesh *m_mesh;
  // Coriolis parameter at vertices and edges (host); set_f_vertex() sends a
  // changed m_f_vertex to the device.
  Real1d m_f_vertex;
  Real1d m_f_edge;

  ShallowWaterModel(MPASMesh *mesh, const ShallowWaterParams &params)
      : ShallowWaterModel(mesh, params, nullptr) {}

protected:
  // linear_h0 != nullptr: LinearShallowWaterModel with that mean thickness
  ShallowWaterModel(MPASMesh *mesh, const ShallowWaterParams &params, const Real *linear_h0)
      : m_params(params), m_mesh(mesh), m_f_vertex("f_vertex", mesh->m_nvertices),
        m_f_edge("f_edge", mesh->m_nedges) {
    const omega_params c = to_c(params);
    if (linear_h0)
      check(omega_model_create_linear(mesh->handle(), &c, *linear_h0, &m_handle));
    else
      check(omega_model_create(mesh->handle(), &c, &m_handle));
    deep_copy(m_f_vertex, params.m_f0);
    deep_copy(m_f_edge, params.m_f0);
    // The hook is registered for every model; the base implementation
    // reports itself on its first call and is then removed, so models
    // without an override keep the fused stage kernels.
    check(omega_model_set_additional_tendency(m_handle, &ShallowWaterModel::hook, this));
  }

public:
  ShallowWaterModel(const ShallowWaterModel &) = delete;
  ShallowWaterModel &operator=(const ShallowWaterModel &) = delete;
  virtual ~ShallowWaterModel() {
    if (m_handle) omega_model_destroy(m_handle);
  }

  virtual void additional_tendency(DeviceArray2d h_tend_cell, DeviceArray2d vn_tend_edge,
                                   ConstDeviceArray2d h_cell, ConstDeviceArray2d vn_edge,
                                   Real t) const {
    (void)h_tend_cell;
    (void)vn_tend_edge;
    (void)h_cell;
    (void)vn_edge;
    (void)t;
    m_base_hook_called = true;
  }

  void compute_tendency(const ShallowWaterState &tend, const ShallowWaterState &state, Real t,
                        AddMode add_mode = AddMode::replace) const {
    check(omega_model_compute_tendency(m_handle, tend.handle(), state.handle(), t,
                                       add_mode == AddMode::increment, m_stream));
    drop_noop_hook();
  }

  // compute_h/vn/tr_tendency (shallow_water_model.hpp:35-40) on whole
  // states: only the named field of tend is written.
  void compute_h_tendency(const ShallowWaterState &tend, const ShallowWaterState &state,
                          AddMode add_mode = AddMode::replace) const {
    field_tendency(0, tend, state, add_mode);
  }
  void compute_vn_tendency(const ShallowWaterState &tend, const ShallowWaterState &state,
                           AddMode add_mode = AddMode::replace) const {
    field_tendency(1, tend, state, add_mode);
  }
  void compute_tr_tendency(const ShallowWaterState &tend, const ShallowWaterState &state,
                           AddMode add_mode = AddMode::replace) const {
    field_tendency(2, tend, state, add_mode);
  }

  // m_aux_state.compute(h, vn, tr, f_vertex, mesh)
  // (shallow_water_auxstate.cpp:12-50) for a state's current buffers.
  void compute_aux(const ShallowWaterState &state) const {
    const DeviceArray3d tr = state.tr_cell_device();
    check(omega_model_compute_aux(m_handle, state.h_cell_device().ptr,
                                  state.vn_edge_device().ptr, tr.ptr, m_stream));
  }

  // The reference's signatures (shallow_water_model.hpp:35-40) on device
  // arrays: they read the auxiliary state of the last compute_aux(),
  // compute_tendency() or step, as the reference reads m_aux_state.
  void compute_h_tendency(DeviceArray2d h_tend_cell, ConstDeviceArray2d h_cell,
                          ConstDeviceArray2d vn_edge, AddMode add_mode) const {
    check(omega_model_compute_h_tendency(m_handle, h_tend_cell.ptr, h_cell.ptr, vn_edge.ptr,
                                         add_mode == AddMode::increment, m_stream));
  }
  void compute_vn_tendency(DeviceArray2d vn_tend_edge, ConstDeviceArray2d h_cell,
                           ConstDeviceArray2d vn_edge, AddMode add_mode) const {
    check(omega_model_compute_vn_tendency(m_handle, vn_tend_edge.ptr, h_cell.ptr, vn_edge.ptr,
                                          add_mode == AddMode::increment, m_stream));
  }
  void compute_tr_tendency(DeviceArray3d tr_tend_cell, ConstDeviceArray3d tr_cell,
                           ConstDeviceArray2d vn_edge, AddMode add_mode) const {
    check(omega_model_compute_tr_tendency(m_handle, tr_tend_cell.ptr, tr_cell.ptr, vn_edge.ptr,
                                          add_mode == AddMode::increment, m_stream));
  }

  // Opt-in bottom drag with m_params.m_drag_coeff (the reference carries the
  // coefficient but applies no drag; its formula is a comment at
  // shallow_water_hierarchical.cpp:479-485).  Off by default.
  void enable_bottom_drag(bool on = true) const {
    int mask = 0;
    check(omega_model_get_terms(m_handle, &mask));
    check(omega_model_set_terms(m_handle, on ? (mask | 128) : (mask & ~128)));
  }

  // Sends m_f_vertex to the device (the reference's kernels read the view
  // directly).
  void set_f_vertex() const { check(omega_model_set_f_vertex(m_handle, m_f_vertex.data())); }

  // Stream of every launch made through this model and its steppers
  // (hipStream_t; nullptr is the null stream).
  void set_stream(void *stream) { m_stream = stream; }
  void *stream() const { return m_stream; }

  omega_model_t handle() const { return m_handle; }

  // Called by steppers after a step: drops the hook once the base
  // implementation has reported itself.
  void drop_noop_hook() const {
    if (m_base_hook_called && !m_hook_dropped) {
      check(omega_model_set_additional_tendency(m_handle, nullptr, nullptr));
      m_hook_dropped = true;
    }
  }

  // For calls that need a model without a hook (adjoint, ensembles): one
  // tendency evaluation of `state` into a scratch state lets the base
  // implementation report itself, so only a real override stays registered.
  void settle_hook(const ShallowWaterState &state) const {
    if (m_hook_dropped || m_hook_settled) return;
    ShallowWaterState scratch(m_mesh, state.ntracers());
    compute_tendency(scratch, state, 0.0);
    m_hook_settled = true;
  }

private:
  omega_model_t m_handle = nullptr;
  void *m_stream = nullptr;
  mutable bool m_base_hook_called = false;
  mutable bool m_hook_dropped = false;
  mutable bool m_hook_settled = false;

  void field_tendency(int field, const ShallowWaterState &tend, const ShallowWaterState &state,
                      AddMode add_mode) const {
    check(omega_model_compute_field_tendency(m_handle, field, tend.handle(), state.handle(),
                                             add_mode == AddMode::increment, m_stream));
  }

  static int hook(void *user, double *ht, double *vt, const double *h, const double *vn,
                  double t, omega_stream_t) {
    const auto *self = static_cast<const ShallowWaterModel *>(user);
    const Int nc = self->m_mesh->m_ncells, ne = self->m_mesh->m_nedges,
              L = self->m_mesh->m_nlayers;
    try {
      self->additional_tendency(DeviceArray2d{ht, nc, L}, DeviceArray2d{vt, ne, L},
                                ConstDeviceArray2d{h, nc, L}, ConstDeviceArray2d{vn, ne, L}, t);
    } catch (...) {
      return 1;
    }
    return 0;
  }
};

// dst += alpha * src for a device field handed to additional_tendency and a
// host field in the original numbering (kind 0 cells, 1 edges); src has one
// column (broadcast over layers) or nlayers columns.
inline void add_host_field(const MPASMesh &mesh, Int kind, DeviceArray2d dst, const Real *src,
                           Int ncols, Real alpha = 1.0, void *stream = nullptr) {
  check(omega_add_host_field(mesh.handle(), kind, dst.ptr, src, ncols, alpha, stream));
}

} // namespace omega
