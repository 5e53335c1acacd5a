// CPU oracle for the shallow-water hot path -- TEST INFRASTRUCTURE ONLY.
//
// This is a plain C++17 (+ optional OpenMP) restatement of the arithmetic of
// mwarusz/omega_playground's default (non-SIMD, non-OMEGA_NO_INNER_IF)
// branches.  It exists to check the HIP implementation and to serve as the
// timed CPU baseline.  Only tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg may load it; nothing in the product path links it.
//
// Every function cites the reference file:line it follows.  Expression order
// (left-to-right products, accumulation order, replace/increment semantics)
// is kept exactly, so with -ffp-contract=off the results are bit-for-bit what
// the reference computes with vector_length == 1.
//
// Arrays are row-major with the vertical level fastest, as the reference's
// Kokkos::LayoutRight views (src/common.hpp:42,52-70):
//   h[ncells][nlayers], vn[nedges][nlayers], tr[ntracers][ncells][nlayers].
#pragma once

#include <cstdint>
#include <vector>

namespace orc {

using Real = double;
using Int = int;

// src/mesh/mpas_mesh.hpp:7-71
struct Mesh {
  static constexpr Int maxedges = 6;

  Int nx = 0, ny = 0;
  Real dc = 0, period_x = 0, period_y = 0;

  Int nlayers = 0;
  Int ncells = 0, nedges = 0, nvertices = 0;

  std::vector<Int> nedges_on_cell, edges_on_cell, cells_on_cell,
      vertices_on_cell, kite_index_on_cell, max_level_cell;
  std::vector<Real> area_cell, lat_cell, lon_cell, x_cell, y_cell, z_cell,
      mesh_density, edge_sign_on_cell;

  std::vector<Int> nedges_on_edge, edges_on_edge, cells_on_edge,
      vertices_on_edge, max_level_edge_bot, max_level_edge_top;
  std::vector<Real> dc_edge, dv_edge, angle_edge, lat_edge, lon_edge, x_edge,
      y_edge, z_edge, weights_on_edge, mesh_scaling_del2, mesh_scaling_del4,
      edge_mask;

  std::vector<Int> edges_on_vertex, cells_on_vertex, max_level_vertex_bot,
      max_level_vertex_top;
  std::vector<Real> area_triangle, lat_vertex, lon_vertex, x_vertex, y_vertex,
      z_vertex, kiteareas_on_vertex, edge_sign_on_vertex;

  void allocate();
  void finalize_mesh();
};

// src/mesh/planar_hexagonal_mesh.cpp:8-62
Mesh *make_planar_hex(Int nx, Int ny, Real dc, Int nlayers);

// src/model/shallow_water_params.hpp:7-18 (field order preserved)
struct Params {
  Real f0;
  Real grav;
  Real drag_coeff;
  Real visc_del2;
  Real visc_del4;
  Real eddy_diff2;
  Real eddy_diff4;
  Int ntracers;
  Int disable_h_tendency;
  Int disable_vn_tendency;
};

enum AddMode { replace = 0, increment = 1 };

// Signature of the additional-tendency hook
// (src/model/shallow_water_model.hpp:42-44).
typedef void (*AdditionalFn)(void *user, Real *h_tend, Real *vn_tend,
                             const Real *h, const Real *vn, Real t);

struct Model {
  const Mesh *mesh;
  Params params;
  bool en_pv = true, en_ke = true, en_ssh = true, en_vdel2 = false,
       en_vdel4 = false, en_tdel2 = false, en_tdel4 = false;
  // opt-in bottom drag (shallow_water_hierarchical.cpp:479-485, commented
  // out there; the active reference model never applies it)
  bool en_drag = false;

  // auxiliary state (src/model/shallow_water_auxstate.hpp:10-22)
  std::vector<Real> rvort_vertex, norm_rvort_vertex, norm_pvort_vertex;
  std::vector<Real> norm_rvort_edge, norm_pvort_edge;
  std::vector<Real> mean_h_edge, flux_h_edge;
  std::vector<Real> ke_cell, vel_div_cell;
  std::vector<Real> norm_tr_cell;
  // term-owned scratch
  std::vector<Real> vel_del2_edge, vel_del2_rvort_vertex, vel_del2_div_cell;
  std::vector<Real> tracer_del2_cell;

  std::vector<Real> f_vertex, f_edge;

  AdditionalFn additional = nullptr;
  void *additional_user = nullptr;

  // LinearShallowWaterModel (src/shallow_water_hierarchical.cpp:676-773):
  // linearised about a resting layer of thickness h0, no tracers
  bool linear = false;
  Real h0 = 0;

  Model(const Mesh *m, const Params &p);
  void aux_compute(const Real *h, const Real *vn, const Real *tr);
  void compute_h_tendency(Real *h_tend, const Real *h, const Real *vn,
                          AddMode mode);
  void compute_vn_tendency(Real *vn_tend, const Real *h, const Real *vn,
                           AddMode mode);
  void compute_tr_tendency(Real *tr_tend, const Real *tr, const Real *vn,
                           AddMode mode);
  void compute_tendency(Real *h_tend, Real *vn_tend, Real *tr_tend,
                        const Real *h, const Real *vn, const Real *tr, Real t,
                        AddMode mode);
  void linear_h_tendency(Real *h_tend, const Real *vn, AddMode mode) const;
  void linear_vn_tendency(Real *vn_tend, const Real *h, const Real *vn,
                          AddMode mode) const;
};

struct Stepper {
  Model *model;
  Int kind; // 0 = RK4, 1 = LSRK
  Int nstages;
  std::vector<Real> rka, rkb, rkc;
  std::vector<Real> h_tend, vn_tend, tr_tend;
  std::vector<Real> h_provis, vn_provis, tr_provis;
  std::vector<Real> h_old, vn_old, tr_old;
  Stepper(Model *m, Int kind, Int nstages);
  void do_step(Real t, Real dt, Real *h, Real *vn, Real *tr);
};

Real energy_integral(const Model &m, const Real *h, const Real *vn);
Real mass_integral(const Model &m, const Real *h);
Real circulation_integral(const Model &m, const Real *vn);

} // namespace orc
