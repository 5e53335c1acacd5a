// MPASMesh of the C++ drop-in API (reference: src/mesh/mpas_mesh.hpp:7-71,
// src/mesh/mpas_mesh.cpp:5-57).
//
// Fields keep the reference's names and shapes as host arrays in the
// original numbering.  finalize_mesh() hands the arrays to the library,
// which derives the signs, kite indices, level bounds and scalings exactly
// as the reference's finalize_mesh and builds the reordered device mesh;
// the derived fields are then copied back into the members.
#pragma once

#include "../common.hpp"

namespace omega {

struct MPASMesh {
  static constexpr Int maxedges = 6;

  Int m_nlayers = 0;
  Int m_nlayers_vec = 0;
  Int m_ncells = 0;
  Int m_nedges = 0;
  Int m_nvertices = 0;

  Int1d m_nedges_on_cell;
  Int2d m_edges_on_cell;
  Int2d m_cells_on_cell;
  Int2d m_vertices_on_cell;

  Int2d m_kite_index_on_cell;
  Int1d m_max_level_cell;

  Real1d m_area_cell;
  Real1d m_lat_cell;
  Real1d m_lon_cell;
  Real1d m_x_cell;
  Real1d m_y_cell;
  Real1d m_z_cell;
  Real1d m_mesh_density;
  Real2d m_edge_sign_on_cell;

  Int1d m_nedges_on_edge;
  Int2d m_edges_on_edge;
  Int2d m_cells_on_edge;
  Int2d m_vertices_on_edge;

  Int1d m_max_level_edge_bot;
  Int1d m_max_level_edge_top;

  Real1d m_dc_edge;
  Real1d m_dv_edge;
  Real1d m_angle_edge;
  Real1d m_lat_edge;
  Real1d m_lon_edge;
  Real1d m_x_edge;
  Real1d m_y_edge;
  Real1d m_z_edge;
  Real2d m_weights_on_edge;

  Real1d m_mesh_scaling_del2;
  Real1d m_mesh_scaling_del4;
  Real2d m_edge_mask;

  Int2d m_edges_on_vertex;
  Int2d m_cells_on_vertex;

  Int1d m_max_level_vertex_bot;
  Int1d m_max_level_vertex_top;

  Real1d m_area_triangle;
  Real1d m_lat_vertex;
  Real1d m_lon_vertex;
  Real1d m_x_vertex;
  Real1d m_y_vertex;
  Real1d m_z_vertex;
  Real2d m_kiteareas_on_vertex;
  Real2d m_edge_sign_on_vertex;

  // Device cell order used when the mesh is built (no reference
  // counterpart; OMEGA_ORDER_HILBERT by default).
  Int m_ordering = OMEGA_ORDER_HILBERT;

  MPASMesh() = default;
  MPASMesh(const MPASMesh &) = delete;
  MPASMesh &operator=(const MPASMesh &) = delete;
  virtual ~MPASMesh() {
    if (m_handle) omega_mesh_destroy(m_handle);
  }

  // Allocates the primary arrays (connectivity, metrics, geometry) for a
  // mesh read from elsewhere; the caller fills them and calls
  // finalize_mesh().
  void allocate(Int ncells, Int nedges, Int nvertices, Int nlayers) {
    m_ncells = ncells;
    m_nedges = nedges;
    m_nvertices = nvertices;
    m_nlayers = m_nlayers_vec = nlayers;
    m_nedges_on_cell = Int1d("nedges_on_cell", ncells);
    m_edges_on_cell = Int2d("edges_on_cell", ncells, maxedges);
    m_cells_on_cell = Int2d("cells_on_cell", ncells, maxedges);
    m_vertices_on_cell = Int2d("vertices_on_cell", ncells, maxedges);
    m_area_cell = Real1d("area_cell", ncells);
    m_lat_cell = Real1d("lat_cell", ncells);
    m_lon_cell = Real1d("lon_cell", ncells);
    m_x_cell = Real1d("x_cell", ncells);
    m_y_cell = Real1d("y_cell", ncells);
    m_z_cell = Real1d("z_cell", ncells);
    m_mesh_density = Real1d("mesh_density", ncells);
    m_nedges_on_edge = Int1d("nedges_on_edge", nedges);
    m_edges_on_edge = Int2d("edges_on_edge", nedges, 2 * maxedges);
    m_cells_on_edge = Int2d("cells_on_edge", nedges, 2);
    m_vertices_on_edge = Int2d("vertices_on_edge", nedges, 2);
    m_dc_edge = Real1d("dc_edge", nedges);
    m_dv_edge = Real1d("dv_edge", nedges);
    m_angle_edge = Real1d("angle_edge", nedges);
    m_lat_edge = Real1d("lat_edge", nedges);
    m_lon_edge = Real1d("lon_edge", nedges);
    m_x_edge = Real1d("x_edge", nedges);
    m_y_edge = Real1d("y_edge", nedges);
    m_z_edge = Real1d("z_edge", nedges);
    m_weights_on_edge = Real2d("weights_on_edge", nedges, 2 * maxedges);
    m_edges_on_vertex = Int2d("edges_on_vertex", nvertices, 3);
    m_cells_on_vertex = Int2d("cells_on_vertex", nvertices, 3);
    m_area_triangle = Real1d("area_triangle", nvertices);
    m_lat_vertex = Real1d("lat_vertex", nvertices);
    m_lon_vertex = Real1d("lon_vertex", nvertices);
    m_x_vertex = Real1d("x_vertex", nvertices);
    m_y_vertex = Real1d("y_vertex", nvertices);
    m_z_vertex = Real1d("z_vertex", nvertices);
    m_kiteareas_on_vertex = Real2d("kiteareas_on_vertex", nvertices, 3);
  }

  // MPASMesh::finalize_mesh (mpas_mesh.cpp:5-57): builds the device mesh
  // from the primary arrays and fills the derived ones.
  void finalize_mesh() {
    omega_mpas_mesh_desc d{};
    d.ncells = m_ncells;
    d.nedges = m_nedges;
    d.nvertices = m_nvertices;
    d.nlayers = m_nlayers;
    d.max_edges = maxedges;
    d.max_edges2 = 2 * maxedges;
    d.nedges_on_cell = m_nedges_on_cell.data();
    d.edges_on_cell = m_edges_on_cell.data();
    d.vertices_on_cell = m_vertices_on_cell.data();
    d.nedges_on_edge = m_nedges_on_edge.data();
    d.edges_on_edge = m_edges_on_edge.data();
    d.cells_on_edge = m_cells_on_edge.data();
    d.vertices_on_edge = m_vertices_on_edge.data();
    d.edges_on_vertex = m_edges_on_vertex.data();
    d.cells_on_vertex = m_cells_on_vertex.data();
    d.area_cell = m_area_cell.data();
    d.dc_edge = m_dc_edge.data();
    d.dv_edge = m_dv_edge.data();
    d.weights_on_edge = m_weights_on_edge.data();
    d.area_triangle = m_area_triangle.data();
    d.kiteareas_on_vertex = m_kiteareas_on_vertex.data();
    d.x_cell = m_x_cell.data();
    d.y_cell = m_y_cell.data();
    d.x_edge = m_x_edge.data();
    d.y_edge = m_y_edge.data();
    d.angle_edge = m_angle_edge.data();
    d.x_vertex = m_x_vertex.data();
    d.y_vertex = m_y_vertex.data();
    omega_mesh_t h = nullptr;
    check(omega_mesh_create_mpas(&d, m_ordering, &h));
    adopt(h, false);
  }

  // Library handle (device mesh), valid after construction/finalize_mesh.
  omega_mesh_t handle() const {
    if (!m_handle) throw std::runtime_error("omega: mesh used before finalize_mesh()");
    return m_handle;
  }

  // The arrays additional_tendency() receives are in the library's device
  // numbering, not in the numbering of the host fields above (m_x_cell...).
  // A HIP kernel launched from the hook indexes these device copies of the
  // geometry, which are in device numbering ("x_cell", "y_cell", "x_edge",
  // "y_edge", "angle_edge", "x_vertex", "y_vertex", "area_cell"; no
  // reference counterpart: the reference's views share one numbering).
  const Real *device_geometry(const char *name) const {
    const double *p = nullptr;
    int64_t n = 0;
    check(omega_mesh_device_array(handle(), name, &p, &n));
    return p;
  }

  // original[i] = index in the host fields of device entity i (kind 0 cells,
  // 1 edges, 2 vertices).
  std::vector<Int> device_to_original(Int kind) const {
    const Int n = kind == 0 ? m_ncells : kind == 1 ? m_nedges : m_nvertices;
    std::vector<int> p((size_t)n);
    check(omega_mesh_permutation(handle(), (int)kind, p.data()));
    return std::vector<Int>(p.begin(), p.end());
  }

protected:
  omega_mesh_t m_handle = nullptr;

  // Takes ownership of a library mesh and copies its host fields; with
  // all_fields the primary arrays are copied too (generated meshes).
  void adopt(omega_mesh_t h, bool all_fields) {
    if (m_handle) omega_mesh_destroy(m_handle);
    m_handle = h;
    int counts[6];
    double reals[3];
    check(omega_mesh_info(h, counts, reals));
    m_ncells = counts[0];
    m_nedges = counts[1];
    m_nvertices = counts[2];
    m_nlayers = m_nlayers_vec = counts[3];
    if (all_fields) {
      fetch(m_nedges_on_cell, "nedges_on_cell");
      fetch(m_edges_on_cell, "edges_on_cell");
      fetch(m_cells_on_cell, "cells_on_cell");
      fetch(m_vertices_on_cell, "vertices_on_cell");
      fetch(m_area_cell, "area_cell");
      fetch(m_lat_cell, "lat_cell");
      fetch(m_lon_cell, "lon_cell");
      fetch(m_x_cell, "x_cell");
      fetch(m_y_cell, "y_cell");
      fetch(m_z_cell, "z_cell");
      fetch(m_mesh_density, "mesh_density");
      fetch(m_nedges_on_edge, "nedges_on_edge");
      fetch(m_edges_on_edge, "edges_on_edge");
      fetch(m_cells_on_edge, "cells_on_edge");
      fetch(m_vertices_on_edge, "vertices_on_edge");
      fetch(m_dc_edge, "dc_edge");
      fetch(m_dv_edge, "dv_edge");
      fetch(m_angle_edge, "angle_edge");
      fetch(m_lat_edge, "lat_edge");
      fetch(m_lon_edge, "lon_edge");
      fetch(m_x_edge, "x_edge");
      fetch(m_y_edge, "y_edge");
      fetch(m_z_edge, "z_edge");
      fetch(m_weights_on_edge, "weights_on_edge");
      fetch(m_edges_on_vertex, "edges_on_vertex");
      fetch(m_cells_on_vertex, "cells_on_vertex");
      fetch(m_area_triangle, "area_triangle");
      fetch(m_lat_vertex, "lat_vertex");
      fetch(m_lon_vertex, "lon_vertex");
      fetch(m_x_vertex, "x_vertex");
      fetch(m_y_vertex, "y_vertex");
      fetch(m_z_vertex, "z_vertex");
      fetch(m_kiteareas_on_vertex, "kiteareas_on_vertex");
    }
    // fields finalize_mesh derives (mpas_mesh.cpp:7-56)
    fetch(m_kite_index_on_cell, "kite_index_on_cell");
    fetch(m_max_level_cell, "max_level_cell");
    fetch(m_edge_sign_on_cell, "edge_sign_on_cell");
    fetch(m_max_level_edge_bot, "max_level_edge_bot");
    fetch(m_max_level_edge_top, "max_level_edge_top");
    fetch(m_mesh_scaling_del2, "mesh_scaling_del2");
    fetch(m_mesh_scaling_del4, "mesh_scaling_del4");
    fetch(m_max_level_vertex_bot, "max_level_vertex_bot");
    fetch(m_max_level_vertex_top, "max_level_vertex_top");
    fetch(m_edge_sign_on_vertex, "edge_sign_on_vertex");
    m_edge_mask = Real2d("edge_mask", m_nedges, m_nlayers);
    for (Int e = 0; e < m_nedges; ++e)
      for (Int k = 0; k < m_nlayers; ++k)
        m_edge_mask(e, k) = k < m_max_level_edge_top(e) ? 1.0 : 0.0;
  }

private:
  template <class T, int R> void fetch(HostArray<T, R> &dst, const char *name) {
    const void *p = nullptr;
    int dtype = 0;
    int64_t n0 = 0, n1 = 0;
    check(omega_mesh_host_array(m_handle, name, &p, &dtype, &n0, &n1));
    if ((dtype == 0) != std::is_same<T, Int>::value)
      throw std::runtime_error(std::string("omega: type mismatch for mesh field ") + name);
    if constexpr (R == 1) dst = HostArray<T, 1>(name, n0 * n1);
    else dst = HostArray<T, 2>(name, n0, n1);
    const T *src = static_cast<const T *>(p);
    std::copy(src, src + n0 * n1, dst.data());
  }
};

} // namespace omega
