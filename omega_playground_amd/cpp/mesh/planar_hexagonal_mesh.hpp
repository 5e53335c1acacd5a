// PlanarHexagonalMesh of the C++ drop-in API (reference:
// src/mesh/planar_hexagonal_mesh.hpp:9-33, planar_hexagonal_mesh.cpp:5-398).
// The library generates the doubly periodic mesh in the reference's
// numbering; its fields are copied into the MPASMesh members.
#pragma once

#include "mpas_mesh.hpp"

namespace omega {

struct PlanarHexagonalMesh : MPASMesh {
  static constexpr Int maxedges = 6;

  Int m_nx;
  Int m_ny;
  Real m_dc;
  Real m_period_x;
  Real m_period_y;

  PlanarHexagonalMesh(Int nx, Int ny, Int nlayers = 1)
      : PlanarHexagonalMesh(nx, ny, 1. / nx, nlayers) {}

  PlanarHexagonalMesh(Int nx, Int ny, Real dc, Int nlayers = 1,
                      Int ordering = OMEGA_ORDER_HILBERT)
      : m_nx(nx), m_ny(ny), m_dc(dc) {
    m_ordering = ordering;
    omega_mesh_t h = nullptr;
    check(omega_mesh_create_planar_hex(nx, ny, dc, nlayers, ordering, &h));
    adopt(h, true);
    int counts[6];
    double reals[3];
    check(omega_mesh_info(h, counts, reals));
    m_period_x = reals[1];
    m_period_y = reals[2];
  }
};

} // namespace omega
