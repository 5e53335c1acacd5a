// CPU oracle for the shallow-water hot path -- TEST INFRASTRUCTURE ONLY.
// See swe_oracle.hpp for scope.  Citations are to mwarusz/omega_playground.
//
// The reference runs each kernel as a Kokkos parallel_for over
// (entity, level-chunk) with vector_length lanes per chunk; with
// vector_length == 1 (OMEGA_NO_SIMD, src/common.hpp:19-21) every (entity,
// level) pair is computed independently.  Here the level loop is innermost and
// the neighbour loop (j) sits outside it, exactly as in the reference's
// non-SIMD branches, so each (entity, level) value sees the same sequence of
// roundings.
#include "swe_oracle.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <stdexcept>
#include <string>

#ifdef _OPENMP
#include <omp.h>
#define ORC_PARALLEL_FOR _Pragma("omp parallel for schedule(static)")
#else
#define ORC_PARALLEL_FOR
#endif

// Planted mistakes for the mutation check of the pinning tests
// (tests/test_oracle_mutants.py builds the oracle with -DORC_MUTANT=n):
//  1 PV flux drops the neighbour's planetary term    2 SSH gradient sign
//  3 del2 curl half divided by dc instead of dv       4 tracer edge value without 1/2
//  5 RK4 weights b0 and b1 swapped                    6 LSRK5 a2 sign flipped
//  7 kinetic energy without 1/4                       8 vertex circulation sign flipped
//  9 linear model: SSH gradient sign                 10 linear model: h0 dropped from div
// 11 linear model: energy weighs K with h, not h0
// 12 bottom drag applied on the top level instead of the deepest one
// 0 (the default) is the restatement itself.
#ifndef ORC_MUTANT
#define ORC_MUTANT 0
#endif

namespace orc {

static constexpr Real pi = M_PI; // src/common.hpp:27
static constexpr int mutant = ORC_MUTANT;

// ---------------------------------------------------------------------------
// Mesh
// ---------------------------------------------------------------------------

void Mesh::allocate() {
  const size_t nc = ncells, ne = nedges, nv = nvertices;
  nedges_on_cell.assign(nc, 0);
  cells_on_cell.assign(nc * maxedges, 0);
  edges_on_cell.assign(nc * maxedges, 0);
  vertices_on_cell.assign(nc * maxedges, 0);
  kite_index_on_cell.assign(nc * maxedges, 0);
  max_level_cell.assign(nc, 0);
  area_cell.assign(nc, 0);
  lat_cell.assign(nc, 0);
  lon_cell.assign(nc, 0);
  x_cell.assign(nc, 0);
  y_cell.assign(nc, 0);
  z_cell.assign(nc, 0);
  mesh_density.assign(nc, 0);
  edge_sign_on_cell.assign(nc * maxedges, 0);

  nedges_on_edge.assign(ne, 0);
  cells_on_edge.assign(ne * 2, 0);
  vertices_on_edge.assign(ne * 2, 0);
  edges_on_edge.assign(ne * 2 * maxedges, 0);
  max_level_edge_bot.assign(ne, 0);
  max_level_edge_top.assign(ne, 0);
  dc_edge.assign(ne, 0);
  dv_edge.assign(ne, 0);
  angle_edge.assign(ne, 0);
  lat_edge.assign(ne, 0);
  lon_edge.assign(ne, 0);
  x_edge.assign(ne, 0);
  y_edge.assign(ne, 0);
  z_edge.assign(ne, 0);
  weights_on_edge.assign(ne * 2 * maxedges, 0);
  mesh_scaling_del2.assign(ne, 0);
  mesh_scaling_del4.assign(ne, 0);
  edge_mask.assign(ne * (size_t)nlayers, 0);

  edges_on_vertex.assign(nv * 3, 0);
  cells_on_vertex.assign(nv * 3, 0);
  max_level_vertex_bot.assign(nv, 0);
  max_level_vertex_top.assign(nv, 0);
  area_triangle.assign(nv, 0);
  lat_vertex.assign(nv, 0);
  lon_vertex.assign(nv, 0);
  x_vertex.assign(nv, 0);
  y_vertex.assign(nv, 0);
  z_vertex.assign(nv, 0);
  kiteareas_on_vertex.assign(nv * 3, 0);
  edge_sign_on_vertex.assign(nv * 3, 0);
}

// src/mesh/mpas_mesh.cpp:5-57
void Mesh::finalize_mesh() {
  for (Int icell = 0; icell < ncells; ++icell) {
    for (Int j = 0; j < nedges_on_cell[icell]; ++j) {
      const Int e = edges_on_cell[icell * maxedges + j];
      edge_sign_on_cell[icell * maxedges + j] =
          cells_on_edge[2 * e + 0] == icell ? 1 : -1;
    }
    for (Int j = 0; j < nedges_on_cell[icell]; ++j) {
      const Int jvertex = vertices_on_cell[icell * maxedges + j];
      for (Int l = 0; l < 3; ++l) {
        if (cells_on_vertex[3 * jvertex + l] == icell) {
          kite_index_on_cell[icell * maxedges + j] = l;
        }
      }
    }
  }
  for (Int ivertex = 0; ivertex < nvertices; ++ivertex) {
    for (Int j = 0; j < 3; ++j) {
      const Int e = edges_on_vertex[3 * ivertex + j];
      edge_sign_on_vertex[3 * ivertex + j] =
          vertices_on_edge[2 * e + 0] == ivertex ? -1 : 1;
    }
  }
  std::fill(mesh_scaling_del2.begin(), mesh_scaling_del2.end(), 1.0);
  std::fill(mesh_scaling_del4.begin(), mesh_scaling_del4.end(), 1.0);
  std::fill(edge_mask.begin(), edge_mask.end(), 1.0);
  std::fill(max_level_cell.begin(), max_level_cell.end(), nlayers);
  std::fill(max_level_edge_bot.begin(), max_level_edge_bot.end(), nlayers);
  std::fill(max_level_edge_top.begin(), max_level_edge_top.end(), nlayers);
  std::fill(max_level_vertex_bot.begin(), max_level_vertex_bot.end(), nlayers);
  std::fill(max_level_vertex_top.begin(), max_level_vertex_top.end(), nlayers);
}

namespace {

// src/mesh/planar_hexagonal_mesh.cpp:64-166
struct HexIndex {
  Int nx, ny;
  Int cellidx(Int icol, Int irow) const { return irow * nx + icol; }
  Int cell_on_cell(Int icol, Int irow, Int nb) const {
    const Int mx = icol == 0 ? nx - 1 : icol - 1;
    const Int px = icol == (nx - 1) ? 0 : icol + 1;
    const Int my = irow == 0 ? ny - 1 : irow - 1;
    const Int py = irow == (ny - 1) ? 0 : irow + 1;
    if (irow % 2 == 0) {
      switch (nb) {
      case 0: return cellidx(mx, irow);
      case 1: return cellidx(mx, my);
      case 2: return cellidx(icol, my);
      case 3: return cellidx(px, irow);
      case 4: return cellidx(icol, py);
      case 5: return cellidx(mx, py);
      }
    } else {
      switch (nb) {
      case 0: return cellidx(mx, irow);
      case 1: return cellidx(icol, my);
      case 2: return cellidx(px, my);
      case 3: return cellidx(px, irow);
      case 4: return cellidx(px, py);
      case 5: return cellidx(icol, py);
      }
    }
    return -1;
  }
  Int edge_on_cell(Int icell, Int icol, Int irow, Int nb) const {
    switch (nb) {
    case 0: return 3 * icell;
    case 1: return 3 * icell + 1;
    case 2: return 3 * icell + 2;
    case 3: return 3 * cell_on_cell(icol, irow, nb);
    case 4: return 3 * cell_on_cell(icol, irow, nb) + 1;
    case 5: return 3 * cell_on_cell(icol, irow, nb) + 2;
    }
    return -1;
  }
  Int vertex_on_cell(Int icell, Int icol, Int irow, Int nb) const {
    switch (nb) {
    case 0: return 2 * icell;
    case 1: return 2 * icell + 1;
    case 2: return 2 * cell_on_cell(icol, irow, 2);
    case 3: return 2 * cell_on_cell(icol, irow, 3) + 1;
    case 4: return 2 * cell_on_cell(icol, irow, 3);
    case 5: return 2 * cell_on_cell(icol, irow, 4) + 1;
    }
    return -1;
  }
};

} // namespace

// src/mesh/planar_hexagonal_mesh.cpp:8-62 and compute_mesh_arrays (:168-398)
Mesh *make_planar_hex(Int nx, Int ny, Real dc, Int nlayers) {
  Mesh *m = new Mesh();
  m->nx = nx;
  m->ny = ny;
  m->dc = dc;
  // vector_length == 1: no layer padding (:11-12)
  m->nlayers = nlayers;
  m->ncells = nx * ny;
  m->nedges = 3 * m->ncells;
  m->nvertices = 2 * m->ncells;
  m->period_x = nx * dc;
  m->period_y = ny * dc * std::sqrt(3) / 2;
  m->allocate();

  const HexIndex hx{nx, ny};
  const Int ME = Mesh::maxedges;
  auto EoC = [&](Int c, Int j) -> Int & { return m->edges_on_cell[c * ME + j]; };
  auto VoC = [&](Int c, Int j) -> Int & { return m->vertices_on_cell[c * ME + j]; };
  auto EoE = [&](Int e, Int j) -> Int & { return m->edges_on_edge[e * 2 * ME + j]; };
  auto W = [&](Int e, Int j) -> Real & { return m->weights_on_edge[e * 2 * ME + j]; };

  // compute_mesh_arrays (:169-321); the scatter writes hit disjoint targets,
  // so a sequential loop gives the parallel_for result.
  for (Int irow = 0; irow < ny; ++irow) {
    for (Int icol = 0; icol < nx; ++icol) {
      const Int icell = hx.cellidx(icol, irow);
      m->nedges_on_cell[icell] = 6;
      for (Int j = 0; j < ME; ++j) {
        m->cells_on_cell[icell * ME + j] = hx.cell_on_cell(icol, irow, j);
        EoC(icell, j) = hx.edge_on_cell(icell, icol, irow, j);
        VoC(icell, j) = hx.vertex_on_cell(icell, icol, irow, j);
      }
      for (Int j = 0; j < 3; ++j) m->cells_on_edge[2 * EoC(icell, j) + 1] = icell;
      for (Int j = 3; j < ME; ++j) m->cells_on_edge[2 * EoC(icell, j) + 0] = icell;
      m->vertices_on_edge[2 * EoC(icell, 0) + 0] = VoC(icell, 1);
      m->vertices_on_edge[2 * EoC(icell, 0) + 1] = VoC(icell, 0);
      m->vertices_on_edge[2 * EoC(icell, 1) + 0] = VoC(icell, 2);
      m->vertices_on_edge[2 * EoC(icell, 1) + 1] = VoC(icell, 1);
      m->vertices_on_edge[2 * EoC(icell, 2) + 0] = VoC(icell, 3);
      m->vertices_on_edge[2 * EoC(icell, 2) + 1] = VoC(icell, 2);

      // edges on edge: slots 0-4 are filled from the cell that sees the edge
      // as its local edge 3/4/5, slots 5-9 from the cell that sees it as
      // local edge 0/1/2 (:200-264)
      static const int eoe_lo[3][5] = {{4, 5, 0, 1, 2}, {5, 0, 1, 2, 3}, {0, 1, 2, 3, 4}};
      static const int eoe_hi[3][5] = {{1, 2, 3, 4, 5}, {2, 3, 4, 5, 0}, {3, 4, 5, 0, 1}};
      for (Int s = 0; s < 3; ++s)
        for (Int j = 0; j < 5; ++j) EoE(EoC(icell, 3 + s), j) = EoC(icell, eoe_lo[s][j]);
      for (Int s = 0; s < 3; ++s)
        for (Int j = 0; j < 5; ++j) EoE(EoC(icell, s), 5 + j) = EoC(icell, eoe_hi[s][j]);

      // TRiSK weights before the 1/sqrt(3) scaling (:266-300)
      static const Real w_lo[3][5] = {{1. / 3, 1. / 6, 0, 1. / 6, 1. / 3},
                                      {1. / 3, -1. / 6, 0, 1. / 6, -1. / 3},
                                      {-1. / 3, -1. / 6, 0, -1. / 6, -1. / 3}};
      for (Int s = 0; s < 3; ++s)
        for (Int j = 0; j < 5; ++j) W(EoC(icell, 3 + s), j) = w_lo[s][j];
      for (Int s = 0; s < 3; ++s)
        for (Int j = 0; j < 5; ++j) W(EoC(icell, s), 5 + j) = w_lo[s][j];

      m->cells_on_vertex[3 * VoC(icell, 1) + 2] = icell;
      m->cells_on_vertex[3 * VoC(icell, 3) + 0] = icell;
      m->cells_on_vertex[3 * VoC(icell, 5) + 1] = icell;
      m->cells_on_vertex[3 * VoC(icell, 0) + 0] = icell;
      m->cells_on_vertex[3 * VoC(icell, 2) + 1] = icell;
      m->cells_on_vertex[3 * VoC(icell, 4) + 2] = icell;

      m->edges_on_vertex[3 * VoC(icell, 0) + 0] = EoC(icell, 0);
      m->edges_on_vertex[3 * VoC(icell, 1) + 0] = EoC(icell, 0);
      m->edges_on_vertex[3 * VoC(icell, 2) + 2] = EoC(icell, 1);
      m->edges_on_vertex[3 * VoC(icell, 1) + 2] = EoC(icell, 1);
      m->edges_on_vertex[3 * VoC(icell, 2) + 1] = EoC(icell, 2);
      m->edges_on_vertex[3 * VoC(icell, 3) + 1] = EoC(icell, 2);
    }
  }

  // scale_weights (:323-327)
  for (Int e = 0; e < m->nedges; ++e)
    for (Int j = 0; j < 2 * ME; ++j) W(e, j) *= 1. / std::sqrt(3);

  // compute_cell_arrays (:329-376)
  for (Int irow = 0; irow < ny; ++irow) {
    for (Int icol = 0; icol < nx; ++icol) {
      const Int icell = hx.cellidx(icol, irow);
      m->area_cell[icell] = dc * dc * std::sqrt(3) / 2;
      m->lat_cell[icell] = 0;
      m->lon_cell[icell] = 0;
      if (irow % 2 == 0) {
        m->x_cell[icell] = dc * icol + dc / 2;
        m->y_cell[icell] = dc * (irow + 1) * std::sqrt(3) / 2;
      } else {
        m->x_cell[icell] = dc * (icol + 1);
        m->y_cell[icell] = dc * (irow + 1) * std::sqrt(3) / 2;
      }
      m->z_cell[icell] = 0;
      const Real xc = m->x_cell[icell], yc = m->y_cell[icell];
      const Int e0 = EoC(icell, 0), e1 = EoC(icell, 1), e2 = EoC(icell, 2);
      m->x_edge[e0] = xc - dc / 2;
      m->y_edge[e0] = yc;
      m->z_edge[e0] = 0;
      m->x_edge[e1] = xc - dc / 2 * std::cos(pi / 3);
      m->y_edge[e1] = yc - dc / 2 * std::sin(pi / 3);
      m->z_edge[e1] = 0;
      m->x_edge[e2] = xc + dc / 2 * std::cos(pi / 3);
      m->y_edge[e2] = yc - dc / 2 * std::sin(pi / 3);
      m->z_edge[e2] = 0;
      m->angle_edge[e0] = 0;
      m->angle_edge[e1] = pi / 3;
      m->angle_edge[e2] = 2 * pi / 3;
      const Int v0 = VoC(icell, 0), v1 = VoC(icell, 1);
      m->x_vertex[v0] = xc - dc / 2;
      m->y_vertex[v0] = yc + dc * std::sqrt(3) / 6;
      m->z_vertex[v0] = 0;
      m->x_vertex[v1] = xc - dc / 2;
      m->y_vertex[v1] = yc - dc * std::sqrt(3) / 6;
      m->z_vertex[v1] = 0;
    }
  }

  // compute_edge_arrays (:378-385)
  for (Int e = 0; e < m->nedges; ++e) {
    m->nedges_on_edge[e] = 10;
    m->dc_edge[e] = dc;
    m->dv_edge[e] = m->dc_edge[e] * std::sqrt(3) / 3;
    m->lat_edge[e] = 0;
    m->lon_edge[e] = 0;
  }

  // compute_vertex_arrays (:387-395)
  for (Int v = 0; v < m->nvertices; ++v) {
    m->lat_vertex[v] = 0;
    m->lon_vertex[v] = 0;
    m->area_triangle[v] = dc * dc * std::sqrt(3) / 4;
    for (Int j = 0; j < 3; ++j) m->kiteareas_on_vertex[3 * v + j] = dc * dc * std::sqrt(3) / 12;
  }

  m->finalize_mesh();
  return m;
}

// ---------------------------------------------------------------------------
// Model
// ---------------------------------------------------------------------------

// src/model/shallow_water_model.cpp:5-41
Model::Model(const Mesh *m, const Params &p) : mesh(m), params(p) {
  const size_t L = m->nlayers;
  en_vdel2 = p.visc_del2 > 0;
  en_vdel4 = p.visc_del4 > 0;
  en_tdel2 = p.eddy_diff2 > 0;
  en_tdel4 = p.eddy_diff4 > 0;
  rvort_vertex.assign(m->nvertices * L, 0);
  norm_rvort_vertex.assign(m->nvertices * L, 0);
  norm_pvort_vertex.assign(m->nvertices * L, 0);
  norm_rvort_edge.assign(m->nedges * L, 0);
  norm_pvort_edge.assign(m->nedges * L, 0);
  mean_h_edge.assign(m->nedges * L, 0);
  flux_h_edge.assign(m->nedges * L, 0);
  ke_cell.assign(m->ncells * L, 0);
  vel_div_cell.assign(m->ncells * L, 0);
  norm_tr_cell.assign((size_t)p.ntracers * m->ncells * L, 0);
  vel_del2_edge.assign(m->nedges * L, 0);
  vel_del2_rvort_vertex.assign(m->nvertices * L, 0);
  vel_del2_div_cell.assign(m->ncells * L, 0);
  tracer_del2_cell.assign((size_t)p.ntracers * m->ncells * L, 0);
  f_vertex.assign(m->nvertices, p.f0);
  f_edge.assign(m->nedges, p.f0);
}

// src/model/shallow_water_auxstate.cpp:12-50 with the per-variable kernels
// in src/model/auxiliary_vars/*.hpp (#else branches).
void Model::aux_compute(const Real *h, const Real *vn, const Real *tr) {
  const Mesh &M = *mesh;
  const Int L = M.nlayers;
  const Int ME = Mesh::maxedges;

  // compute_vertex_auxiliarys -> VorticityAuxVars::compute_vort_vertex
  // (vorticity_aux.hpp:85-116)
  ORC_PARALLEL_FOR
  for (Int iv = 0; iv < M.nvertices; ++iv) {
    const Real inv_area_triangle = 1. / M.area_triangle[iv];
    for (Int k = 0; k < L; ++k) {
      Real rvort = 0, thick = 0;
      for (Int j = 0; j < 3; ++j) {
        const Int jedge = M.edges_on_vertex[3 * iv + j];
        const Int jcell = M.cells_on_vertex[3 * iv + j];
        rvort += M.dc_edge[jedge] * inv_area_triangle *
                 (mutant == 8 ? -1.0 : 1.0) * M.edge_sign_on_vertex[3 * iv + j] *
                 vn[(size_t)jedge * L + k];
        thick += M.kiteareas_on_vertex[3 * iv + j] * inv_area_triangle *
                 h[(size_t)jcell * L + k];
      }
      const Real inv_thick = 1. / thick;
      rvort_vertex[(size_t)iv * L + k] = rvort;
      norm_rvort_vertex[(size_t)iv * L + k] = inv_thick * rvort;
      norm_pvort_vertex[(size_t)iv * L + k] = inv_thick * f_vertex[iv];
    }
  }

  // compute_cell_auxiliarys -> KineticAuxVars::compute_kinetic_cell
  // (kinetic_aux.hpp:44-73) and TracerAuxVars::compute_norm_tracer_cell
  // (tracer_aux.hpp:31-51)
  const Int ntr = params.ntracers;
  ORC_PARALLEL_FOR
  for (Int ic = 0; ic < M.ncells; ++ic) {
    const Real inv_area_cell = 1. / M.area_cell[ic];
    for (Int k = 0; k < L; ++k) {
      Real ke = 0, div = 0;
      for (Int j = 0; j < M.nedges_on_cell[ic]; ++j) {
        const Int jedge = M.edges_on_cell[ic * ME + j];
        const Real area_edge = M.dv_edge[jedge] * M.dc_edge[jedge];
        const Real v = vn[(size_t)jedge * L + k];
        ke += inv_area_cell * area_edge * v * v * (mutant == 7 ? 1.0 : 0.25);
        div += inv_area_cell * M.dv_edge[jedge] * M.edge_sign_on_cell[ic * ME + j] * v;
      }
      vel_div_cell[(size_t)ic * L + k] = div;
      ke_cell[(size_t)ic * L + k] = ke;
    }
    for (Int l = 0; l < ntr; ++l) {
      for (Int k = 0; k < L; ++k) {
        const Real inv_h_cell = 1. / h[(size_t)ic * L + k];
        const size_t idx = ((size_t)l * M.ncells + ic) * L + k;
        norm_tr_cell[idx] = tr[idx] * inv_h_cell;
      }
    }
  }

  // compute_edge_auxiliarys -> ThicknessAuxVars::compute_thickness_edge
  // (thickness_aux.hpp:33-52) and VorticityAuxVars::compute_vort_edge
  // (vorticity_aux.hpp:118-132)
  ORC_PARALLEL_FOR
  for (Int ie = 0; ie < M.nedges; ++ie) {
    const Int c0 = M.cells_on_edge[2 * ie], c1 = M.cells_on_edge[2 * ie + 1];
    const Int v0 = M.vertices_on_edge[2 * ie], v1 = M.vertices_on_edge[2 * ie + 1];
    for (Int k = 0; k < L; ++k) {
      const Real mh = 0.5 * (h[(size_t)c0 * L + k] + h[(size_t)c1 * L + k]);
      mean_h_edge[(size_t)ie * L + k] = mh;
      flux_h_edge[(size_t)ie * L + k] = mh;
      norm_rvort_edge[(size_t)ie * L + k] =
          0.5 * (norm_rvort_vertex[(size_t)v0 * L + k] + norm_rvort_vertex[(size_t)v1 * L + k]);
      norm_pvort_edge[(size_t)ie * L + k] =
          0.5 * (norm_pvort_vertex[(size_t)v0 * L + k] + norm_pvort_vertex[(size_t)v1 * L + k]);
    }
  }
}

// src/model/shallow_water_model.cpp:43-86 (#else) with
// ThicknessHorzAdvOnCell (tendency_terms/thickness_hadv_cell.hpp:48-72)
void Model::compute_h_tendency(Real *h_tend, const Real *h, const Real *vn,
                               AddMode mode) {
  (void)h;
  const Mesh &M = *mesh;
  const Int L = M.nlayers, ME = Mesh::maxedges;
  ORC_PARALLEL_FOR
  for (Int ic = 0; ic < M.ncells; ++ic) {
    const Real inv_area_cell = 1. / M.area_cell[ic];
    for (Int k = 0; k < L; ++k) {
      Real &ht = h_tend[(size_t)ic * L + k];
      if (mode == replace) ht = 0;
      Real accum = 0;
      for (Int j = 0; j < M.nedges_on_cell[ic]; ++j) {
        const Int jedge = M.edges_on_cell[ic * ME + j];
        accum -= M.dv_edge[jedge] * inv_area_cell * M.edge_sign_on_cell[ic * ME + j] *
                 flux_h_edge[(size_t)jedge * L + k] * vn[(size_t)jedge * L + k];
      }
      ht += accum;
    }
  }
}

// src/model/shallow_water_model.cpp:88-219 (#else) with the edge terms in
// tendency_terms/{pv_flux_edge,ke_grad_edge,ssh_grad_edge,vel_diff_edge,
// vel_hyperdiff_edge}.hpp (#else branches)
void Model::compute_vn_tendency(Real *vn_tend, const Real *h, const Real *vn,
                                AddMode mode) {
  const Mesh &M = *mesh;
  const Int L = M.nlayers, ME = Mesh::maxedges;

  if (en_vdel4) {
    // compute_vel_del2 (vel_hyperdiff_edge.hpp:140-164)
    ORC_PARALLEL_FOR
    for (Int ie = 0; ie < M.nedges; ++ie) {
      const Int c0 = M.cells_on_edge[2 * ie], c1 = M.cells_on_edge[2 * ie + 1];
      const Int v0 = M.vertices_on_edge[2 * ie], v1 = M.vertices_on_edge[2 * ie + 1];
      const Real dc_edge_inv = 1. / M.dc_edge[ie];
      const Real dv_edge_inv = 1. / std::max(M.dv_edge[ie], 0.25 * M.dc_edge[ie]);
      for (Int k = 0; k < L; ++k) {
        const Real del2u =
            ((vel_div_cell[(size_t)c1 * L + k] - vel_div_cell[(size_t)c0 * L + k]) * dc_edge_inv -
             (rvort_vertex[(size_t)v1 * L + k] - rvort_vertex[(size_t)v0 * L + k]) * dv_edge_inv);
        vel_del2_edge[(size_t)ie * L + k] = del2u;
      }
    }
    // compute_vel_del2_rvort (vel_hyperdiff_edge.hpp:166-187)
    ORC_PARALLEL_FOR
    for (Int iv = 0; iv < M.nvertices; ++iv) {
      const Real inv_area_triangle = 1. / M.area_triangle[iv];
      for (Int k = 0; k < L; ++k) {
        Real rv = 0;
        for (Int j = 0; j < 3; ++j) {
          const Int jedge = M.edges_on_vertex[3 * iv + j];
          rv += M.dc_edge[jedge] * inv_area_triangle * M.edge_sign_on_vertex[3 * iv + j] *
                vel_del2_edge[(size_t)jedge * L + k];
        }
        vel_del2_rvort_vertex[(size_t)iv * L + k] = rv;
      }
    }
    // compute_vel_del2_div (vel_hyperdiff_edge.hpp:189-210)
    ORC_PARALLEL_FOR
    for (Int ic = 0; ic < M.ncells; ++ic) {
      const Real inv_area_cell = 1. / M.area_cell[ic];
      for (Int k = 0; k < L; ++k) {
        Real acc = 0;
        for (Int j = 0; j < M.nedges_on_cell[ic]; ++j) {
          const Int jedge = M.edges_on_cell[ic * ME + j];
          acc += M.dv_edge[jedge] * inv_area_cell * M.edge_sign_on_cell[ic * ME + j] *
                 vel_del2_edge[(size_t)jedge * L + k];
        }
        vel_del2_div_cell[(size_t)ic * L + k] = acc;
      }
    }
  }

  const Real grav = params.grav;
  const Real visc_del2 = params.visc_del2, visc_del4 = params.visc_del4;
  ORC_PARALLEL_FOR
  for (Int ie = 0; ie < M.nedges; ++ie) {
    const Int c0 = M.cells_on_edge[2 * ie], c1 = M.cells_on_edge[2 * ie + 1];
    const Int v0 = M.vertices_on_edge[2 * ie], v1 = M.vertices_on_edge[2 * ie + 1];
    const Real inv_dc_edge = 1. / M.dc_edge[ie];
    const Real inv_dv_edge = 1. / M.dv_edge[ie];
    for (Int k = 0; k < L; ++k) {
      Real &vt = vn_tend[(size_t)ie * L + k];
      if (mode == replace) vt = 0;

      if (en_pv) { // pv_flux_edge.hpp:56-84
        Real qt = 0;
        for (Int j = 0; j < M.nedges_on_edge[ie]; ++j) {
          const Int jedge = M.edges_on_edge[ie * 2 * ME + j];
          const Real norm_vort =
              (norm_rvort_edge[(size_t)ie * L + k] + norm_pvort_edge[(size_t)ie * L + k] +
               norm_rvort_edge[(size_t)jedge * L + k] +
               (mutant == 1 ? 0.0 : norm_pvort_edge[(size_t)jedge * L + k])) *
              0.5;
          qt += M.weights_on_edge[ie * 2 * ME + j] * flux_h_edge[(size_t)jedge * L + k] *
                vn[(size_t)jedge * L + k] * norm_vort;
        }
        vt += qt;
      }
      if (en_ke) { // ke_grad_edge.hpp:37-51
        vt -= (ke_cell[(size_t)c1 * L + k] - ke_cell[(size_t)c0 * L + k]) * inv_dc_edge;
      }
      if (en_ssh) { // ssh_grad_edge.hpp:37-50
        vt -= (mutant == 2 ? -grav : grav) * (h[(size_t)c1 * L + k] - h[(size_t)c0 * L + k]) *
              inv_dc_edge;
      }
      // opt-in bottom drag (term bit 128): the formula left as a comment at
      // src/shallow_water_hierarchical.cpp:479-485, on the edge's deepest
      // active level, with h_drag_edge = mean_h_edge
      if (en_drag && k == (mutant == 12 ? 0 : M.max_level_edge_top[ie] - 1)) {
        const Real inv_h_drag_edge = 1. / mean_h_edge[(size_t)ie * L + k];
        const Real drag_force =
            -params.drag_coeff *
            std::sqrt(ke_cell[(size_t)c0 * L + k] + ke_cell[(size_t)c1 * L + k]) *
            vn[(size_t)ie * L + k] * inv_h_drag_edge;
        vt += drag_force;
      }
      if (en_vdel2) { // vel_diff_edge.hpp:61-86
        const Real del2u =
            ((vel_div_cell[(size_t)c1 * L + k] - vel_div_cell[(size_t)c0 * L + k]) * inv_dc_edge -
             (rvort_vertex[(size_t)v1 * L + k] - rvort_vertex[(size_t)v0 * L + k]) *
                 (mutant == 3 ? inv_dc_edge : inv_dv_edge));
        vt += M.edge_mask[(size_t)ie * L + k] * visc_del2 * M.mesh_scaling_del2[ie] * del2u;
      }
      if (en_vdel4) { // vel_hyperdiff_edge.hpp:212-237
        const Real del2u = ((vel_del2_div_cell[(size_t)c1 * L + k] -
                             vel_del2_div_cell[(size_t)c0 * L + k]) *
                                inv_dc_edge -
                            (vel_del2_rvort_vertex[(size_t)v1 * L + k] -
                             vel_del2_rvort_vertex[(size_t)v0 * L + k]) *
                                inv_dv_edge);
        vt -= M.edge_mask[(size_t)ie * L + k] * visc_del4 * M.mesh_scaling_del4[ie] * del2u;
      }
    }
  }
}

// src/model/shallow_water_model.cpp:221-311 (#else) with
// tendency_terms/{tracer_hadv_cell,tracer_diff_cell,tracer_hyperdiff_cell}.hpp
void Model::compute_tr_tendency(Real *tr_tend, const Real *tr, const Real *vn,
                                AddMode mode) {
  (void)tr;
  const Mesh &M = *mesh;
  const Int L = M.nlayers, ME = Mesh::maxedges, ntr = params.ntracers;
  const Int NC = M.ncells;
  // the reference passes m_mean_h_edge as flux_h and m_flux_h_edge as mean_h
  // (shallow_water_model.cpp:225-228); both hold the same values.
  const std::vector<Real> &flux_h = mean_h_edge;
  const std::vector<Real> &mean_h = flux_h_edge;

  if (en_tdel4) { // compute_tracer_del2 (tracer_hyperdiff_cell.hpp:94-129)
    ORC_PARALLEL_FOR
    for (Int lc = 0; lc < ntr * NC; ++lc) {
      const Int l = lc / NC, ic = lc % NC;
      const Real inv_area_cell = 1. / M.area_cell[ic];
      for (Int k = 0; k < L; ++k) {
        Real d2 = 0;
        for (Int j = 0; j < M.nedges_on_cell[ic]; ++j) {
          const Int jedge = M.edges_on_cell[ic * ME + j];
          const Real inv_dc_edge = 1. / M.dc_edge[jedge];
          const Int jc0 = M.cells_on_edge[2 * jedge], jc1 = M.cells_on_edge[2 * jedge + 1];
          const Real grad = (norm_tr_cell[((size_t)l * NC + jc1) * L + k] -
                             norm_tr_cell[((size_t)l * NC + jc0) * L + k]) *
                            inv_dc_edge;
          d2 += M.dv_edge[jedge] * inv_area_cell * M.edge_sign_on_cell[ic * ME + j] *
                mean_h[(size_t)jedge * L + k] * grad;
        }
        tracer_del2_cell[((size_t)l * NC + ic) * L + k] = d2;
      }
    }
  }

  const Real eddy_diff2 = params.eddy_diff2, eddy_diff4 = params.eddy_diff4;
  ORC_PARALLEL_FOR
  for (Int lc = 0; lc < ntr * NC; ++lc) {
    const Int l = lc / NC, ic = lc % NC;
    const Real inv_area_cell = 1. / M.area_cell[ic];
    for (Int k = 0; k < L; ++k) {
      Real &tt = tr_tend[((size_t)l * NC + ic) * L + k];
      if (mode == replace) tt = 0;
      { // TracerHorzAdvOnCell (tracer_hadv_cell.hpp:60-90)
        Real accum = 0;
        for (Int j = 0; j < M.nedges_on_cell[ic]; ++j) {
          const Int jedge = M.edges_on_cell[ic * ME + j];
          const Int jc0 = M.cells_on_edge[2 * jedge], jc1 = M.cells_on_edge[2 * jedge + 1];
          const Real norm_tr_edge = (norm_tr_cell[((size_t)l * NC + jc0) * L + k] +
                                     norm_tr_cell[((size_t)l * NC + jc1) * L + k]) *
                                    (mutant == 4 ? 1.0 : 0.5);
          accum -= M.dv_edge[jedge] * inv_area_cell * M.edge_sign_on_cell[ic * ME + j] *
                   flux_h[(size_t)jedge * L + k] * norm_tr_edge * vn[(size_t)jedge * L + k];
        }
        tt += accum;
      }
      if (en_tdel2) { // TracerDiffusionOnCell (tracer_diff_cell.hpp:61-92)
        Real accum = 0;
        for (Int j = 0; j < M.nedges_on_cell[ic]; ++j) {
          const Int jedge = M.edges_on_cell[ic * ME + j];
          const Real inv_dc_edge = 1. / M.dc_edge[jedge];
          const Int jc0 = M.cells_on_edge[2 * jedge], jc1 = M.cells_on_edge[2 * jedge + 1];
          const Real grad = (norm_tr_cell[((size_t)l * NC + jc1) * L + k] -
                             norm_tr_cell[((size_t)l * NC + jc0) * L + k]) *
                            inv_dc_edge;
          accum += eddy_diff2 * inv_area_cell * M.dv_edge[jedge] *
                   M.edge_sign_on_cell[ic * ME + j] * mean_h[(size_t)jedge * L + k] *
                   M.mesh_scaling_del2[jedge] * grad;
        }
        tt += accum;
      }
      if (en_tdel4) { // TracerHyperDiffusionOnCell (tracer_hyperdiff_cell.hpp:131-161)
        Real accum = 0;
        for (Int j = 0; j < M.nedges_on_cell[ic]; ++j) {
          const Int jedge = M.edges_on_cell[ic * ME + j];
          const Real inv_dc_edge = 1. / M.dc_edge[jedge];
          const Int jc0 = M.cells_on_edge[2 * jedge], jc1 = M.cells_on_edge[2 * jedge + 1];
          const Real grad = (tracer_del2_cell[((size_t)l * NC + jc1) * L + k] -
                             tracer_del2_cell[((size_t)l * NC + jc0) * L + k]) *
                            inv_dc_edge;
          accum -= eddy_diff4 * inv_area_cell * M.dv_edge[jedge] *
                   M.edge_sign_on_cell[ic * ME + j] * M.mesh_scaling_del4[jedge] * grad;
        }
        tt += accum;
      }
    }
  }
}

// LinearShallowWaterModel::compute_h_tendency
// (src/shallow_water_hierarchical.cpp:680-707): -h0 div(vn)
void Model::linear_h_tendency(Real *h_tend, const Real *vn, AddMode mode) const {
  const Mesh &M = *mesh;
  const Int L = M.nlayers, ME = Mesh::maxedges;
  ORC_PARALLEL_FOR
  for (Int ic = 0; ic < M.ncells; ++ic) {
    for (Int k = 0; k < L; ++k) {
      Real accum = 0;
      for (Int j = 0; j < M.nedges_on_cell[ic]; ++j) {
        const Int jedge = M.edges_on_cell[ic * ME + j];
        accum += M.dv_edge[jedge] * M.edge_sign_on_cell[ic * ME + j] * vn[(size_t)jedge * L + k];
      }
      const Real x = (mutant == 10 ? -1.0 : -h0) * accum / M.area_cell[ic];
      Real &t = h_tend[(size_t)ic * L + k];
      if (mode == increment) t += x;
      else t = x;
    }
  }
}

// LinearShallowWaterModel::compute_vn_tendency
// (src/shallow_water_hierarchical.cpp:709-741): f_edge * (TRiSK tangential
// velocity) - g grad(h)
void Model::linear_vn_tendency(Real *vn_tend, const Real *h, const Real *vn,
                               AddMode mode) const {
  const Mesh &M = *mesh;
  const Int L = M.nlayers, ME2 = 2 * Mesh::maxedges;
  const Real grav = params.grav;
  ORC_PARALLEL_FOR
  for (Int ie = 0; ie < M.nedges; ++ie) {
    for (Int k = 0; k < L; ++k) {
      Real vt = 0;
      for (Int j = 0; j < M.nedges_on_edge[ie]; ++j) {
        const Int jedge = M.edges_on_edge[ie * ME2 + j];
        vt += M.weights_on_edge[ie * ME2 + j] * vn[(size_t)jedge * L + k];
      }
      const Int ic0 = M.cells_on_edge[2 * ie], ic1 = M.cells_on_edge[2 * ie + 1];
      const Real grad_h = (h[(size_t)ic1 * L + k] - h[(size_t)ic0 * L + k]) / M.dc_edge[ie];
      const Real x = mutant == 9 ? f_edge[ie] * vt + grav * grad_h : f_edge[ie] * vt - grav * grad_h;
      Real &t = vn_tend[(size_t)ie * L + k];
      if (mode == increment) t += x;
      else t = x;
    }
  }
}

// src/model/shallow_water_model.cpp:313-347; the linear model keeps the
// base class's sequence with no auxiliary variables and no tracers
// (src/shallow_water_hierarchical.cpp:20-56)
void Model::compute_tendency(Real *h_tend, Real *vn_tend, Real *tr_tend,
                             const Real *h, const Real *vn, const Real *tr,
                             Real t, AddMode mode) {
  if (linear) {
    if (!params.disable_h_tendency) linear_h_tendency(h_tend, vn, mode);
    if (!params.disable_vn_tendency) linear_vn_tendency(vn_tend, h, vn, mode);
    if (additional) additional(additional_user, h_tend, vn_tend, h, vn, t);
    return;
  }
  aux_compute(h, vn, tr);
  if (!params.disable_h_tendency) compute_h_tendency(h_tend, h, vn, mode);
  if (!params.disable_vn_tendency) compute_vn_tendency(vn_tend, h, vn, mode);
  if (params.ntracers > 0) compute_tr_tendency(tr_tend, tr, vn, mode);
  if (additional) additional(additional_user, h_tend, vn_tend, h, vn, t);
}

// ---------------------------------------------------------------------------
// Time steppers
// ---------------------------------------------------------------------------

// src/time_integration/rk4.cpp:8-25 and lsrk.cpp:5-34
Stepper::Stepper(Model *m, Int kind_, Int nstages_) : model(m), kind(kind_) {
  const Mesh &M = *m->mesh;
  const size_t L = M.nlayers, ntr = m->params.ntracers;
  h_tend.assign(M.ncells * L, 0);
  vn_tend.assign(M.nedges * L, 0);
  tr_tend.assign(ntr * M.ncells * L, 0);
  if (kind == 0) {
    nstages = 4;
    rka = {1. / 2, 1. / 2, 1};
    rkb = {1. / 6, 1. / 3, 1. / 3, 1. / 6};
    if (mutant == 5) std::swap(rkb[0], rkb[1]);
    rkc = {1. / 2, 1. / 2, 1};
    h_provis.assign(M.ncells * L, 0);
    vn_provis.assign(M.nedges * L, 0);
    tr_provis.assign(ntr * M.ncells * L, 0);
    h_old.assign(M.ncells * L, 0);
    vn_old.assign(M.nedges * L, 0);
    tr_old.assign(ntr * M.ncells * L, 0);
  } else {
    nstages = nstages_;
    rka.assign(nstages, 0.);
    rkb.assign(nstages, 0.);
    rkc.assign(nstages, 0.);
    if (nstages == 5) {
      rka = {0., -567301805773. / 1357537059087., -2404267990393. / 2016746695238.,
             -3550918686646. / 2091501179385., -1275806237668. / 842570457699.};
      rkb = {1432997174477. / 9575080441755., 5161836677717. / 13612068292357.,
             1720146321549. / 2090206949498., 3134564353537. / 4481467310338.,
             2277821191437. / 14882151754819.};
      rkc = {0., 1432997174477. / 9575080441755., 2526269341429. / 6820363962896.,
             2006345519317. / 3224310063776., 2802321613138. / 2924317926251.};
      if (mutant == 6) rka[2] = -rka[2];
    }
    if (nstages == 1) {
      rka = {0};
      rkb = {1};
      rkc = {0};
    }
  }
}

static void axpy_inplace(Real *y, Real alpha, const Real *x, size_t n) {
  ORC_PARALLEL_FOR
  for (long i = 0; i < (long)n; ++i) y[i] += alpha * x[i];
}

void Stepper::do_step(Real t, Real dt, Real *h, Real *vn, Real *tr) {
  const Mesh &M = *model->mesh;
  const size_t L = M.nlayers, ntr = model->params.ntracers;
  const size_t nh = M.ncells * L, nv = M.nedges * L, nt = ntr * M.ncells * L;
  if (kind == 0) { // rk4.cpp:28-104
    std::copy(h, h + nh, h_old.begin());
    std::copy(vn, vn + nv, vn_old.begin());
    std::copy(tr, tr + nt, tr_old.begin());
    model->compute_tendency(h_tend.data(), vn_tend.data(), tr_tend.data(), h, vn, tr, t, replace);
    for (Int stage = 0; stage < 4; ++stage) {
      const Real rkb_stage = rkb[stage];
      // state += dt * rkb * tend; the product is evaluated as (dt*rkb)*tend
      const Real db = dt * rkb_stage;
      axpy_inplace(h, db, h_tend.data(), nh);
      axpy_inplace(vn, db, vn_tend.data(), nv);
      if (ntr > 0) axpy_inplace(tr, db, tr_tend.data(), nt);
      if (stage < 3) {
        const Real stagetime = t + rkc[stage] * dt;
        const Real da = dt * rka[stage];
        ORC_PARALLEL_FOR
        for (long i = 0; i < (long)nh; ++i) h_provis[i] = h_old[i] + da * h_tend[i];
        ORC_PARALLEL_FOR
        for (long i = 0; i < (long)nv; ++i) vn_provis[i] = vn_old[i] + da * vn_tend[i];
        if (ntr > 0) {
          ORC_PARALLEL_FOR
          for (long i = 0; i < (long)nt; ++i) tr_provis[i] = tr_old[i] + da * tr_tend[i];
        }
        model->compute_tendency(h_tend.data(), vn_tend.data(), tr_tend.data(), h_provis.data(),
                                vn_provis.data(), tr_provis.data(), stagetime, replace);
      }
    }
  } else { // lsrk.cpp:36-89
    for (Int stage = 0; stage < nstages; ++stage) {
      const Real rka_stage = rka[stage];
      ORC_PARALLEL_FOR
      for (long i = 0; i < (long)nh; ++i) h_tend[i] *= rka_stage;
      ORC_PARALLEL_FOR
      for (long i = 0; i < (long)nv; ++i) vn_tend[i] *= rka_stage;
      if (ntr > 0) {
        ORC_PARALLEL_FOR
        for (long i = 0; i < (long)nt; ++i) tr_tend[i] *= rka_stage;
      }
      const Real stagetime = t + rkc[stage] * dt;
      model->compute_tendency(h_tend.data(), vn_tend.data(), tr_tend.data(), h, vn, tr,
                              stagetime, increment);
      const Real db = dt * rkb[stage];
      axpy_inplace(h, db, h_tend.data(), nh);
      axpy_inplace(vn, db, vn_tend.data(), nv);
      if (ntr > 0) axpy_inplace(tr, db, tr_tend.data(), nt);
    }
  }
}

// ---------------------------------------------------------------------------
// Diagnostics (src/model/shallow_water_diagnostics.cpp:5-89); summed
// sequentially in entity order.
// ---------------------------------------------------------------------------

Real energy_integral(const Model &m, const Real *h, const Real *vn) {
  const Mesh &M = *m.mesh;
  const Int L = M.nlayers, ME = Mesh::maxedges;
  const Real grav = m.params.grav;
  Real total = 0;
  for (Int ic = 0; ic < M.ncells; ++ic) {
    for (Int k = 0; k < M.max_level_cell[ic]; ++k) {
      Real K = 0;
      for (Int j = 0; j < M.nedges_on_cell[ic]; ++j) {
        const Int jedge = M.edges_on_cell[ic * ME + j];
        const Real area_edge = M.dv_edge[jedge] * M.dc_edge[jedge];
        const Real v = vn[(size_t)jedge * L + k];
        K += area_edge * v * v / 4;
      }
      K /= M.area_cell[ic];
      const Real hc = h[(size_t)ic * L + k];
      // the linear model weighs K with h0 (shallow_water_hierarchical.cpp:767-769)
      total += M.area_cell[ic] * (grav * hc * hc / 2 + (m.linear && mutant != 11 ? m.h0 : hc) * K);
    }
  }
  return total;
}

Real mass_integral(const Model &m, const Real *h) {
  const Mesh &M = *m.mesh;
  const Int L = M.nlayers;
  Real total = 0;
  for (Int ic = 0; ic < M.ncells; ++ic)
    for (Int k = 0; k < M.max_level_cell[ic]; ++k) total += M.area_cell[ic] * h[(size_t)ic * L + k];
  return total;
}

Real circulation_integral(const Model &m, const Real *vn) {
  const Mesh &M = *m.mesh;
  const Int L = M.nlayers;
  Real total = 0;
  for (Int iv = 0; iv < M.nvertices; ++iv) {
    for (Int k = 0; k < M.max_level_vertex_bot[iv]; ++k) {
      Real cir_i = -0.;
      for (Int j = 0; j < 3; ++j) {
        const Int jedge = M.edges_on_vertex[3 * iv + j];
        cir_i += M.dc_edge[jedge] * M.edge_sign_on_vertex[3 * iv + j] * vn[(size_t)jedge * L + k];
      }
      total += cir_i + m.f_vertex[iv] * M.area_triangle[iv];
    }
  }
  return total;
}

} // namespace orc

// ---------------------------------------------------------------------------
// C ABI used by the Python test harness (ctypes).
// ---------------------------------------------------------------------------
using namespace orc;

namespace {
struct ArrayRef {
  void *ptr;
  int dtype; // 0 = int32, 1 = float64
  long n0, n1;
};

template <class T> ArrayRef ref(std::vector<T> &v, long n0, long n1) {
  return ArrayRef{v.data(), std::is_same<T, double>::value ? 1 : 0, n0, n1};
}

bool mesh_array(Mesh *m, const std::string &name, ArrayRef &out) {
  const long NC = m->ncells, NE = m->nedges, NV = m->nvertices, L = m->nlayers;
  const long ME = Mesh::maxedges;
#define ORC_MA(field, a, b)                                                    \
  if (name == #field) { out = ref(m->field, a, b); return true; }
  ORC_MA(nedges_on_cell, NC, 1) ORC_MA(edges_on_cell, NC, ME) ORC_MA(cells_on_cell, NC, ME)
  ORC_MA(vertices_on_cell, NC, ME) ORC_MA(kite_index_on_cell, NC, ME) ORC_MA(max_level_cell, NC, 1)
  ORC_MA(area_cell, NC, 1) ORC_MA(lat_cell, NC, 1) ORC_MA(lon_cell, NC, 1) ORC_MA(x_cell, NC, 1)
  ORC_MA(y_cell, NC, 1) ORC_MA(z_cell, NC, 1) ORC_MA(mesh_density, NC, 1)
  ORC_MA(edge_sign_on_cell, NC, ME)
  ORC_MA(nedges_on_edge, NE, 1) ORC_MA(edges_on_edge, NE, 2 * ME) ORC_MA(cells_on_edge, NE, 2)
  ORC_MA(vertices_on_edge, NE, 2) ORC_MA(max_level_edge_bot, NE, 1) ORC_MA(max_level_edge_top, NE, 1)
  ORC_MA(dc_edge, NE, 1) ORC_MA(dv_edge, NE, 1) ORC_MA(angle_edge, NE, 1) ORC_MA(lat_edge, NE, 1)
  ORC_MA(lon_edge, NE, 1) ORC_MA(x_edge, NE, 1) ORC_MA(y_edge, NE, 1) ORC_MA(z_edge, NE, 1)
  ORC_MA(weights_on_edge, NE, 2 * ME) ORC_MA(mesh_scaling_del2, NE, 1)
  ORC_MA(mesh_scaling_del4, NE, 1) ORC_MA(edge_mask, NE, L)
  ORC_MA(edges_on_vertex, NV, 3) ORC_MA(cells_on_vertex, NV, 3) ORC_MA(max_level_vertex_bot, NV, 1)
  ORC_MA(max_level_vertex_top, NV, 1) ORC_MA(area_triangle, NV, 1) ORC_MA(lat_vertex, NV, 1)
  ORC_MA(lon_vertex, NV, 1) ORC_MA(x_vertex, NV, 1) ORC_MA(y_vertex, NV, 1) ORC_MA(z_vertex, NV, 1)
  ORC_MA(kiteareas_on_vertex, NV, 3) ORC_MA(edge_sign_on_vertex, NV, 3)
#undef ORC_MA
  return false;
}

bool model_array(Model *m, const std::string &name, ArrayRef &out) {
  const long NC = m->mesh->ncells, NE = m->mesh->nedges, NV = m->mesh->nvertices;
  const long L = m->mesh->nlayers, NT = m->params.ntracers;
#define ORC_AA(field, a, b)                                                    \
  if (name == #field) { out = ref(m->field, a, b); return true; }
  ORC_AA(rvort_vertex, NV, L) ORC_AA(norm_rvort_vertex, NV, L) ORC_AA(norm_pvort_vertex, NV, L)
  ORC_AA(norm_rvort_edge, NE, L) ORC_AA(norm_pvort_edge, NE, L) ORC_AA(mean_h_edge, NE, L)
  ORC_AA(flux_h_edge, NE, L) ORC_AA(ke_cell, NC, L) ORC_AA(vel_div_cell, NC, L)
  ORC_AA(norm_tr_cell, NT * NC, L) ORC_AA(vel_del2_edge, NE, L)
  ORC_AA(vel_del2_rvort_vertex, NV, L) ORC_AA(vel_del2_div_cell, NC, L)
  ORC_AA(tracer_del2_cell, NT * NC, L) ORC_AA(f_vertex, NV, 1) ORC_AA(f_edge, NE, 1)
#undef ORC_AA
  return false;
}
} // namespace

extern "C" {

void *orc_mesh_planar_hex(int nx, int ny, double dc, int nlayers) {
  return make_planar_hex(nx, ny, dc, nlayers);
}
void orc_mesh_free(void *m) { delete static_cast<Mesh *>(m); }
void orc_mesh_info(void *mp, int *ints, double *reals) {
  Mesh *m = static_cast<Mesh *>(mp);
  ints[0] = m->nx; ints[1] = m->ny; ints[2] = m->nlayers;
  ints[3] = m->ncells; ints[4] = m->nedges; ints[5] = m->nvertices;
  reals[0] = m->dc; reals[1] = m->period_x; reals[2] = m->period_y;
}
void *orc_mesh_array(void *mp, const char *name, int *dtype, long *n0, long *n1) {
  ArrayRef r{};
  if (!mesh_array(static_cast<Mesh *>(mp), name, r)) return nullptr;
  *dtype = r.dtype; *n0 = r.n0; *n1 = r.n1;
  return r.ptr;
}
// Re-derives the mesh's sign/kite arrays after a caller edited connectivity
// (used to build general MPAS meshes in tests).
void orc_mesh_finalize(void *mp) { static_cast<Mesh *>(mp)->finalize_mesh(); }
// An empty MPASMesh of the given sizes (mpas_mesh.hpp:7-71): the caller fills
// the primary arrays through orc_mesh_array and calls orc_mesh_finalize.
void *orc_mesh_alloc(int ncells, int nedges, int nvertices, int nlayers) {
  Mesh *m = new Mesh;
  m->ncells = ncells;
  m->nedges = nedges;
  m->nvertices = nvertices;
  m->nlayers = nlayers;
  m->allocate();
  return m;
}

void *orc_model_create(void *mesh, const Params *p) {
  return new Model(static_cast<Mesh *>(mesh), *p);
}
// Turns a model into LinearShallowWaterModel with mean thickness h0
// (shallow_water_hierarchical.cpp:676-678).
void orc_model_set_linear(void *mp, double h0) {
  Model *m = static_cast<Model *>(mp);
  m->linear = true;
  m->h0 = h0;
}
void orc_model_free(void *m) { delete static_cast<Model *>(m); }
void *orc_model_array(void *mp, const char *name, int *dtype, long *n0, long *n1) {
  ArrayRef r{};
  if (!model_array(static_cast<Model *>(mp), name, r)) return nullptr;
  *dtype = r.dtype; *n0 = r.n0; *n1 = r.n1;
  return r.ptr;
}
void orc_model_set_additional(void *mp, AdditionalFn fn, void *user) {
  Model *m = static_cast<Model *>(mp);
  m->additional = fn;
  m->additional_user = user;
}
// Term switches for isolating one operator in tests: bit 0 pv, 1 ke, 2 ssh,
// 3 vel del2, 4 vel del4, 5 tracer del2, 6 tracer del4.
void orc_model_set_terms(void *mp, int mask) {
  Model *m = static_cast<Model *>(mp);
  m->en_pv = mask & 1; m->en_ke = mask & 2; m->en_ssh = mask & 4; m->en_vdel2 = mask & 8;
  m->en_vdel4 = mask & 16; m->en_tdel2 = mask & 32; m->en_tdel4 = mask & 64;
  m->en_drag = mask & 128;
}
int orc_model_get_terms(void *mp) {
  Model *m = static_cast<Model *>(mp);
  return (m->en_pv ? 1 : 0) | (m->en_ke ? 2 : 0) | (m->en_ssh ? 4 : 0) | (m->en_vdel2 ? 8 : 0) |
         (m->en_vdel4 ? 16 : 0) | (m->en_tdel2 ? 32 : 0) | (m->en_tdel4 ? 64 : 0) |
         (m->en_drag ? 128 : 0);
}
void orc_aux_compute(void *mp, const double *h, const double *vn, const double *tr) {
  static_cast<Model *>(mp)->aux_compute(h, vn, tr);
}
// one field of compute_tendency after orc_aux_compute: field 0
// compute_h_tendency(tend, a=h, vn), 1 compute_vn_tendency(tend, a=h, vn),
// 2 compute_tr_tendency(tend, a=tr, vn)
void orc_compute_field_tendency(void *mp, int field, double *tend, const double *a,
                                const double *vn, int add_mode) {
  Model *m = static_cast<Model *>(mp);
  const AddMode mode = add_mode ? increment : replace;
  if (field == 0) m->compute_h_tendency(tend, a, vn, mode);
  else if (field == 1) m->compute_vn_tendency(tend, a, vn, mode);
  else m->compute_tr_tendency(tend, a, vn, mode);
}
void orc_compute_tendency(void *mp, double *h_tend, double *vn_tend, double *tr_tend,
                          const double *h, const double *vn, const double *tr, double t,
                          int add_mode) {
  static_cast<Model *>(mp)->compute_tendency(h_tend, vn_tend, tr_tend, h, vn, tr, t,
                                             add_mode ? increment : replace);
}
void *orc_stepper_create(void *mp, int kind, int nstages) {
  return new Stepper(static_cast<Model *>(mp), kind, nstages);
}
void orc_stepper_free(void *s) { delete static_cast<Stepper *>(s); }
void orc_stepper_coeffs(void *sp, double *a, double *b, double *c) {
  Stepper *s = static_cast<Stepper *>(sp);
  std::copy(s->rka.begin(), s->rka.end(), a);
  std::copy(s->rkb.begin(), s->rkb.end(), b);
  std::copy(s->rkc.begin(), s->rkc.end(), c);
}
int orc_stepper_nstages(void *sp) { return static_cast<Stepper *>(sp)->nstages; }
void orc_do_steps(void *sp, double t0, double dt, int nsteps, int t_from_step_plus_one,
                  double *h, double *vn, double *tr) {
  Stepper *s = static_cast<Stepper *>(sp);
  for (int step = 0; step < nsteps; ++step) {
    const double t = (step + (t_from_step_plus_one ? 1 : 0)) * dt + t0;
    s->do_step(t, dt, h, vn, tr);
  }
}
double orc_energy(void *mp, const double *h, const double *vn) {
  return energy_integral(*static_cast<Model *>(mp), h, vn);
}
double orc_mass(void *mp, const double *h) { return mass_integral(*static_cast<Model *>(mp), h); }
double orc_circulation(void *mp, const double *vn) {
  return circulation_integral(*static_cast<Model *>(mp), vn);
}
void orc_set_num_threads(int n) {
#ifdef _OPENMP
  omp_set_num_threads(n);
#else
  (void)n;
#endif
}
int orc_get_max_threads() {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}
}
