// Host-side mesh construction, space-filling-curve ordering and upload.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>

#include "internal.hpp"

namespace omega_hip {

void HostMesh::resize() {
  const size_t nc = ncells, ne = nedges, nv = nvertices, me = max_edges, me2 = max_edges2;
  nedges_on_cell.assign(nc, 0);
  edges_on_cell.assign(nc * me, 0);
  cells_on_cell.assign(nc * me, 0);
  vertices_on_cell.assign(nc * me, 0);
  kite_index_on_cell.assign(nc * me, 0);
  max_level_cell.assign(nc, 0);
  for (auto *v : {&area_cell, &lat_cell, &lon_cell, &x_cell, &y_cell, &z_cell, &mesh_density})
    v->assign(nc, 0.0);
  edge_sign_on_cell.assign(nc * me, 0.0);
  nedges_on_edge.assign(ne, 0);
  edges_on_edge.assign(ne * me2, 0);
  cells_on_edge.assign(ne * 2, 0);
  vertices_on_edge.assign(ne * 2, 0);
  max_level_edge_bot.assign(ne, 0);
  max_level_edge_top.assign(ne, 0);
  for (auto *v : {&dc_edge, &dv_edge, &angle_edge, &lat_edge, &lon_edge, &x_edge, &y_edge,
                  &z_edge, &mesh_scaling_del2, &mesh_scaling_del4})
    v->assign(ne, 0.0);
  weights_on_edge.assign(ne * me2, 0.0);
  edges_on_vertex.assign(nv * 3, 0);
  cells_on_vertex.assign(nv * 3, 0);
  max_level_vertex_bot.assign(nv, 0);
  max_level_vertex_top.assign(nv, 0);
  for (auto *v : {&area_triangle, &lat_vertex, &lon_vertex, &x_vertex, &y_vertex, &z_vertex})
    v->assign(nv, 0.0);
  kiteareas_on_vertex.assign(nv * 3, 0.0);
  edge_sign_on_vertex.assign(nv * 3, 0.0);
}

// Derived connectivity of MPASMesh::finalize_mesh (src/mesh/mpas_mesh.cpp:5-57):
// the orientation of each cell edge (+1 when the cell is the edge's first
// cell), the kite slot of each cell vertex, the orientation of each vertex
// edge (-1 when the vertex is the edge's first vertex), unit mesh scalings
// and full-depth level limits.
void HostMesh::derive_signs_and_levels() {
  const int me = max_edges;
  for (int c = 0; c < ncells; ++c) {
    for (int j = 0; j < nedges_on_cell[c]; ++j) {
      const int e = edges_on_cell[(size_t)c * me + j];
      edge_sign_on_cell[(size_t)c * me + j] = cells_on_edge[2 * (size_t)e] == c ? 1.0 : -1.0;
      const int v = vertices_on_cell[(size_t)c * me + j];
      for (int l = 0; l < 3; ++l)
        if (cells_on_vertex[3 * (size_t)v + l] == c) kite_index_on_cell[(size_t)c * me + j] = l;
    }
  }
  for (int v = 0; v < nvertices; ++v)
    for (int j = 0; j < 3; ++j) {
      const int e = edges_on_vertex[3 * (size_t)v + j];
      edge_sign_on_vertex[3 * (size_t)v + j] = vertices_on_edge[2 * (size_t)e] == v ? -1.0 : 1.0;
    }
  std::fill(max_level_cell.begin(), max_level_cell.end(), nlayers);
  std::fill(max_level_edge_bot.begin(), max_level_edge_bot.end(), nlayers);
  std::fill(max_level_vertex_bot.begin(), max_level_vertex_bot.end(), nlayers);
  std::fill(max_level_vertex_top.begin(), max_level_vertex_top.end(), nlayers);
}

// ---------------------------------------------------------------------------
// Doubly periodic planar hexagonal mesh in the reference's numbering
// (src/mesh/planar_hexagonal_mesh.cpp:8-398).
//
// Cell (col, row) has index row*nx + col; rows alternate between two
// horizontal offsets.  Local directions d = 0..5 are W, SW, SE, E, NE, NW.
// Cell c owns the edges toward W, SW, SE (ids 3c, 3c+1, 3c+2) and its
// upper-left and lower-left corners (ids 2c, 2c+1); every other edge or
// corner of c is owned by the neighbour across it.
// ---------------------------------------------------------------------------
namespace {

struct Lattice {
  int nx, ny;
  // column/row offsets of the six neighbours, even rows then odd rows
  static constexpr int dcol[2][6] = {{-1, -1, 0, 1, 0, -1}, {-1, 0, 1, 1, 1, 0}};
  static constexpr int drow[6] = {0, -1, -1, 0, 1, 1};
  int neighbour(int c, int d) const {
    const int col = c % nx, row = c / nx;
    const int nc_ = (col + dcol[row & 1][d] + nx) % nx;
    const int nr_ = (row + drow[d] + ny) % ny;
    return nr_ * nx + nc_;
  }
};
constexpr int Lattice::dcol[2][6];
constexpr int Lattice::drow[6];

// Edge in direction d of cell c: owned by c for d < 3, otherwise edge d-3 of
// the neighbour across it.  Corner k of cell c (k = 0 upper-left, 1
// lower-left, 2 bottom, 3 lower-right, 4 upper-right, 5 top), as an
// (owner direction, owner corner) pair: -1 means c itself.
constexpr int corner_owner[6][2] = {{-1, 0}, {-1, 1}, {2, 0}, {3, 1}, {3, 0}, {4, 1}};

// Edges-on-edge rows: the five edges of the cell that sees edge e as its
// local edge s come in local-slot order starting after s (slots 5..9 when
// s < 3, slots 0..4 when s >= 3), with the TRiSK weights (before the common
// 1/sqrt(3) factor) of the regular hexagon:
constexpr double kTriskW[3][5] = {{1. / 3, 1. / 6, 0, 1. / 6, 1. / 3},
                                  {1. / 3, -1. / 6, 0, 1. / 6, -1. / 3},
                                  {-1. / 3, -1. / 6, 0, -1. / 6, -1. / 3}};

// (corner, slot) pairs that place c in cells_on_vertex of its corners and
// (corner, slot, local edge) triples for edges_on_vertex of the corners the
// owned edges touch.
constexpr int cov_write[6][2] = {{1, 2}, {3, 0}, {5, 1}, {0, 0}, {2, 1}, {4, 2}};
constexpr int eov_write[6][3] = {{0, 0, 0}, {1, 0, 0}, {2, 2, 1}, {1, 2, 1}, {2, 1, 2}, {3, 1, 2}};

} // namespace

void build_planar_hex(HostMesh &m, int nx, int ny, double dc, int nlayers) {
  OMEGA_REQUIRE(nx >= 3 && ny >= 3, "planar hex mesh needs nx, ny >= 3");
  OMEGA_REQUIRE(nlayers >= 1, "nlayers must be >= 1");
  OMEGA_REQUIRE(dc > 0, "dc must be positive");
  m.nx = nx;
  m.ny = ny;
  m.dc = dc;
  m.nlayers = nlayers;
  m.ncells = nx * ny;
  m.nedges = 3 * m.ncells;
  m.nvertices = 2 * m.ncells;
  m.max_edges = 6;
  m.max_edges2 = 12;
  m.period_x = nx * dc;
  m.period_y = ny * dc * std::sqrt(3) / 2;
  m.resize();
  const Lattice lat{nx, ny};
  const int NC = m.ncells;

  auto edge_of = [&](int c, int d) {
    return d < 3 ? 3 * c + d : 3 * lat.neighbour(c, d) + (d - 3);
  };
  auto corner_of = [&](int c, int k) {
    const int od = corner_owner[k][0];
    return 2 * (od < 0 ? c : lat.neighbour(c, od)) + corner_owner[k][1];
  };

  for (int c = 0; c < NC; ++c) {
    m.nedges_on_cell[c] = 6;
    for (int d = 0; d < 6; ++d) {
      m.cells_on_cell[6 * c + d] = lat.neighbour(c, d);
      m.edges_on_cell[6 * c + d] = edge_of(c, d);
      m.vertices_on_cell[6 * c + d] = corner_of(c, d);
    }
  }
  for (int c = 0; c < NC; ++c) {
    const int *E = &m.edges_on_cell[6 * c];
    const int *V = &m.vertices_on_cell[6 * c];
    for (int d = 0; d < 3; ++d) {
      // owned edge d joins the neighbour (first cell) and c (second cell);
      // its vertices are corner d+1 then corner d
      m.cells_on_edge[2 * E[d] + 1] = c;
      m.cells_on_edge[2 * E[d + 3] + 0] = c;
      m.vertices_on_edge[2 * E[d] + 0] = V[d + 1];
      m.vertices_on_edge[2 * E[d] + 1] = V[d];
    }
    for (int s = 0; s < 6; ++s) {
      const int e = E[s];
      const int base = s < 3 ? 5 : 0;
      for (int j = 0; j < 5; ++j) {
        m.edges_on_edge[12 * e + base + j] = E[(s + 1 + j) % 6];
        m.weights_on_edge[12 * e + base + j] = kTriskW[s % 3][j];
      }
    }
    for (const auto &w : cov_write) m.cells_on_vertex[3 * V[w[0]] + w[1]] = c;
    for (const auto &w : eov_write) m.edges_on_vertex[3 * V[w[0]] + w[1]] = E[w[2]];
  }
  for (int e = 0; e < m.nedges; ++e) {
    m.nedges_on_edge[e] = 10;
    for (int j = 0; j < 12; ++j) m.weights_on_edge[12 * e + j] *= 1. / std::sqrt(3);
  }

  // Geometry, evaluated with the reference's operation order so that the
  // metric terms are bit-identical (planar_hexagonal_mesh.cpp:329-395).
  const double pi = M_PI;
  for (int c = 0; c < NC; ++c) {
    const int col = c % nx, row = c / nx;
    m.area_cell[c] = dc * dc * std::sqrt(3) / 2;
    m.x_cell[c] = (row % 2 == 0) ? dc * col + dc / 2 : dc * (col + 1);
    m.y_cell[c] = dc * (row + 1) * std::sqrt(3) / 2;
    const double xc = m.x_cell[c], yc = m.y_cell[c];
    const int *E = &m.edges_on_cell[6 * c];
    const int *V = &m.vertices_on_cell[6 * c];
    m.x_edge[E[0]] = xc - dc / 2;
    m.y_edge[E[0]] = yc;
    m.x_edge[E[1]] = xc - dc / 2 * std::cos(pi / 3);
    m.y_edge[E[1]] = yc - dc / 2 * std::sin(pi / 3);
    m.x_edge[E[2]] = xc + dc / 2 * std::cos(pi / 3);
    m.y_edge[E[2]] = yc - dc / 2 * std::sin(pi / 3);
    m.angle_edge[E[0]] = 0;
    m.angle_edge[E[1]] = pi / 3;
    m.angle_edge[E[2]] = 2 * pi / 3;
    m.x_vertex[V[0]] = xc - dc / 2;
    m.y_vertex[V[0]] = yc + dc * std::sqrt(3) / 6;
    m.x_vertex[V[1]] = xc - dc / 2;
    m.y_vertex[V[1]] = yc - dc * std::sqrt(3) / 6;
  }
  for (int e = 0; e < m.nedges; ++e) {
    m.dc_edge[e] = dc;
    m.dv_edge[e] = m.dc_edge[e] * std::sqrt(3) / 3;
  }
  for (int v = 0; v < m.nvertices; ++v) {
    m.area_triangle[v] = dc * dc * std::sqrt(3) / 4;
    for (int j = 0; j < 3; ++j) m.kiteareas_on_vertex[3 * v + j] = dc * dc * std::sqrt(3) / 12;
  }
  std::fill(m.mesh_scaling_del2.begin(), m.mesh_scaling_del2.end(), 1.0);
  std::fill(m.mesh_scaling_del4.begin(), m.mesh_scaling_del4.end(), 1.0);
  std::fill(m.max_level_edge_top.begin(), m.max_level_edge_top.end(), nlayers);
  m.derive_signs_and_levels();
}

// ---------------------------------------------------------------------------
// Space-filling-curve orderings
// ---------------------------------------------------------------------------
namespace {

// Index of (x, y) along the Hilbert curve filling a 2^order square
// (the iterative rotate-and-reflect construction).
uint64_t hilbert_index(uint32_t order, uint32_t x, uint32_t y) {
  uint64_t d = 0;
  for (uint32_t s = 1u << (order - 1); s > 0; s >>= 1) {
    const uint32_t rx = (x & s) ? 1 : 0;
    const uint32_t ry = (y & s) ? 1 : 0;
    d += (uint64_t)s * s * ((3 * rx) ^ ry);
    if (ry == 0) {
      if (rx == 1) {
        x = s - 1 - x;
        y = s - 1 - y;
      }
      std::swap(x, y);
    }
  }
  return d;
}

uint64_t morton_index(uint32_t x, uint32_t y) {
  uint64_t d = 0;
  for (int b = 0; b < 32; ++b)
    d |= ((uint64_t)((x >> b) & 1) << (2 * b)) | ((uint64_t)((y >> b) & 1) << (2 * b + 1));
  return d;
}

} // namespace

std::vector<int> order_cells(const HostMesh &m, int ordering) {
  std::vector<int> n2o(m.ncells);
  std::iota(n2o.begin(), n2o.end(), 0);
  if (ordering == OMEGA_ORDER_ORIGINAL) return n2o;
  // Lattice coordinates: (col, row) for planar meshes, else cell centres
  // quantised onto a square grid of about one point per cell.
  std::vector<uint32_t> gx(m.ncells), gy(m.ncells);
  uint32_t extent = 1;
  if (m.nx > 0) {
    for (int c = 0; c < m.ncells; ++c) {
      gx[c] = c % m.nx;
      gy[c] = c / m.nx;
    }
    extent = std::max(m.nx, m.ny);
  } else {
    if (m.x_cell.empty() || m.y_cell.empty()) return n2o;
    const auto [xmin, xmax] = std::minmax_element(m.x_cell.begin(), m.x_cell.end());
    const auto [ymin, ymax] = std::minmax_element(m.y_cell.begin(), m.y_cell.end());
    const double span = std::max(*xmax - *xmin, *ymax - *ymin);
    extent = std::max<uint32_t>(1, (uint32_t)std::ceil(std::sqrt((double)m.ncells)));
    const double scale = span > 0 ? (extent - 1) / span : 0.0;
    for (int c = 0; c < m.ncells; ++c) {
      gx[c] = (uint32_t)std::lround((m.x_cell[c] - *xmin) * scale);
      gy[c] = (uint32_t)std::lround((m.y_cell[c] - *ymin) * scale);
    }
  }
  uint32_t order = 1;
  while ((1u << order) < extent) ++order;
  std::vector<uint64_t> key(m.ncells);
  for (int c = 0; c < m.ncells; ++c)
    key[c] = ordering == OMEGA_ORDER_MORTON ? morton_index(gx[c], gy[c])
                                            : hilbert_index(order, gx[c], gy[c]);
  std::stable_sort(n2o.begin(), n2o.end(), [&](int a, int b) { return key[a] < key[b]; });
  return n2o;
}

// ---------------------------------------------------------------------------
// Device layout
//
// Cells follow the curve.  Each edge and vertex is owned by the cell that
// lists it in its lowest local slot (ties: lower original cell index); edges
// and vertices are numbered by (owner position on the curve, slot), so a
// block that takes a stretch of cells also finds their edges and vertices in
// contiguous ranges [eofs[c0], eofs[c1]) and [vofs[c0], vofs[c1]).  With the
// original ordering this reproduces the reference numbering of the planar
// mesh.  Neighbour lists keep their local order, so every sum runs over the
// same terms in the same order as in the reference.
// ---------------------------------------------------------------------------

// x[i*stride + j] == x[j] bit for bit for every row i
template <class T> bool rows_equal(const std::vector<T> &x, size_t stride) {
  if (stride == 0 || x.size() < stride || x.size() % stride) return false;
  for (size_t i = stride; i < x.size(); i += stride)
    if (std::memcmp(&x[i], &x[0], stride * sizeof(T)) != 0) return false;
  return true;
}
template bool rows_equal(const std::vector<double> &, size_t);
template bool rows_equal(const std::vector<int> &, size_t);

void DeviceMesh::build_device(int ord) {
  HostMesh &m = host;
  ordering = ord;
  const int NC = m.ncells, NE = m.nedges, NV = m.nvertices, ME = m.max_edges, L = m.nlayers;
  cell_n2o = order_cells(m, ord);
  cell_o2n.assign(NC, -1);
  for (int i = 0; i < NC; ++i) cell_o2n[cell_n2o[i]] = i;

  // owners
  std::vector<int> eown(NE, -1), eslot(NE, 99), vown(NV, -1), vslot(NV, 99);
  for (int c = 0; c < NC; ++c) {
    for (int j = 0; j < m.nedges_on_cell[c]; ++j) {
      const int e = m.edges_on_cell[(size_t)c * ME + j];
      const int v = m.vertices_on_cell[(size_t)c * ME + j];
      OMEGA_REQUIRE(e >= 0 && e < NE && v >= 0 && v < NV, "mesh connectivity out of range");
      if (j < eslot[e] || (j == eslot[e] && c < eown[e])) { eslot[e] = j; eown[e] = c; }
      if (j < vslot[v] || (j == vslot[v] && c < vown[v])) { vslot[v] = j; vown[v] = c; }
    }
  }
  for (int e = 0; e < NE; ++e) OMEGA_REQUIRE(eown[e] >= 0, "edge not referenced by any cell");
  for (int v = 0; v < NV; ++v) OMEGA_REQUIRE(vown[v] >= 0, "vertex not referenced by any cell");

  auto order_by_owner = [&](int n, const std::vector<int> &own, const std::vector<int> &slot,
                            std::vector<int> &n2o, std::vector<int> &o2n, std::vector<int> &ofs) {
    n2o.resize(n);
    std::iota(n2o.begin(), n2o.end(), 0);
    std::stable_sort(n2o.begin(), n2o.end(), [&](int a, int b) {
      const int ra = cell_o2n[own[a]], rb = cell_o2n[own[b]];
      return ra != rb ? ra < rb : slot[a] < slot[b];
    });
    o2n.assign(n, -1);
    for (int i = 0; i < n; ++i) o2n[n2o[i]] = i;
    ofs.assign(NC + 1, 0);
    for (int i = 0; i < n; ++i) ofs[cell_o2n[own[n2o[i]]] + 1]++;
    for (int c = 0; c < NC; ++c) ofs[c + 1] += ofs[c];
  };
  order_by_owner(NE, eown, eslot, edge_n2o, edge_o2n, eofs_h);
  order_by_owner(NV, vown, vslot, vert_n2o, vert_o2n, vofs_h);

  int noe_max = 0;
  for (int e = 0; e < NE; ++e) noe_max = std::max(noe_max, m.nedges_on_edge[e]);
  me2 = std::max(1, noe_max);
  const int ME2 = m.max_edges2;

  noc_h.assign(NC, 0);
  eoc_h.assign((size_t)NC * ME, 0);
  cdiv_h.assign((size_t)NC * ME, 0.0);
  cke_h.assign((size_t)NC * ME, 0.0);
  for (int i = 0; i < NC; ++i) {
    const int c = cell_n2o[i];
    const double inv_area = 1. / m.area_cell[c];
    noc_h[i] = m.nedges_on_cell[c];
    for (int j = 0; j < noc_h[i]; ++j) {
      const int e = m.edges_on_cell[(size_t)c * ME + j];
      eoc_h[(size_t)i * ME + j] = edge_o2n[e];
      // (dv * inv_area) * sign: the leading factors of every cell-edge
      // product in the reference (e.g. thickness_hadv_cell.hpp:62-64)
      cdiv_h[(size_t)i * ME + j] = m.dv_edge[e] * inv_area * m.edge_sign_on_cell[(size_t)c * ME + j];
      // inv_area * (dv*dc): leading factors of the kinetic energy sum
      // (kinetic_aux.hpp:61-67)
      const double area_edge = m.dv_edge[e] * m.dc_edge[e];
      cke_h[(size_t)i * ME + j] = inv_area * area_edge;
    }
  }

  coe_h.assign((size_t)NE * 2, 0);
  voe_h.assign((size_t)NE * 2, 0);
  noe_h.assign(NE, 0);
  eoe_h.assign((size_t)NE * me2, 0);
  w_h.assign((size_t)NE * me2, 0.0);
  inv_dc_h.assign(NE, 0.0);
  inv_dv_h.assign(NE, 0.0);
  inv_dv4_h.assign(NE, 0.0);
  s2_h.assign(NE, 1.0);
  s4_h.assign(NE, 1.0);
  maxlev_h.assign(NE, L);
  for (int i = 0; i < NE; ++i) {
    const int e = edge_n2o[i];
    for (int s = 0; s < 2; ++s) {
      coe_h[2 * (size_t)i + s] = cell_o2n[m.cells_on_edge[2 * (size_t)e + s]];
      voe_h[2 * (size_t)i + s] = vert_o2n[m.vertices_on_edge[2 * (size_t)e + s]];
    }
    noe_h[i] = m.nedges_on_edge[e];
    for (int j = 0; j < noe_h[i]; ++j) {
      eoe_h[(size_t)i * me2 + j] = edge_o2n[m.edges_on_edge[(size_t)e * ME2 + j]];
      w_h[(size_t)i * me2 + j] = m.weights_on_edge[(size_t)e * ME2 + j];
    }
    inv_dc_h[i] = 1. / m.dc_edge[e];
    inv_dv_h[i] = 1. / m.dv_edge[e];
    inv_dv4_h[i] = 1. / std::max(m.dv_edge[e], 0.25 * m.dc_edge[e]);
    s2_h[i] = m.mesh_scaling_del2[e];
    s4_h[i] = m.mesh_scaling_del4[e];
    maxlev_h[i] = std::min(L, m.max_level_edge_top[e]);
  }

  eov_h.assign((size_t)NV * 3, 0);
  cov_h.assign((size_t)NV * 3, 0);
  crv_h.assign((size_t)NV * 3, 0.0);
  cth_h.assign((size_t)NV * 3, 0.0);
  for (int i = 0; i < NV; ++i) {
    const int v = vert_n2o[i];
    const double inv_at = 1. / m.area_triangle[v];
    for (int j = 0; j < 3; ++j) {
      const int e = m.edges_on_vertex[3 * (size_t)v + j];
      eov_h[3 * (size_t)i + j] = edge_o2n[e];
      cov_h[3 * (size_t)i + j] = cell_o2n[m.cells_on_vertex[3 * (size_t)v + j]];
      // (dc * inv_area_triangle) * sign and kite * inv_area_triangle
      // (vorticity_aux.hpp:94-106)
      crv_h[3 * (size_t)i + j] = m.dc_edge[e] * inv_at * m.edge_sign_on_vertex[3 * (size_t)v + j];
      cth_h[3 * (size_t)i + j] = m.kiteareas_on_vertex[3 * (size_t)v + j] * inv_at;
    }
  }

  build_hex_records(*this);
  // tables the pair kernels may read from one cell for both (bit-identical
  // at every stencil slot: e = 3c+K, v = 2c+x on the hexagonal numbering)
  coef_uni = hex_ok && rows_equal(inv_dc_h, 1) && rows_equal(inv_dv_h, 1) &&
             rows_equal(inv_dv4_h, 1) &&
             rows_equal(maxlev_h, 1) && rows_equal(s2_h, 1) && rows_equal(s4_h, 1) &&
             rows_equal(w_h, 3 * (size_t)me2) && rows_equal(cdiv_h, ME) &&
             rows_equal(cke_h, ME) && rows_equal(crv_h, 6) && rows_equal(cth_h, 6);
  const size_t rows = std::max({(size_t)NC, (size_t)NE, (size_t)NV});
  off32 = rows < (size_t(1) << 24) && rows * (size_t)L * sizeof(double) < (size_t(1) << 32);
  uploaded = false;
}

// Copies the device-numbered tables to the GPU (first use by a model or
// state); mesh construction itself touches no device memory.
void DeviceMesh::upload() {
  if (uploaded) return;
  const HostMesh &m = host;
  const int NC = m.ncells, NV = m.nvertices, ME = m.max_edges;
  noc.upload(noc_h); eoc.upload(eoc_h); cdiv.upload(cdiv_h); cke.upload(cke_h);
  coe.upload(coe_h); voe.upload(voe_h); noe.upload(noe_h); eoe.upload(eoe_h); w.upload(w_h);
  inv_dc.upload(inv_dc_h); inv_dv.upload(inv_dv_h); inv_dv4.upload(inv_dv4_h);
  s2.upload(s2_h); s4.upload(s4_h); maxlev_e.upload(maxlev_h);
  eov.upload(eov_h); cov.upload(cov_h); crv.upload(crv_h); cth.upload(cth_h);
  eofs.upload(eofs_h); vofs.upload(vofs_h);
  if (hex_ok) {
    hexrec.upload(hexrec_h);
    if (hexu_ok) hexu.upload(hexu_h);
    coc.upload(coc_h);
  }

  auto perm_upload = [](DBuf<Real> &dst, const std::vector<double> &src,
                        const std::vector<int> &n2o) {
    if (src.empty()) return;
    std::vector<double> tmp(n2o.size());
    for (size_t i = 0; i < n2o.size(); ++i) tmp[i] = src[n2o[i]];
    dst.upload(tmp);
  };
  perm_upload(x_cell, m.x_cell, cell_n2o);
  perm_upload(y_cell, m.y_cell, cell_n2o);
  perm_upload(area_cell, m.area_cell, cell_n2o);
  perm_upload(x_edge, m.x_edge, edge_n2o);
  perm_upload(y_edge, m.y_edge, edge_n2o);
  perm_upload(angle_edge, m.angle_edge, edge_n2o);
  perm_upload(x_vertex, m.x_vertex, vert_n2o);
  perm_upload(y_vertex, m.y_vertex, vert_n2o);
  perm_upload(area_tri, m.area_triangle, vert_n2o);
  std::vector<double> ae((size_t)NC * ME, 0.0), dcs((size_t)NV * 3, 0.0);
  for (int i = 0; i < NC; ++i) {
    const int c = cell_n2o[i];
    for (int j = 0; j < noc_h[i]; ++j) {
      const int e = m.edges_on_cell[(size_t)c * ME + j];
      ae[(size_t)i * ME + j] = m.dv_edge[e] * m.dc_edge[e];
    }
  }
  for (int i = 0; i < NV; ++i) {
    const int v = vert_n2o[i];
    for (int j = 0; j < 3; ++j) {
      const int e = m.edges_on_vertex[3 * (size_t)v + j];
      dcs[3 * (size_t)i + j] = m.dc_edge[e] * m.edge_sign_on_vertex[3 * (size_t)v + j];
    }
  }
  ae_c.upload(ae);
  dcs_v.upload(dcs);
  uploaded = true;
}

MeshView DeviceMesh::view() const {
  MeshView v{};
  v.nc = host.ncells;
  v.ne = host.nedges;
  v.nv = host.nvertices;
  v.L = host.nlayers;
  v.me = host.max_edges;
  v.me2 = me2;
  v.noc = noc.p; v.eoc = eoc.p; v.coe = coe.p; v.voe = voe.p; v.noe = noe.p; v.eoe = eoe.p;
  v.eov = eov.p; v.cov = cov.p; v.eofs = eofs.p; v.vofs = vofs.p; v.maxlev_e = maxlev_e.p;
  v.cdiv = cdiv.p; v.cke = cke.p; v.w = w.p; v.inv_dc = inv_dc.p; v.inv_dv = inv_dv.p;
  v.inv_dv4 = inv_dv4.p; v.s2 = s2.p; v.s4 = s4.p; v.crv = crv.p; v.cth = cth.p;
  v.hex_ok = hex_ok ? 1 : 0;
  v.hexu_ok = hexu_ok ? 1 : 0;
  v.hexd_ok = hex_ok && hexd_ok ? 1 : 0;
  v.coef_uni = hex_ok && coef_uni ? 1 : 0;
  v.off32 = off32 ? 1 : 0;
  v.term_sets = 1;
  v.hexrec = hexrec.p;
  v.hexu = hexu.p;
  v.coc = coc.p;
  return v;
}

} // namespace omega_hip
