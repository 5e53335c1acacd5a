// LDS-tiled tendency kernels: per-tile connectivity in tile-local numbering.
//
// A tile is a run of consecutive device cells (a compact patch, since cells
// follow a space-filling curve) together with the edges it owns.  For each
// tile the host lists every entity whose rows the tile's tendencies read --
// its "union" sets -- and rewrites the stencils in indices into those sets.
// A block then loads each needed row once into LDS (coalesced, one row per
// lane group) and evaluates all stencils from LDS, instead of gathering the
// same neighbour rows from the cache hierarchy once per use.
#pragma once

#include "internal.hpp"

namespace omega_hip {

// Device view of the tile tables (all indices int32).
struct TileView {
  int ntiles, tcells;     // tiles and cells per tile
  // edge phase (vn tendency of the owned edges)
  const int *ue_ofs;      // [ntiles+1] union edges: owned edges first
  const int *ue;          // global edge ids
  const int4 *ue_cv;      // per union edge: local (cell0, cell1, vertex0, vertex1)
  const int *uc_ofs;      // [ntiles+1] union cells: cells of owned edges first
  const int *uc;
  const int *nkc;         // [ntiles] cells of owned edges (prefix of uc)
  const int *uv_ofs;      // [ntiles+1] union vertices: vertices of owned edges first
  const int *uv;
  const int *eoe_l;       // [ne][me2] local union-edge index of edges_on_edge
  // cell phase (h and tracer tendencies of the owned cells)
  const int *ce_ofs;      // [ntiles+1] edges of the owned cells
  const int *ce;
  const int *cc_ofs;      // [ntiles+1] owned cells first, then their neighbours
  const int *cc;
  const int4 *eoc_l;      // [nc][me] local (edge, cell0, cell1, 0) of each cell edge
  // largest set sizes (LDS allocation); max_eo: owned edges
  int max_ue, max_uc, max_nkc, max_uv, max_ce, max_cc, max_eo;
  // regular hexagonal mesh: per owned edge, its ten edges-on-edge, ten
  // distinct vertices and ten distinct cells (hex_pattern.hpp order) as
  // one-byte indices into the tile's union edges / vertices / cells
  // (32 bytes per edge, two uint4); null when unavailable
  const uint4 *hexl;
};

struct TileTables {
  int tcells = 0;
  int ntiles = 0;
  DBuf<int> ue_ofs, ue, uc_ofs, uc, nkc, uv_ofs, uv, eoe_l, ce_ofs, ce, cc_ofs, cc;
  DBuf<int4> ue_cv, eoc_l;
  DBuf<uint4> hexl;
  bool hexl_ok = false;
  int max_ue = 0, max_uc = 0, max_nkc = 0, max_uv = 0, max_ce = 0, max_cc = 0, max_eo = 0;
  TileView view() const;
};

// Builds (and verifies) the tables for tiles of `tcells` cells; without
// upload only the set sizes are filled in (host only).
void build_tile_tables(const DeviceMesh &dm, int tcells, TileTables &t, bool upload = true);

// Fused tendency + stage update from LDS tiles: one launch for the owned
// edges, one for the owned cells.  levels: levels per block (8 or 16).
void launch_tend_tile(const MeshView &m, const TileView &t, const Launch &lc,
                      const ModelCoef &mc, const StageArgs &a, const DiagOut &d, hipStream_t s);

// Hexagonal-mesh variant: the vn tendency of the owned edges from LDS rows
// (edges-on-edge velocities, cell thicknesses, vertex PV products) with the
// compile-time stencil pattern, then the cell tendencies as in the
// pattern-specialised kernels.  Usable when the tables carry hexl records.
bool hextile_usable(const MeshView &m, const TileTables &t, const Launch &lc);
void launch_tend_hextile(const MeshView &m, const TileView &t, const Launch &lc,
                         const ModelCoef &mc, const StageArgs &a, const DiagOut &d,
                         hipStream_t s);
// Cell part of the pattern-specialised tendency kernel alone (kernels_hex.hip).
void launch_tend_hex_cells(const MeshView &m, const Launch &lc, const ModelCoef &mc,
                           const StageArgs &a, const DiagOut &d, hipStream_t s);

} // namespace omega_hip
