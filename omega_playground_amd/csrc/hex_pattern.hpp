// Local stencil pattern of the regular hexagonal C-grid.
//
// On the doubly periodic hexagonal mesh every edge's ten edges-on-edge
// touch exactly ten distinct vertices and ten distinct cells.  Numbering
// those vertices and cells in the order they first appear when walking the
// edge's edges_on_edge list (vertex 0 then 1, cell 0 then 1 of each), the
// slot-to-local-index map is the same for all edges of one class, where the
// class is the edge's position (W, SW or SE) among its owner cell's edges.
// The tables below are that map; the host checks every edge against them
// before the pattern-specialised kernels are used (hex_layout.cpp), so a
// mesh that does not follow them falls back to the general kernels.
#pragma once

namespace omega_hip {
namespace hex {

constexpr int kSlots = 10;
constexpr int kUniq = 10;

// [class][slot] local index of vertices_on_edge[ej][0..1] and
// cells_on_edge[ej][0..1]
constexpr int LV0[3][kSlots] = {{0, 1, 3, 4, 5, 6, 7, 7, 8, 9},
                                {0, 2, 3, 4, 4, 6, 6, 7, 8, 0},
                                {0, 2, 3, 3, 4, 5, 6, 7, 9, 1}};
constexpr int LV1[3][kSlots] = {{1, 2, 2, 3, 4, 5, 6, 8, 9, 0},
                                {1, 1, 2, 3, 5, 5, 7, 8, 9, 9},
                                {1, 0, 2, 4, 5, 6, 7, 8, 8, 9}};
constexpr int LC0[3][kSlots] = {{0, 0, 3, 4, 5, 5, 7, 6, 6, 6},
                                {0, 2, 3, 4, 0, 5, 6, 6, 6, 1},
                                {0, 2, 3, 1, 1, 6, 6, 6, 9, 0}};
constexpr int LC1[3][kSlots] = {{1, 2, 0, 0, 0, 6, 6, 8, 9, 1},
                                {1, 0, 0, 0, 5, 6, 7, 8, 9, 6},
                                {1, 1, 1, 4, 5, 5, 7, 8, 6, 6}};
// local index of the edge's own vertices_on_edge[e][0..1] and
// cells_on_edge[e][0..1]
constexpr int OWNV[3][2] = {{5, 0}, {5, 0}, {5, 1}};
constexpr int OWNC[3][2] = {{0, 6}, {0, 6}, {1, 6}};

// Record per edge: edges_on_edge (10), unique vertices (10), unique cells (10)
constexpr int kRecInts = 30;

// Cell edges j < 3 are owned by the cell and list it second
// (cells_on_edge = (neighbour, cell)); edges j >= 3 list it first.
constexpr bool cell_is_second(int j) { return j < 3; }

} // namespace hex
} // namespace omega_hip
