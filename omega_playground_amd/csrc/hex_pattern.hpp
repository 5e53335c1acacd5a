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

namespace omega_hip {
namespace hexu {

// Union stencil of one cell on the regular hexagonal mesh: the cell's three
// owned edges (classes 0..2) read 19 distinct edges, 16 distinct vertices
// and 14 distinct cells.  Numbering them in first-appearance order (owned
// edges first; then, per owned edge in class order, its edges-on-edge and
// their vertices 0/1 and cells 0/1 slot by slot, then its own vertices and
// cells; then the cell's six edges and their cells), the maps below are the
// same for every cell; the host verifies that before these kernels run.
constexpr int kNE = 19, kNV = 16, kNC = 14;
constexpr int kRec = kNE + kNV + kNC; // ints per cell record

// [class][slot] union index of edges_on_edge and of its vertices / cells
constexpr int TE[3][10] = {{3, 4, 5, 6, 7, 1, 2, 8, 9, 10},
                           {7, 11, 12, 13, 14, 2, 8, 9, 10, 0},
                           {14, 15, 16, 17, 18, 8, 9, 10, 0, 1}};
constexpr int TV0[3][10] = {{0, 1, 3, 4, 5, 6, 7, 7, 8, 9},
                            {5, 10, 11, 12, 12, 7, 7, 8, 9, 5},
                            {12, 13, 14, 14, 15, 7, 8, 9, 5, 6}};
constexpr int TV1[3][10] = {{1, 2, 2, 3, 4, 5, 6, 8, 9, 0},
                            {4, 4, 10, 11, 6, 6, 8, 9, 0, 0},
                            {6, 12, 13, 15, 7, 8, 9, 0, 0, 5}};
constexpr int TC0[3][10] = {{0, 0, 3, 4, 5, 5, 7, 6, 6, 6},
                            {5, 4, 10, 11, 5, 7, 6, 6, 6, 0},
                            {5, 11, 12, 7, 7, 6, 6, 6, 0, 5}};
constexpr int TC1[3][10] = {{1, 2, 0, 0, 0, 6, 6, 8, 9, 1},
                            {0, 5, 5, 5, 7, 6, 8, 9, 1, 6},
                            {7, 7, 7, 13, 8, 8, 9, 1, 6, 6}};
// [class][0/1] union index of the owned edge's own vertices and cells
constexpr int OV[3][2] = {{5, 0}, {6, 5}, {7, 6}};
constexpr int OC[3][2] = {{0, 6}, {5, 6}, {7, 6}};
// the cell's six edges (edges_on_cell order) and their cells 0/1
constexpr int CE[6] = {0, 1, 2, 8, 9, 10};
constexpr int CC0[6] = {0, 5, 7, 6, 6, 6};
constexpr int CC1[6] = {6, 6, 6, 8, 9, 1};
constexpr int SELF = 6; // union index of the cell itself

// the seven cells of the tracer stencil (the cell and its neighbours), as
// positions into a compact array, and the union index of each position
constexpr int kNT = 7;
constexpr int TCELL[kNT] = {0, 1, 5, 6, 7, 8, 9};
constexpr int tpos(int u) {
  return u == 0 ? 0 : u == 1 ? 1 : u == 5 ? 2 : u == 6 ? 3 : u == 7 ? 4 : u == 8 ? 5 : 6;
}

} // namespace hexu
namespace hexd {

// Rows of the diagnostics of one cell: the two vertices a cell owns touch
// the cell's six edges plus two more edges, and only cells among the cell
// and its six neighbours.  Per owned vertex (device vertices 2c, 2c+1):
// [vertex][j] position of edges_on_vertex[j] in (edges_on_cell[0..6),
// extra edge 6, extra edge 7) and of cells_on_vertex[j] in (cell,
// cells_on_cell[0..6)).  The host verifies every cell before the kernels
// that rely on these maps run.
constexpr int VE[2][3] = {{0, 5, 6}, {0, 7, 1}};
constexpr int VC[2][3] = {{0, 6, 1}, {1, 2, 0}};

} // namespace hexd
} // namespace omega_hip
