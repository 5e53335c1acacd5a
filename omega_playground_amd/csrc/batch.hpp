// Ensembles: members of different meshes, layer counts, tracer counts and
// coefficients advanced together, one launch per stage phase for the whole
// ensemble, through per-member descriptor tables.
#pragma once

#include "internal.hpp"

namespace omega_hip {

// Everything one member needs for one stage; a table of these per
// (step parity, stage) lives in device memory.
struct BatchMember {
  MeshView m;
  ModelCoef mc;
  DiagOut d;
  StageArgs a;
  int dterms;   // diagnostics terms (0: no diagnostics pass needed)
  int vdel4;    // momentum del4 pre-pass needed
  int lx, rows; // lanes per row and rows (entities) per block
  int chunks;   // level chunks (blocks per cell segment)
  int block0;   // first block of this member in its launch
  int hex;      // member runs the pattern-specialised kernels
};

// One launch per stage phase over every member's table entries; blk[b] =
// member of entry b.  persistent > 0 launches that many blocks (when fewer
// than the entries), each walking a contiguous run of entries; 0 launches
// one block per entry.  kinds: bit 0 the ensemble has general-kernel
// members, bit 1 pattern-specialised members.
void launch_batch_diag(const BatchMember *desc, const int *blk, int nblocks, int threads, int N,
                       int persistent, int kinds, hipStream_t s);
void launch_batch_vdel4(const BatchMember *desc, const int *blk, int nblocks, int threads, int N,
                        int persistent, hipStream_t s);
// tset > 0: every member of a pattern-specialised launch (N = 2, kinds ==
// 2) has both tendencies and the term set tset, one of term_set_batch's, and
// every 64k-layer member the shared-coefficient and 32-bit-offset forms
void launch_batch_tend(const BatchMember *desc, const int *blk, int nblocks, int threads, int N,
                       int persistent, int kinds, int tset, hipStream_t s);
bool term_set_batch(int terms);

} // namespace omega_hip
