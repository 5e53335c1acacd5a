// Records for the pattern-specialised kernels on the regular hexagonal mesh.
//
// For every edge the record holds its ten edges-on-edge and the ten distinct
// vertices and cells they touch, in first-appearance order; every edge must
// reproduce the tables of hex_pattern.hpp exactly, every cell must own three
// edges (its local edges 0..2, in class order) and two vertices, and every
// cell edge must list the cell second for j < 3 and first for j >= 3.  Any
// mismatch leaves hex_ok false and the general kernels are used.
#include <algorithm>

#include "hex_pattern.hpp"
#include "internal.hpp"

namespace omega_hip {

namespace {

int local_index(std::vector<int> &list, int id) {
  for (size_t i = 0; i < list.size(); ++i)
    if (list[i] == id) return (int)i;
  list.push_back(id);
  return (int)list.size() - 1;
}

} // namespace

bool build_hex_records(DeviceMesh &dm) {
  const HostMesh &m = dm.host;
  const int NC = m.ncells, NE = m.nedges, ME = m.max_edges;
  dm.hex_ok = false;
  dm.hexd_ok = false;
  if (ME < 6 || dm.me2 != hex::kSlots) return false;
  // cells_on_cell in device numbering, derived from the cell's edges
  dm.coc_h.assign((size_t)NC * 6, 0);
  for (int c = 0; c < NC; ++c) {
    if (dm.noc_h[c] != 6) return false;
    if (dm.eofs_h[c + 1] - dm.eofs_h[c] != 3 || dm.vofs_h[c + 1] - dm.vofs_h[c] != 2) return false;
    for (int j = 0; j < 6; ++j) {
      const int e = dm.eoc_h[(size_t)c * ME + j];
      if (j < 3 && e != dm.eofs_h[c] + j) return false;
      const int a0 = dm.coe_h[2 * (size_t)e], a1 = dm.coe_h[2 * (size_t)e + 1];
      const int other = hex::cell_is_second(j) ? a0 : a1;
      const int self = hex::cell_is_second(j) ? a1 : a0;
      if (self != c || other == c) return false;
      dm.coc_h[(size_t)c * 6 + j] = other;
    }
  }
  dm.hexrec_h.assign((size_t)NE * hex::kRecInts, 0);
  std::vector<int> uv, uc;
  for (int c = 0; c < NC; ++c) {
    for (int k = 0; k < 3; ++k) {
      const int e = dm.eofs_h[c] + k;
      if (dm.noe_h[e] != hex::kSlots) return false;
      uv.clear();
      uc.clear();
      int *rec = &dm.hexrec_h[(size_t)e * hex::kRecInts];
      for (int j = 0; j < hex::kSlots; ++j) {
        const int ej = dm.eoe_h[(size_t)e * dm.me2 + j];
        rec[j] = ej;
        const int v0 = local_index(uv, dm.voe_h[2 * (size_t)ej]);
        const int v1 = local_index(uv, dm.voe_h[2 * (size_t)ej + 1]);
        const int q0 = local_index(uc, dm.coe_h[2 * (size_t)ej]);
        const int q1 = local_index(uc, dm.coe_h[2 * (size_t)ej + 1]);
        if (v0 != hex::LV0[k][j] || v1 != hex::LV1[k][j] || q0 != hex::LC0[k][j] ||
            q1 != hex::LC1[k][j])
          return false;
      }
      if ((int)uv.size() != hex::kUniq || (int)uc.size() != hex::kUniq) return false;
      for (int s = 0; s < 2; ++s) {
        if (uv[hex::OWNV[k][s]] != dm.voe_h[2 * (size_t)e + s]) return false;
        if (uc[hex::OWNC[k][s]] != dm.coe_h[2 * (size_t)e + s]) return false;
      }
      std::copy(uv.begin(), uv.end(), rec + 10);
      std::copy(uc.begin(), uc.end(), rec + 20);
    }
  }
  dm.hex_ok = true;
  // diagnostics rows: every cell's owned vertices must follow the hexd maps
  dm.hexd_ok = true;
  for (int c = 0; c < NC && dm.hexd_ok; ++c) {
    int E[8], Cl[7];
    for (int j = 0; j < 6; ++j) E[j] = dm.eoc_h[(size_t)c * ME + j];
    Cl[0] = c;
    for (int j = 0; j < 6; ++j) Cl[1 + j] = dm.coc_h[(size_t)c * 6 + j];
    const int v0 = dm.vofs_h[c];
    E[6] = dm.eov_h[3 * (size_t)v0 + 2];
    E[7] = dm.eov_h[3 * (size_t)(v0 + 1) + 1];
    for (int i = 0; i < 2 && dm.hexd_ok; ++i)
      for (int j = 0; j < 3; ++j) {
        const int v = v0 + i;
        if (dm.eov_h[3 * (size_t)v + j] != E[hexd::VE[i][j]] ||
            dm.cov_h[3 * (size_t)v + j] != Cl[hexd::VC[i][j]])
          dm.hexd_ok = false;
      }
  }
  // per-cell union stencil: every cell must reproduce the hexu tables
  dm.hexu_ok = false;
  dm.hexu_h.assign((size_t)NC * hexu::kRec, 0);
  std::vector<int> ue, uv2, uc2;
  auto check = [](bool ok) {
    if (!ok) throw 0;
  };
  try {
    for (int c = 0; c < NC; ++c) {
      ue = {dm.eofs_h[c], dm.eofs_h[c] + 1, dm.eofs_h[c] + 2};
      uv2.clear();
      uc2.clear();
      for (int k = 0; k < 3; ++k) {
        const int e = dm.eofs_h[c] + k;
        for (int j = 0; j < hex::kSlots; ++j) {
          const int ej = dm.eoe_h[(size_t)e * dm.me2 + j];
          check(local_index(ue, ej) == hexu::TE[k][j]);
          check(local_index(uv2, dm.voe_h[2 * (size_t)ej]) == hexu::TV0[k][j]);
          check(local_index(uv2, dm.voe_h[2 * (size_t)ej + 1]) == hexu::TV1[k][j]);
          check(local_index(uc2, dm.coe_h[2 * (size_t)ej]) == hexu::TC0[k][j]);
          check(local_index(uc2, dm.coe_h[2 * (size_t)ej + 1]) == hexu::TC1[k][j]);
        }
        for (int s = 0; s < 2; ++s) {
          check(local_index(uv2, dm.voe_h[2 * (size_t)e + s]) == hexu::OV[k][s]);
          check(local_index(uc2, dm.coe_h[2 * (size_t)e + s]) == hexu::OC[k][s]);
        }
      }
      for (int j = 0; j < 6; ++j) {
        const int e = dm.eoc_h[(size_t)c * ME + j];
        check(local_index(ue, e) == hexu::CE[j]);
        check(local_index(uc2, dm.coe_h[2 * (size_t)e]) == hexu::CC0[j]);
        check(local_index(uc2, dm.coe_h[2 * (size_t)e + 1]) == hexu::CC1[j]);
      }
      check(local_index(uc2, c) == hexu::SELF);
      check((int)ue.size() == hexu::kNE && (int)uv2.size() == hexu::kNV &&
            (int)uc2.size() == hexu::kNC);
      int *rec = &dm.hexu_h[(size_t)c * hexu::kRec];
      std::copy(ue.begin(), ue.end(), rec);
      std::copy(uv2.begin(), uv2.end(), rec + hexu::kNE);
      std::copy(uc2.begin(), uc2.end(), rec + hexu::kNE + hexu::kNV);
    }
  } catch (int) {
    dm.hexu_h.clear();
    return true;
  }
  dm.hexu_ok = true;
  return true;
}

} // namespace omega_hip
