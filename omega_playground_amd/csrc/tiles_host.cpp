// Host construction of the LDS tile tables (see tiles.hpp).
#include <algorithm>
#include <cstdint>
#include <cstring>
#include <unordered_map>

#include "hex_pattern.hpp"
#include "tiles.hpp"

namespace omega_hip {

namespace {

// Ordered union: `first` in its given order, then the remaining ids of
// `rest` ascending, without duplicates.
std::vector<int> ordered_union(const std::vector<int> &first, std::vector<int> rest) {
  std::vector<int> out = first;
  std::vector<int> sf = first;
  std::sort(sf.begin(), sf.end());
  std::sort(rest.begin(), rest.end());
  rest.erase(std::unique(rest.begin(), rest.end()), rest.end());
  for (int x : rest)
    if (!std::binary_search(sf.begin(), sf.end(), x)) out.push_back(x);
  return out;
}

std::vector<int> unique_in_order(const std::vector<int> &v) {
  std::vector<int> out;
  for (int x : v)
    if (std::find(out.begin(), out.end(), x) == out.end()) out.push_back(x);
  return out;
}

} // namespace

TileView TileTables::view() const {
  TileView v{};
  v.ntiles = ntiles;
  v.tcells = tcells;
  v.ue_ofs = ue_ofs.p;
  v.ue = ue.p;
  v.ue_cv = ue_cv.p;
  v.uc_ofs = uc_ofs.p;
  v.uc = uc.p;
  v.nkc = nkc.p;
  v.uv_ofs = uv_ofs.p;
  v.uv = uv.p;
  v.eoe_l = eoe_l.p;
  v.ce_ofs = ce_ofs.p;
  v.ce = ce.p;
  v.cc_ofs = cc_ofs.p;
  v.cc = cc.p;
  v.eoc_l = eoc_l.p;
  v.max_ue = max_ue;
  v.max_uc = max_uc;
  v.max_nkc = max_nkc;
  v.max_uv = max_uv;
  v.max_ce = max_ce;
  v.max_cc = max_cc;
  v.max_eo = max_eo;
  v.hexl = hexl_ok ? hexl.p : nullptr;
  return v;
}

void build_tile_tables(const DeviceMesh &dm, int tcells, TileTables &t, bool upload) {
  const HostMesh &hm = dm.host;
  const int NC = hm.ncells, NE = hm.nedges, ME = hm.max_edges, ME2 = dm.me2;
  OMEGA_REQUIRE(tcells >= 1, "tile size must be >= 1");
  const int ntiles = (NC + tcells - 1) / tcells;
  std::vector<int> ue_ofs{0}, ue, uc_ofs{0}, uc, nkc, uv_ofs{0}, uv, ce_ofs{0}, ce, cc_ofs{0}, cc;
  std::vector<int4> ue_cv, eoc_l((size_t)NC * ME, int4{0, 0, 0, 0});
  std::vector<int> eoe_l((size_t)NE * ME2, 0);
  // one-byte local stencil records of the hexagonal pattern (8 words/edge)
  bool hexl_ok = dm.hex_ok;
  std::vector<uint32_t> hexl(hexl_ok ? (size_t)NE * 8 : 0, 0u);
  t.max_ue = t.max_uc = t.max_nkc = t.max_uv = t.max_ce = t.max_cc = t.max_eo = 0;
  for (int ti = 0; ti < ntiles; ++ti) {
    const int c0 = ti * tcells, c1 = std::min(NC, c0 + tcells);
    const int e0 = dm.eofs_h[c0], e1 = dm.eofs_h[c1];
    // ---- edge phase
    std::vector<int> owned;
    for (int e = e0; e < e1; ++e) owned.push_back(e);
    std::vector<int> nb;
    for (int e = e0; e < e1; ++e)
      for (int j = 0; j < dm.noe_h[e]; ++j) nb.push_back(dm.eoe_h[(size_t)e * ME2 + j]);
    const std::vector<int> U = ordered_union(owned, nb);
    std::vector<int> kc_raw, vv_raw, uc_rest, uv_rest;
    for (int e = e0; e < e1; ++e)
      for (int s = 0; s < 2; ++s) {
        kc_raw.push_back(dm.coe_h[2 * (size_t)e + s]);
        vv_raw.push_back(dm.voe_h[2 * (size_t)e + s]);
      }
    std::vector<int> KC = unique_in_order(kc_raw), VV = unique_in_order(vv_raw);
    for (int x : U)
      for (int s = 0; s < 2; ++s) {
        uc_rest.push_back(dm.coe_h[2 * (size_t)x + s]);
        uv_rest.push_back(dm.voe_h[2 * (size_t)x + s]);
      }
    const std::vector<int> UC = ordered_union(KC, uc_rest), UV = ordered_union(VV, uv_rest);
    std::unordered_map<int, int> lu, lc, lv;
    for (int i = 0; i < (int)U.size(); ++i) lu[U[i]] = i;
    for (int i = 0; i < (int)UC.size(); ++i) lc[UC[i]] = i;
    for (int i = 0; i < (int)UV.size(); ++i) lv[UV[i]] = i;
    for (int x : U) {
      const int a0 = lc.at(dm.coe_h[2 * (size_t)x]), a1 = lc.at(dm.coe_h[2 * (size_t)x + 1]);
      const int b0 = lv.at(dm.voe_h[2 * (size_t)x]), b1 = lv.at(dm.voe_h[2 * (size_t)x + 1]);
      ue_cv.push_back(int4{a0, a1, b0, b1});
    }
    for (int e = e0; e < e1; ++e) {
      OMEGA_REQUIRE(lu.at(e) == e - e0, "tile tables: owned edges must lead the union");
      for (int j = 0; j < dm.noe_h[e]; ++j)
        eoe_l[(size_t)e * ME2 + j] = lu.at(dm.eoe_h[(size_t)e * ME2 + j]);
      const int4 q = ue_cv[ue_ofs.back() + (e - e0)];
      OMEGA_REQUIRE(q.x < (int)KC.size() && q.y < (int)KC.size(),
                    "tile tables: cells of owned edges must lead the union");
      OMEGA_REQUIRE(q.z < (int)VV.size() && q.w < (int)VV.size(),
                    "tile tables: vertices of owned edges must lead the union");
    }
    if (hexl_ok) {
      if (U.size() > 256 || UC.size() > 256 || UV.size() > 256) hexl_ok = false;
      for (int e = e0; e < e1 && hexl_ok; ++e) {
        const int *rec = &dm.hexrec_h[(size_t)e * hex::kRecInts];
        uint32_t *dst = &hexl[(size_t)e * 8];
        for (int b = 0; b < hex::kRecInts; ++b) {
          const std::unordered_map<int, int> &mp = b < 10 ? lu : b < 20 ? lv : lc;
          const auto it = mp.find(rec[b]);
          if (it == mp.end()) {
            hexl_ok = false;
            break;
          }
          dst[b >> 2] |= (uint32_t)it->second << (8 * (b & 3));
        }
      }
    }
    ue.insert(ue.end(), U.begin(), U.end());
    ue_ofs.push_back((int)ue.size());
    uc.insert(uc.end(), UC.begin(), UC.end());
    uc_ofs.push_back((int)uc.size());
    nkc.push_back((int)KC.size());
    uv.insert(uv.end(), UV.begin(), UV.end());
    uv_ofs.push_back((int)uv.size());
    t.max_ue = std::max(t.max_ue, (int)U.size());
    t.max_eo = std::max(t.max_eo, e1 - e0);
    t.max_uc = std::max(t.max_uc, (int)UC.size());
    t.max_nkc = std::max(t.max_nkc, (int)KC.size());
    t.max_uv = std::max(t.max_uv, (int)UV.size());
    // ---- cell phase
    std::vector<int> ownc, nbc, ced;
    for (int c = c0; c < c1; ++c) {
      ownc.push_back(c);
      for (int j = 0; j < dm.noc_h[c]; ++j) {
        const int e = dm.eoc_h[(size_t)c * ME + j];
        ced.push_back(e);
        nbc.push_back(dm.coe_h[2 * (size_t)e]);
        nbc.push_back(dm.coe_h[2 * (size_t)e + 1]);
      }
    }
    const std::vector<int> CE = ordered_union({}, ced), CC = ordered_union(ownc, nbc);
    std::unordered_map<int, int> le, lcc;
    for (int i = 0; i < (int)CE.size(); ++i) le[CE[i]] = i;
    for (int i = 0; i < (int)CC.size(); ++i) lcc[CC[i]] = i;
    for (int c = c0; c < c1; ++c) {
      OMEGA_REQUIRE(lcc.at(c) == c - c0, "tile tables: owned cells must lead the union");
      for (int j = 0; j < dm.noc_h[c]; ++j) {
        const int e = dm.eoc_h[(size_t)c * ME + j];
        eoc_l[(size_t)c * ME + j] = int4{le.at(e), lcc.at(dm.coe_h[2 * (size_t)e]),
                                         lcc.at(dm.coe_h[2 * (size_t)e + 1]), 0};
      }
    }
    ce.insert(ce.end(), CE.begin(), CE.end());
    ce_ofs.push_back((int)ce.size());
    cc.insert(cc.end(), CC.begin(), CC.end());
    cc_ofs.push_back((int)cc.size());
    t.max_ce = std::max(t.max_ce, (int)CE.size());
    t.max_cc = std::max(t.max_cc, (int)CC.size());
  }
  t.ntiles = ntiles;
  t.hexl_ok = hexl_ok && upload;
  if (!upload) return;
  t.tcells = tcells;
  t.ue_ofs.upload(ue_ofs);
  t.ue.upload(ue);
  t.ue_cv.upload(ue_cv);
  t.uc_ofs.upload(uc_ofs);
  t.uc.upload(uc);
  t.nkc.upload(nkc);
  t.uv_ofs.upload(uv_ofs);
  t.uv.upload(uv);
  t.eoe_l.upload(eoe_l);
  t.ce_ofs.upload(ce_ofs);
  t.ce.upload(ce);
  t.cc_ofs.upload(cc_ofs);
  t.cc.upload(cc);
  t.eoc_l.upload(eoc_l);
  if (t.hexl_ok) {
    std::vector<uint4> h4(hexl.size() / 4);
    std::memcpy(h4.data(), hexl.data(), hexl.size() * sizeof(uint32_t));
    t.hexl.upload(h4);
  } else {
    t.hexl.release();
  }
}

} // namespace omega_hip
