// Discrete adjoint, host side: transposed tables, the tendency
// vector-Jacobian product, the step adjoints of the RK4 and low-storage RK
// steppers (stage recomputation from the step's input state), a binomial
// checkpointing tape over many steps, and a least-squares misfit.
//
// The reference has no adjoint; the forward operators it differentiates are
// those of src/model/shallow_water_model.cpp:43-347 and the steppers of
// src/time_integration/rk4.cpp:28-104 and lsrk.cpp:36-89.
#include <algorithm>
#include <cstring>
#include <memory>

#include "adjoint.hpp"
#include "objects.hpp"

namespace omega_hip {

// ---------------------------------------------------------------------------
// Transposed tables
// ---------------------------------------------------------------------------
void build_adj_layout(const DeviceMesh &dm, AdjLayout &a) {
  const HostMesh &hm = dm.host;
  const int NC = hm.ncells, NE = hm.nedges, NV = hm.nvertices, ME = hm.max_edges,
            ME2 = dm.me2;
  std::vector<Real> ecd(2 * (size_t)NE), ecke(2 * (size_t)NE), ecrv(2 * (size_t)NE),
      ccth((size_t)NC * ME, 0.0), area_e(NE);
  std::vector<int> voc((size_t)NC * ME, 0), cslot(2 * (size_t)NE, 0);
  for (int e = 0; e < NE; ++e) {
    for (int s = 0; s < 2; ++s) {
      const int c = dm.coe_h[2 * (size_t)e + s];
      int slot = -1, hits = 0;
      for (int j = 0; j < dm.noc_h[c]; ++j)
        if (dm.eoc_h[(size_t)c * ME + j] == e) {
          slot = j;
          ++hits;
        }
      OMEGA_REQUIRE(hits == 1, "adjoint: an edge must appear exactly once in each of its cells "
                               "(mesh too small?)");
      cslot[2 * (size_t)e + s] = slot;
      ecd[2 * (size_t)e + s] = dm.cdiv_h[(size_t)c * ME + slot];
      ecke[2 * (size_t)e + s] = dm.cke_h[(size_t)c * ME + slot];
      const int v = dm.voe_h[2 * (size_t)e + s];
      slot = -1;
      hits = 0;
      for (int j = 0; j < 3; ++j)
        if (dm.eov_h[3 * (size_t)v + j] == e) {
          slot = j;
          ++hits;
        }
      OMEGA_REQUIRE(hits == 1, "adjoint: an edge must appear exactly once in each of its "
                               "vertices (mesh too small?)");
      ecrv[2 * (size_t)e + s] = dm.crv_h[3 * (size_t)v + slot];
    }
    const int eo = dm.edge_n2o[e];
    area_e[e] = 0.5 * hm.dc_edge[eo] * hm.dv_edge[eo];
  }
  for (int i = 0; i < NC; ++i) {
    const int c = dm.cell_n2o[i];
    for (int j = 0; j < dm.noc_h[i]; ++j) {
      const int v = dm.vert_o2n[hm.vertices_on_cell[(size_t)c * ME + j]];
      voc[(size_t)i * ME + j] = v;
      int slot = -1, hits = 0;
      for (int q = 0; q < 3; ++q)
        if (dm.cov_h[3 * (size_t)v + q] == i) {
          slot = q;
          ++hits;
        }
      OMEGA_REQUIRE(hits == 1, "adjoint: a cell must appear exactly once in each of its "
                               "vertices (mesh too small?)");
      ccth[(size_t)i * ME + j] = dm.cth_h[3 * (size_t)v + slot];
    }
  }
  // transposed edges_on_edge: for each edge x, the edges t listing x, in
  // increasing (t, slot) order (neighbouring edges' lists then start with
  // the same rows, which a wavefront's lanes load together)
  std::vector<int> tofs((size_t)NE + 1, 0);
  for (int t = 0; t < NE; ++t)
    for (int j = 0; j < dm.noe_h[t]; ++j) ++tofs[dm.eoe_h[(size_t)t * ME2 + j] + 1];
  for (int e = 0; e < NE; ++e) tofs[e + 1] += tofs[e];
  std::vector<int> fill(tofs.begin(), tofs.end() - 1), tedge(tofs[NE]);
  std::vector<Real> tw(tofs[NE]);
  for (int t = 0; t < NE; ++t)
    for (int j = 0; j < dm.noe_h[t]; ++j) {
      const int x = dm.eoe_h[(size_t)t * ME2 + j];
      tedge[fill[x]] = t;
      tw[fill[x]] = dm.w_h[(size_t)t * ME2 + j];
      ++fill[x];
    }
  a.fixed = true;
  for (int c = 0; c < NC && a.fixed; ++c) a.fixed = dm.noc_h[c] == 6;
  for (int e = 0; e < NE && a.fixed; ++e)
    a.fixed = dm.noe_h[e] == 10 && tofs[e + 1] - tofs[e] == 10;
  // slot-uniform form: the hexagonal numbering (cell c owns edges 3c..3c+2
  // and vertices 2c, 2c+1) and every coefficient the VJP reads equal at each
  // stencil slot, bit for bit, so the kernels read them once per slot
  bool uni = a.fixed && dm.hex_ok && dm.coef_uni && ME == 6 && ME2 == 10 &&
             dm.coc_h.size() == (size_t)NC * 6 && rows_equal(dm.inv_dv4_h, 1) &&
             rows_equal(ecd, 6) && rows_equal(ecke, 6) && rows_equal(ecrv, 6) &&
             rows_equal(ccth, 6);
  for (int c = 0; c <= NC && uni; ++c)
    uni = dm.eofs_h[c] == 3 * c && dm.vofs_h[c] == 2 * c;
  int ufirst[6] = {0}, ukslot[6] = {0}, uvsg[6] = {0};
  for (int c = 0; c < NC && uni; ++c)
    for (int j = 0; j < 6 && uni; ++j) {
      const int e = dm.eoc_h[(size_t)c * 6 + j];
      const int first = dm.coe_h[2 * (size_t)e] == c ? 1 : 0;
      const int other = dm.coe_h[2 * (size_t)e + first];
      if (c == 0) {
        ufirst[j] = first;
        ukslot[j] = e % 3;
      }
      uni = first == ufirst[j] && e % 3 == ukslot[j] && other == dm.coc_h[(size_t)c * 6 + j] &&
            dm.coe_h[2 * (size_t)e + 1 - first] == c;
    }
  for (int v = 0; v < NV && uni; ++v)
    for (int j = 0; j < 3 && uni; ++j) {
      const int e = dm.eov_h[3 * (size_t)v + j];
      const int sg = dm.voe_h[2 * (size_t)e + 1] == v ? 1 : 0;
      if (v < 2) uvsg[3 * v + j] = sg;
      uni = sg == uvsg[3 * (v & 1) + j];
    }
  a.uni = uni;
  std::memcpy(a.ufirst, ufirst, sizeof ufirst);
  std::memcpy(a.ukslot, ukslot, sizeof ukslot);
  std::memcpy(a.uvsg, uvsg, sizeof uvsg);
  a.ecd.upload(ecd);
  a.ecke.upload(ecke);
  a.ecrv.upload(ecrv);
  a.ccth.upload(ccth);
  a.voc.upload(voc);
  a.tofs.upload(tofs);
  a.tedge.upload(tedge);
  a.tw.upload(tw);
  a.area_e.upload(area_e);
  a.cslot_h = std::move(cslot);
  a.built = true;
}

namespace {

// Builds the mesh's transposed tables, the model's transposed tracer
// coefficients and the cotangent scratch, once.
void ensure_adjoint(omega_model_s *M) {
  OMEGA_REQUIRE(!M->linear, "adjoint: not available for the linear model");
  OMEGA_REQUIRE(M->add_fn == nullptr,
                "adjoint: not available while an additional-tendency hook is set");
  OMEGA_REQUIRE(!(M->terms & T_DRAG), "adjoint: the bottom drag term has no adjoint");
  DeviceMesh &dm = M->mesh->dm;
  dm.upload();
  if (!M->mesh->adj.built) build_adj_layout(dm, M->mesh->adj);
  const HostMesh &hm = dm.host;
  const size_t NC = hm.ncells, NE = hm.nedges, NV = hm.nvertices, L = hm.nlayers,
               ME = hm.max_edges, ntr = M->params.ntracers;
  if (!M->ect2.p) {
    std::vector<Real> e2(2 * NE), e4(2 * NE);
    const std::vector<int> &cs = M->mesh->adj.cslot_h;
    for (size_t e = 0; e < NE; ++e)
      for (int s = 0; s < 2; ++s) {
        const size_t c = dm.coe_h[2 * e + s];
        e2[2 * e + s] = M->ct2_h[c * ME + cs[2 * e + s]];
        e4[2 * e + s] = M->ct4_h[c * ME + cs[2 * e + s]];
      }
    M->ect2.upload(e2);
    M->ect4.upload(e4);
    M->ect_uniform = rows_equal(e2, 6) && rows_equal(e4, 6);
  }
  if (!M->a_sbar.p) {
    M->a_se.alloc(NE * L);
    M->a_fu.alloc(NE * L);
    M->a_px.alloc(NE * L);
    if (ntr) M->a_xa.alloc(ntr * NE * L);
    M->a_sbar.alloc(NE * L);
    M->a_fbar.alloc(NE * L);
    M->a_e2bar.alloc(NE * L);
    M->a_zbar.alloc(NV * L);
    M->a_tbar.alloc(NV * L);
    M->a_z2bar.alloc(NV * L);
    M->a_kbar.alloc(NC * L);
    M->a_dbar.alloc(NC * L);
    M->a_d2bar.alloc(NC * L);
    if (ntr) M->a_gbar.alloc(ntr * NC * L);
    M->a_nparts = adj_partial_slots(dm.view(), M->lc);
    M->a_partial.alloc(4 * (size_t)M->a_nparts);
    M->a_pgrad.alloc(4);
  }
  if (M->terms & T_VDEL4) ensure_vdel4(M);
}

AdjScratch scratch(omega_model_s *M) {
  return AdjScratch{M->a_sbar.p, M->a_zbar.p,  M->a_tbar.p,  M->a_kbar.p,
                    M->a_dbar.p, M->a_gbar.p,  M->a_fbar.p,  M->a_d2bar.p,
                    M->a_z2bar.p, M->a_e2bar.p, M->a_partial.p, M->a_pgrad.p,
                    M->a_se.p, M->a_fu.p, M->a_px.p, M->a_xa.p};
}

// out (+)= J(y)^T k for the tendency at state y (pointers in device numbering;
// epi says how the result is combined);
// the mixing-coefficient gradient accumulates in M->a_pgrad.
// pre: the rows the stage kernel wrote at y (the pre-pass is skipped);
// part: bit 0 adds the coefficient-gradient partials to those of an earlier
// VJP, bit 1 leaves them for a later VJP to reduce (AdjArgs::part_add/keep)
void tendency_vjp_ptrs(omega_model_s *M, const Real *const y[3], const Real *const k[3],
                       Real *const out[3], hipStream_t s, const AdjEpi &epi = AdjEpi{},
                       bool have_diag = false, const omega_model_s::PreSet *pre = nullptr,
                       int part = 0) {
  if (!have_diag) run_diag(M, y[0], y[1], y[2], s);
  AdjArgs a{};
  a.terms = M->terms;
  a.do_h = !M->params.disable_h_tendency;
  a.do_vn = !M->params.disable_vn_tendency;
  a.ntr = M->params.ntracers;
  a.grav = M->params.grav;
  a.h = y[0];
  a.vn = y[1];
  a.tr = y[2];
  a.c = CotIO{k[0], k[1], k[2], out[0], out[1], out[2]};
  a.nparts = M->a_nparts;
  a.epi = epi;
  a.part_add = part & 1;
  a.part_keep = (part >> 1) & 1;
  const AdjCoef ac{M->ect2.p, M->ect4.p, M->ect_uniform ? 1 : 0};
  AdjScratch sc = scratch(M);
  if (pre) {
    sc.se_e = pre->se.p;
    sc.fu_e = pre->fu.p;
    sc.own_e = pre->own.p;
  }
  launch_tendency_vjp(M->view(), M->mesh->adj.view(), M->lc, M->coef(), ac, a, M->diag(), sc,
                      s, pre != nullptr);
}

void zero_pgrad(omega_model_s *M, hipStream_t s) { launch_fill(M->a_pgrad.p, 0.0, 4, s); }

// param_grad[0..3] += the accumulated gradient (synchronises the stream).
void add_pgrad(omega_model_s *M, double *param_grad, hipStream_t s) {
  if (!param_grad) return;
  double g[4];
  OMEGA_HIP_CHECK(hipMemcpyAsync(g, M->a_pgrad.p, sizeof g, hipMemcpyDeviceToHost, s));
  OMEGA_HIP_CHECK(hipStreamSynchronize(s));
  for (int i = 0; i < 4; ++i) param_grad[i] += g[i];
}

void copy_fields(Real *const dst[3], const Real *const src[3], const size_t n[3],
                 hipStream_t s) {
  for (int f = 0; f < 3; ++f)
    if (n[f])
      OMEGA_HIP_CHECK(
          hipMemcpyAsync(dst[f], src[f], n[f] * sizeof(Real), hipMemcpyDeviceToDevice, s));
}

void ensure_sized(DBuf<Real> (&b)[3], const size_t n[3]) {
  for (int f = 0; f < 3; ++f)
    if (b[f].n != n[f]) b[f].alloc(n[f]);
}

// Diagnostic sets 1..nsets for the stage inputs of a step adjoint, sized
// like the model's own buffers.
void ensure_diag_sets(omega_model_s *M, int nsets) {
  if (M->terms & T_VDEL4) ensure_vdel4(M);
  auto fit = [](DBuf<Real> &b, const DBuf<Real> &like) {
    if (b.n != like.n) b.alloc(like.n);
  };
  for (int i = 0; i < nsets; ++i) {
    omega_model_s::DiagSet &d = M->diag_sets[i];
    fit(d.rvort_v, M->rvort_v);
    fit(d.inv_thick_v, M->inv_thick_v);
    fit(d.ke_c, M->ke_c);
    fit(d.div_c, M->div_c);
    fit(d.trdel2_c, M->trdel2_c);
    fit(d.vdel2_e, M->vdel2_e);
    fit(d.vdel2_rv_v, M->vdel2_rv_v);
    fit(d.vdel2_div_c, M->vdel2_div_c);
  }
}

// The tendency VJP's rows at stage inputs 0..nsets-1 when the stage kernels
// write them; nullptr entries otherwise.
std::array<const omega_model_s::PreSet *, 4> ensure_pre_sets(omega_model_s *M, int nsets) {
  std::array<const omega_model_s::PreSet *, 4> out{};
  if (!stage_writes_pre(M)) return out;
  const size_t n = (size_t)M->mesh->dm.host.nedges * M->mesh->dm.host.nlayers;
  for (int i = 0; i < nsets; ++i) {
    omega_model_s::PreSet &p = M->pre_sets[i];
    for (DBuf<Real> *b : {&p.se, &p.fu, &p.own})
      if (b->n != n) b->alloc(n);
    out[i] = &p;
  }
  return out;
}

PreOut pre_out(const omega_model_s::PreSet *p) {
  return p ? PreOut{p->se.p, p->fu.p, p->own.p} : PreOut{};
}

// Selects the diagnostic set for the duration of a step adjoint; every
// other caller always sees set 0.
struct DiagSlotScope {
  omega_model_s *M;
  explicit DiagSlotScope(omega_model_s *m) : M(m) {}
  void use(int slot) { M->diag_slot = slot; }
  ~DiagSlotScope() { M->diag_slot = 0; }
};

} // namespace

// ---------------------------------------------------------------------------
// Step adjoints
// ---------------------------------------------------------------------------

// cot <- (d y_{n+1} / d y_n)^T cot for one step from y (device pointers,
// sizes n).  The stage states are recomputed from y with the forward
// kernels, so they are bit-identical to those of the forward step.
void step_vjp_ptrs(omega_stepper_s *St, const Real *const y[3], Real *const cot[3],
                   const size_t n[3], double dt, hipStream_t s) {
  omega_model_s *M = St->model;
  ensure_adjoint(M);
  for (int i = 0; i < 4; ++i) ensure_sized(St->adj_y[i], n);
  ensure_sized(St->adj_t, n);
  ensure_sized(St->adj_gy, n);
  ensure_sized(St->adj_pb, n);
  ensure_sized(St->adj_kb, n);
  auto P = [&](DBuf<Real> (&b)[3]) {
    return std::array<Real *, 3>{b[0].p, b[1].p, b[2].p};
  };
  auto K = [&](DBuf<Real> (&b)[3]) {
    return std::array<const Real *, 3>{b[0].p, b[1].p, b[2].p};
  };
  const std::array<const Real *, 3> Y{y[0], y[1], y[2]};
  const std::array<Real *, 3> C{cot[0], cot[1], cot[2]};
  // the diagnostics of stage input s (computed by the recomputation) stay in
  // set s for the VJP at that input; the last stage input's are computed by
  // its VJP
  DiagSlotScope slots(M);
  if (St->kind == OMEGA_STEPPER_RK4) {
    ensure_diag_sets(M, 3);
    // the stage kernels at P0..P2 also write the VJP's rows there, where they can
    const auto pre = ensure_pre_sets(M, 3);
    // stage inputs P1..P3 (P_{i+1} = y + dt*a_i*k_i), as rk4.cpp:168-189
    for (int st = 0; st < 3; ++st) {
      slots.use(st);
      StageArgs a{};
      a.epi = EPI_PROVIS;
      a.da = dt * St->rka[st];
      const std::array<const Real *, 3> in = st == 0 ? Y : K(St->adj_y[st - 1]);
      const auto out = P(St->adj_y[st]);
      a.h = fio(in[0], nullptr, nullptr, y[0], out[0]);
      a.vn = fio(in[1], nullptr, nullptr, y[1], out[1]);
      a.tr = fio(in[2], nullptr, nullptr, y[2], out[2]);
      const PreOut po = pre_out(pre[st]);
      run_stage(M, a, s, 7, true, pre[st] ? &po : nullptr);
    }
    // reverse sweep with gy = the incoming cotangent (left in place until the
    // last VJP): acc = gy + sum_s J(P_s)^T kbar_s and, in the same kernels,
    // kbar_{s-1} = dt*b_{s-1}*gy + dt*a_{s-1}*J(P_s)^T kbar_s.  kbar_3 =
    // dt*b_3*gy enters as gy with the result scaled; two kbar buffers
    // alternate, since each VJP reads its kbar at neighbours.
    const auto acc = P(St->adj_gy);
    std::array<Real *, 3> kin = P(St->adj_kb), kout = P(St->adj_pb);
    const std::array<const Real *, 3> gy{C[0], C[1], C[2]};
    for (int st = 3; st >= 1; --st) {
      const auto ps = K(St->adj_y[st - 1]);
      const std::array<const Real *, 3> kc =
          st == 3 ? gy : std::array<const Real *, 3>{kin[0], kin[1], kin[2]};
      AdjEpi epi{};
      epi.rk4 = 1;
      epi.db = dt * St->rkb[st - 1];
      epi.da = dt * St->rka[st - 1];
      if (st == 3) epi.vs = dt * St->rkb[3];
      for (int f = 0; f < 3; ++f) {
        epi.gy[f] = gy[f];
        epi.kbo[f] = kout[f];
        if (st == 3) epi.base[f] = gy[f];
      }
      slots.use(st);
      // the four VJPs' coefficient-gradient partials are reduced once, by the last
      tendency_vjp_ptrs(M, ps.data(), kc.data(), acc.data(), s, epi, st < 3,
                        st < 3 ? pre[st] : nullptr, st == 3 ? 2 : 3);
      std::swap(kin, kout);
    }
    // cot = acc + J(y)^T kbar_0
    const std::array<const Real *, 3> kc{kin[0], kin[1], kin[2]};
    AdjEpi last{};
    for (int f = 0; f < 3; ++f) last.base[f] = acc[f];
    slots.use(0);
    tendency_vjp_ptrs(M, Y.data(), kc.data(), C.data(), s, last, true, pre[0], 1);
  } else {
    // stage inputs Y_1..Y_{S-1} (lsrk.cpp:36-89: tend = a_s*tend + F(Y_s),
    // Y_{s+1} = Y_s + dt*b_s*tend)
    const int S = St->nstages;
    OMEGA_REQUIRE(S >= 1 && S <= 5, "step adjoint: unsupported number of stages");
    const auto tb = P(St->adj_t);
    ensure_diag_sets(M, S - 1);
    const auto pre = ensure_pre_sets(M, S - 1);
    for (int st = 0; st + 1 < S; ++st) {
      slots.use(st);
      StageArgs a{};
      a.epi = EPI_LSRK;
      a.a = St->rka[st];
      a.base_from_tend = St->rka[st] != 0.0;
      a.db = dt * St->rkb[st];
      const std::array<const Real *, 3> in = st == 0 ? Y : K(St->adj_y[st - 1]);
      const auto out = P(St->adj_y[st]);
      a.h = fio(in[0], tb[0], nullptr, nullptr, out[0]);
      a.vn = fio(in[1], tb[1], nullptr, nullptr, out[1]);
      a.tr = fio(in[2], tb[2], nullptr, nullptr, out[2]);
      const PreOut po = pre_out(pre[st]);
      run_stage(M, a, s, 7, true, pre[st] ? &po : nullptr);
    }
    // Tbar_s = dt*b_s*Ybar_{s+1} + a_{s+1}*Tbar_{s+1};  Ybar_s = Ybar_{s+1} + J(Y_s)^T Tbar_s.
    // The VJP at Y_s writes Ybar_s and, in the same kernels, Tbar_{s-1} =
    // a_s*Tbar_s + dt*b_{s-1}*Ybar_s.  Tbar_{S-1} = dt*b_{S-1}*Ybar_S enters as
    // the incoming cotangent with the result scaled; Ybar accumulates in acc
    // (the incoming cotangent is read by that first VJP at neighbours) and
    // the last VJP writes it to cot; two Tbar buffers alternate.
    const auto acc = P(St->adj_gy);
    std::array<Real *, 3> kin = P(St->adj_kb), kout = P(St->adj_pb);
    const std::array<const Real *, 3> gy{C[0], C[1], C[2]};
    for (int st = S - 1; st >= 0; --st) {
      const bool first = st == S - 1, last = st == 0;
      const std::array<const Real *, 3> ys = st == 0 ? Y : K(St->adj_y[st - 1]);
      const std::array<const Real *, 3> kc =
          first ? gy : std::array<const Real *, 3>{kin[0], kin[1], kin[2]};
      AdjEpi epi{};
      if (first) epi.vs = dt * St->rkb[st];
      if (!last) {
        epi.lsrk = 1;
        epi.db = dt * St->rkb[st - 1];
        epi.da = first ? St->rka[st] * (dt * St->rkb[st]) : St->rka[st];
      }
      for (int f = 0; f < 3; ++f) {
        epi.base[f] = first ? gy[f] : acc[f];
        epi.kbo[f] = kout[f];
      }
      Real *const *out = (last && !first) ? C.data() : acc.data();
      slots.use(st);
      // the S VJPs' coefficient-gradient partials are reduced once, by the last
      tendency_vjp_ptrs(M, ys.data(), kc.data(), out, s, epi, st + 1 < S,
                        st + 1 < S ? pre[st] : nullptr, (first ? 0 : 1) | (last ? 0 : 2));
      std::swap(kin, kout);
    }
    if (S == 1) copy_fields(C.data(), K(St->adj_gy).data(), n, s);
  }
}

} // namespace omega_hip

// ---------------------------------------------------------------------------
// Binomial checkpointing tape
// ---------------------------------------------------------------------------
using namespace omega_hip;

struct omega_tape_s {
  omega_stepper_s *stepper = nullptr;
  int nsteps = 0;
  int nsnap_req = 1;
  int64_t byte_cap = 0;
  int nsnap = 0;                      // snapshots in use (y_0 included)
  size_t n[3] = {0, 0, 0};
  int ntr = 0;
  std::vector<std::array<DBuf<Real>, 3>> slots;
  std::unique_ptr<omega_state_s> work; // state advanced during the reverse sweep
  std::vector<std::pair<int, int>> descent; // (step index, slot) stored by the forward sweep
  double t0 = 0, dt = 0;
  bool have_forward = false;
  int64_t forward_steps = 0;
};

namespace {

// Minimal number of forward steps to reverse l steps from a stored state
// with c snapshots (the stored one included): r*l - C(c+r, c+1) with r the
// least integer such that C(c+r, c) >= l (Griewank's binomial bound).
int64_t binom(int64_t n, int64_t k) {
  if (k < 0 || k > n) return 0;
  k = std::min(k, n - k);
  long double r = 1;
  for (int64_t i = 1; i <= k; ++i) r = r * (n - k + i) / i;
  return (int64_t)(r + 0.5L);
}

int64_t revolve_cost(int64_t l, int c) {
  if (l <= 1) return 0;
  if (c <= 1) return l * (l - 1) / 2;
  int64_t r = 0;
  while (binom(c + r, c) < l) ++r;
  return r * l - binom(c + r, c + 1);
}

// Optimal first checkpoint position (steps from the stored state) for l
// steps and c snapshots (c >= 2).
int revolve_split(int l, int c) {
  int best = 1;
  int64_t bc = -1;
  for (int m = 1; m < l; ++m) {
    const int64_t v = m + revolve_cost(l - m, c - 1) + revolve_cost(m, c);
    if (bc < 0 || v < bc) {
      bc = v;
      best = m;
    }
  }
  return best;
}

std::array<Real *, 3> state_ptrs(omega_state_s *S) { return {S->ptr(0), S->ptr(1), S->ptr(2)}; }
std::array<Real *, 3> slot_ptrs(std::array<DBuf<Real>, 3> &b) { return {b[0].p, b[1].p, b[2].p}; }

void store(omega_tape_s *T, int slot, omega_state_s *S, hipStream_t s) {
  const auto src = state_ptrs(S);
  const std::array<const Real *, 3> c{src[0], src[1], src[2]};
  auto dst = slot_ptrs(T->slots[slot]);
  copy_fields(dst.data(), c.data(), T->n, s);
}

void restore(omega_tape_s *T, int slot, omega_state_s *S, hipStream_t s) {
  auto src = slot_ptrs(T->slots[slot]);
  const std::array<const Real *, 3> c{src[0], src[1], src[2]};
  auto dst = state_ptrs(S);
  copy_fields(dst.data(), c.data(), T->n, s);
}

void advance(omega_tape_s *T, omega_state_s *S, int from, int count, hipStream_t s) {
  for (int i = 0; i < count; ++i)
    do_step_maybe_graph(T->stepper, S, T->t0 + (from + i) * T->dt, T->dt, s);
  T->forward_steps += count;
}

void step_adjoint_at(omega_tape_s *T, omega_state_s *Y, omega_state_s *C, hipStream_t s) {
  const auto y = state_ptrs(Y);
  const std::array<const Real *, 3> yc{y[0], y[1], y[2]};
  const auto c = state_ptrs(C);
  step_vjp_ptrs(T->stepper, yc.data(), c.data(), T->n, T->dt, s);
}

// Reverses steps [a, b): slot sa holds y_a, cot holds dJ/dy_b on entry and
// dJ/dy_a on exit; c snapshots are available from slot sa on.  With desc set
// the first checkpoint of every nested right half was stored by the forward
// sweep.
void reverse_range(omega_tape_s *T, int a, int b, int sa, int c, bool desc, omega_state_s *C,
                   hipStream_t s) {
  omega_state_s *W = T->work.get();
  if (b - a == 1) {
    restore(T, sa, W, s);
    step_adjoint_at(T, W, C, s);
    return;
  }
  if (c <= 1) {
    for (int i = b - 1; i >= a; --i) {
      restore(T, sa, W, s);
      advance(T, W, a, i - a, s);
      step_adjoint_at(T, W, C, s);
    }
    return;
  }
  const int m = revolve_split(b - a, c);
  const int sn = sa + 1;
  if (!desc) {
    restore(T, sa, W, s);
    advance(T, W, a, m, s);
    store(T, sn, W, s);
  }
  reverse_range(T, a + m, b, sn, c - 1, desc, C, s);
  reverse_range(T, a, a + m, sa, c, false, C, s);
}

void tape_forward(omega_tape_s *T, double t0, double dt, omega_state_s *S, hipStream_t s) {
  omega_model_s *M = T->stepper->model;
  check_state(M, S, "tape forward");
  ensure_adjoint(M);
  // snapshot storage, sized to this state
  const size_t bytes = (S->n[0] + S->n[1] + S->n[2]) * sizeof(Real);
  int nsnap = T->nsnap_req;
  if (T->byte_cap > 0) nsnap = (int)std::min<int64_t>(nsnap, T->byte_cap / (int64_t)bytes);
  OMEGA_REQUIRE(nsnap >= 1, "tape: the byte cap does not hold one state snapshot");
  nsnap = std::min(nsnap, std::max(1, T->nsteps));
  const bool resize = T->nsnap != nsnap || T->n[0] != S->n[0] || T->n[1] != S->n[1] ||
                      T->n[2] != S->n[2];
  if (resize) {
    T->slots.clear();
    T->slots.resize(nsnap);
    for (auto &sl : T->slots)
      for (int f = 0; f < 3; ++f) sl[f].alloc(S->n[f]);
    T->nsnap = nsnap;
    for (int f = 0; f < 3; ++f) T->n[f] = S->n[f];
    T->ntr = S->ntr;
    T->work = std::make_unique<omega_state_s>();
    T->work->mesh = S->mesh;
    T->work->ntr = S->ntr;
    for (int f = 0; f < 3; ++f) {
      T->work->n[f] = S->n[f];
      T->work->buf[0][f].alloc(S->n[f]);
    }
  }
  T->t0 = t0;
  T->dt = dt;
  T->forward_steps = 0;
  // first descent of the reversal: checkpoints the forward sweep stores
  T->descent.clear();
  int a = 0, c = T->nsnap, slot = 0;
  while (c > 1 && T->nsteps - a > 1) {
    a += revolve_split(T->nsteps - a, c);
    T->descent.emplace_back(a, ++slot);
    --c;
  }
  store(T, 0, S, s);
  size_t next = 0;
  for (int i = 0; i < T->nsteps; ++i) {
    do_step_maybe_graph(T->stepper, S, t0 + i * dt, dt, s);
    ++T->forward_steps;
    if (next < T->descent.size() && T->descent[next].first == i + 1)
      store(T, T->descent[next++].second, S, s);
  }
  T->have_forward = true;
}

void tape_reverse(omega_tape_s *T, omega_state_s *C, double *param_grad, hipStream_t s) {
  OMEGA_REQUIRE(T->have_forward, "tape reverse: run the forward sweep first");
  omega_model_s *M = T->stepper->model;
  check_state(M, C, "tape reverse");
  OMEGA_REQUIRE(C->n[0] == T->n[0] && C->n[1] == T->n[1] && C->n[2] == T->n[2],
                "tape reverse: cotangent shape differs from the recorded state");
  ensure_adjoint(M);
  zero_pgrad(M, s);
  if (T->nsteps > 0) reverse_range(T, 0, T->nsteps, 0, T->nsnap, true, C, s);
  add_pgrad(M, param_grad, s);
  OMEGA_HIP_CHECK(hipStreamSynchronize(s));
}

// J = sum over fields of w_f/2 * sum_i area_i * sum_k (y - obs)^2, cot = dJ/dy.
double misfit(omega_model_s *M, omega_state_s *S, omega_state_s *O, const double w[3],
              omega_state_s *C, hipStream_t s) {
  ensure_adjoint(M);
  DeviceMesh &dm = M->mesh->dm;
  const HostMesh &hm = dm.host;
  const int L = hm.nlayers, NC = hm.ncells, NE = hm.nedges;
  const int ntr = std::min(S->ntr, std::min(O->ntr, C->ntr));
  const Real *area[3] = {dm.area_cell.p, M->mesh->adj.area_e.p, dm.area_cell.p};
  const int nent[3] = {NC, NE, NC};
  const int nrep[3] = {1, 1, ntr};
  int off[4] = {0, 0, 0, 0};
  for (int f = 0; f < 3; ++f)
    off[f + 1] = off[f] + (nrep[f] ? misfit_blocks((size_t)nrep[f] * nent[f] * L) : 0);
  if ((int)M->a_mpart.n < off[3] + 1) M->a_mpart.alloc(off[3] + 1);
  for (int f = 0; f < 3; ++f) {
    if (!nrep[f]) continue;
    launch_misfit(S->ptr(f), O->ptr(f), area[f], nent[f], L, nrep[f], w[f], C->ptr(f),
                  M->a_mpart.p + off[f], s);
  }
  // extra tracers of the cotangent (beyond those compared) are zeroed
  if (C->n[2] > (size_t)ntr * NC * L)
    launch_fill(C->ptr(2) + (size_t)ntr * NC * L, 0.0, C->n[2] - (size_t)ntr * NC * L, s);
  launch_sum(M->a_mpart.p, off[3], M->a_mpart.p + off[3], 0, s);
  double J = 0;
  OMEGA_HIP_CHECK(
      hipMemcpyAsync(&J, M->a_mpart.p + off[3], sizeof J, hipMemcpyDeviceToHost, s));
  OMEGA_HIP_CHECK(hipStreamSynchronize(s));
  return J;
}

} // namespace

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
extern "C" {

int omega_model_tendency_vjp(omega_model_t M, omega_state_t state, omega_state_t tend_cot,
                             omega_state_t state_cot, double *param_grad,
                             omega_stream_t stream) {
  return guarded([&] {
    OMEGA_REQUIRE(M, "null model");
    check_state(M, state, "tendency_vjp state");
    check_state(M, tend_cot, "tendency_vjp tendency cotangent");
    check_state(M, state_cot, "tendency_vjp state cotangent");
    ensure_adjoint(M);
    hipStream_t s = as_stream(stream);
    zero_pgrad(M, s);
    const Real *y[3] = {state->ptr(0), state->ptr(1), state->ptr(2)};
    const Real *k[3] = {tend_cot->ptr(0), tend_cot->ptr(1), tend_cot->ptr(2)};
    Real *o[3] = {state_cot->ptr(0), state_cot->ptr(1), state_cot->ptr(2)};
    tendency_vjp_ptrs(M, y, k, o, s);
    add_pgrad(M, param_grad, s);
  });
}

int omega_stepper_step_vjp(omega_stepper_t St, double t, double dt, omega_state_t state,
                           omega_state_t cot, double *param_grad, omega_stream_t stream) {
  (void)t;
  return guarded([&] {
    OMEGA_REQUIRE(St, "null stepper");
    omega_model_s *M = St->model;
    check_state(M, state, "step_vjp state");
    check_state(M, cot, "step_vjp cotangent");
    OMEGA_REQUIRE(cot->n[2] == state->n[2], "step_vjp: cotangent and state tracer counts differ");
    ensure_adjoint(M);
    hipStream_t s = as_stream(stream);
    zero_pgrad(M, s);
    const Real *y[3] = {state->ptr(0), state->ptr(1), state->ptr(2)};
    Real *c[3] = {cot->ptr(0), cot->ptr(1), cot->ptr(2)};
    step_vjp_ptrs(St, y, c, state->n, dt, s);
    add_pgrad(M, param_grad, s);
  });
}

int omega_tape_create(omega_stepper_t St, int nsteps, int n_snapshots, int64_t byte_cap,
                      omega_tape_t *out) {
  return guarded([&] {
    OMEGA_REQUIRE(St && out, "null argument");
    OMEGA_REQUIRE(nsteps >= 0, "nsteps must be >= 0");
    OMEGA_REQUIRE(n_snapshots >= 1, "n_snapshots must be >= 1");
    OMEGA_REQUIRE(byte_cap >= 0, "byte_cap must be >= 0");
    auto T = std::make_unique<omega_tape_s>();
    T->stepper = St;
    T->nsteps = nsteps;
    T->nsnap_req = n_snapshots;
    T->byte_cap = byte_cap;
    *out = T.release();
  });
}

int omega_tape_destroy(omega_tape_t T) {
  return guarded([&] { delete T; });
}

int omega_tape_forward(omega_tape_t T, double t0, double dt, omega_state_t state,
                       omega_stream_t stream) {
  return guarded([&] {
    OMEGA_REQUIRE(T, "null tape");
    tape_forward(T, t0, dt, state, as_stream(stream));
  });
}

int omega_tape_reverse(omega_tape_t T, omega_state_t cot, double *param_grad,
                       omega_stream_t stream) {
  return guarded([&] {
    OMEGA_REQUIRE(T, "null tape");
    tape_reverse(T, cot, param_grad, as_stream(stream));
  });
}

int omega_tape_info(omega_tape_t T, int *n_snapshots, int64_t *forward_steps,
                    int64_t *snapshot_bytes) {
  return guarded([&] {
    OMEGA_REQUIRE(T, "null tape");
    if (n_snapshots) *n_snapshots = T->nsnap;
    if (forward_steps) *forward_steps = T->forward_steps;
    if (snapshot_bytes)
      *snapshot_bytes = (int64_t)T->nsnap * (int64_t)(T->n[0] + T->n[1] + T->n[2]) *
                        (int64_t)sizeof(Real);
  });
}

int64_t omega_revolve_cost(int64_t nsteps, int n_snapshots) {
  if (nsteps < 0 || n_snapshots < 1) return -1;
  return revolve_cost(nsteps, n_snapshots);
}

int omega_misfit(omega_model_t M, omega_state_t state, omega_state_t obs, const double *weights,
                 omega_state_t cot, double *J, omega_stream_t stream) {
  return guarded([&] {
    OMEGA_REQUIRE(M && weights && J, "null argument");
    check_state(M, state, "misfit state");
    check_state(M, cot, "misfit cotangent");
    OMEGA_REQUIRE(obs && obs->mesh == M->mesh, "misfit: observation state of another mesh");
    *J = misfit(M, state, obs, weights, cot, as_stream(stream));
  });
}

int omega_adjoint_run(omega_stepper_t St, double t0, double dt, int nsteps, int n_snapshots,
                      int64_t byte_cap, omega_state_t state, omega_state_t obs,
                      const double *weights, double *J, omega_state_t grad, double *param_grad,
                      omega_stream_t stream) {
  return guarded([&] {
    OMEGA_REQUIRE(St && weights && J, "null argument");
    omega_model_s *M = St->model;
    check_state(M, state, "adjoint_run state");
    check_state(M, grad, "adjoint_run gradient");
    OMEGA_REQUIRE(grad->n[2] == state->n[2], "adjoint_run: gradient and state tracer counts differ");
    OMEGA_REQUIRE(obs && obs->mesh == M->mesh, "adjoint_run: observation state of another mesh");
    hipStream_t s = as_stream(stream);
    // run on a copy so the caller's initial state is kept
    omega_state_s Y;
    Y.mesh = state->mesh;
    Y.ntr = state->ntr;
    for (int f = 0; f < 3; ++f) {
      Y.n[f] = state->n[f];
      Y.buf[0][f].alloc(state->n[f]);
    }
    const Real *src[3] = {state->ptr(0), state->ptr(1), state->ptr(2)};
    Real *dst[3] = {Y.ptr(0), Y.ptr(1), Y.ptr(2)};
    copy_fields(dst, src, state->n, s);
    omega_tape_s T;
    T.stepper = St;
    T.nsteps = nsteps;
    T.nsnap_req = std::max(1, n_snapshots);
    T.byte_cap = byte_cap;
    tape_forward(&T, t0, dt, &Y, s);
    *J = misfit(M, &Y, obs, weights, grad, s);
    tape_reverse(&T, grad, param_grad, s);
  });
}

} // extern "C"
