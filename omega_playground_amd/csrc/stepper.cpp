// Host-side orchestration: model coefficients, one tendency evaluation
// (diagnostics + fused tendency/update launch), the RK4 and low-storage RK
// steppers and their HIP-graph capture.
#include <algorithm>
#include <cmath>
#include <mutex>

#include "objects.hpp"

namespace omega_hip {

void model_build_coefficients(omega_model_s *M) {
  const DeviceMesh &dm = M->mesh->dm;
  const HostMesh &hm = dm.host;
  const int NC = hm.ncells, NE = hm.nedges, NV = hm.nvertices, ME = hm.max_edges;
  const size_t L = hm.nlayers, ntr = M->params.ntracers;
  std::vector<Real> f(NV, M->params.f0);
  M->f_v.upload(f);
  std::vector<Real> cv2(NE), cv4(NE);
  for (int e = 0; e < NE; ++e) {
    // mask*visc*scaling with mask = 1 (vel_diff_edge.hpp:83-84,
    // vel_hyperdiff_edge.hpp:234-235)
    cv2[e] = M->params.visc_del2 * dm.s2_h[e];
    cv4[e] = M->params.visc_del4 * dm.s4_h[e];
  }
  M->cv2.upload(cv2);
  M->cv4.upload(cv4);
  std::vector<Real> ct2((size_t)NC * ME, 0.0), ct4((size_t)NC * ME, 0.0);
  for (int i = 0; i < NC; ++i) {
    const int c = dm.cell_n2o[i];
    const double inv_area = 1. / hm.area_cell[c];
    for (int j = 0; j < dm.noc_h[i]; ++j) {
      const int eo = hm.edges_on_cell[(size_t)c * ME + j];
      const int en = dm.edge_o2n[eo];
      const double sign = hm.edge_sign_on_cell[(size_t)c * ME + j];
      // leading factors of tracer_diff_cell.hpp:82-84 and
      // tracer_hyperdiff_cell.hpp:152-154
      ct2[(size_t)i * ME + j] = M->params.eddy_diff2 * inv_area * hm.dv_edge[eo] * sign;
      ct4[(size_t)i * ME + j] =
          M->params.eddy_diff4 * inv_area * hm.dv_edge[eo] * sign * dm.s4_h[en];
    }
  }
  M->ct2.upload(ct2);
  M->ct2_h = ct2;
  M->ct4_h = ct4;
  M->ct4.upload(ct4);
  M->f_uniform = rows_equal(f, 1);
  M->mix_uniform = rows_equal(cv2, 1) && rows_equal(cv4, 1) && rows_equal(ct2, ME) &&
                   rows_equal(ct4, ME);
  M->rvort_v.alloc((size_t)NV * L);
  M->inv_thick_v.alloc((size_t)NV * L);
  M->ke_c.alloc((size_t)NC * L);
  M->div_c.alloc((size_t)NC * L);
  if (ntr > 0) M->trdel2_c.alloc(ntr * NC * L);
  M->partial.alloc((size_t)reduce_blocks(std::max(NC, NV)) + 1);
  M->result.alloc(4);
}

// Tables of the linear model: dv*sign per cell edge (the leading factors of
// shallow_water_hierarchical.cpp:696-697), dc and f per edge (:718-735).
void model_build_linear(omega_model_s *M) {
  const DeviceMesh &dm = M->mesh->dm;
  const HostMesh &hm = dm.host;
  const int NC = hm.ncells, NE = hm.nedges, ME = hm.max_edges;
  std::vector<Real> cdv((size_t)NC * ME, 0.0), dc(NE), fe(NE, M->params.f0);
  for (int i = 0; i < NC; ++i) {
    const int c = dm.cell_n2o[i];
    for (int j = 0; j < dm.noc_h[i]; ++j) {
      const int eo = hm.edges_on_cell[(size_t)c * ME + j];
      cdv[(size_t)i * ME + j] = hm.dv_edge[eo] * hm.edge_sign_on_cell[(size_t)c * ME + j];
    }
  }
  for (int i = 0; i < NE; ++i) dc[i] = hm.dc_edge[dm.edge_n2o[i]];
  M->lin_cdv.upload(cdv);
  M->lin_dc.upload(dc);
  M->lin_fe.upload(fe);
}

void ensure_vdel4(omega_model_s *M) {
  const HostMesh &hm = M->mesh->dm.host;
  const size_t L = hm.nlayers;
  if (!M->vdel2_e.p) {
    M->vdel2_e.alloc((size_t)hm.nedges * L);
    M->vdel2_rv_v.alloc((size_t)hm.nvertices * L);
    M->vdel2_div_c.alloc((size_t)hm.ncells * L);
  }
}

hipStream_t capture_stream() {
  static std::mutex mu;
  static std::map<int, hipStream_t> streams;
  int dev = 0;
  OMEGA_HIP_CHECK(hipGetDevice(&dev));
  std::lock_guard<std::mutex> g(mu);
  auto it = streams.find(dev);
  if (it != streams.end()) return it->second;
  hipStream_t cs = nullptr;
  OMEGA_HIP_CHECK(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
  streams.emplace(dev, cs);
  return cs;
}

// The opt-in bottom drag is evaluated by the general kernels only.
bool general_only(const omega_model_s *M) { return (M->terms & T_DRAG) != 0; }

// Diagnostics of a state (vertex vorticity and thickness, cell kinetic
// energy and divergence, tracer Laplacian, momentum del4 pre-pass) into the
// model's buffers; only what the enabled terms read.  Records the arrays as
// the auxiliary state's source.
bool diag_needed(const omega_model_s *M) {
  // diagnostics are needed by the vn tendency and by tracer del4 only
  return !M->params.disable_vn_tendency || ((M->terms & T_TDEL4) && M->params.ntracers > 0);
}

void run_diag_into(omega_model_s *M, const Real *h, const Real *vn, const Real *tr,
                   const DiagOut &d, hipStream_t s) {
  if (!diag_needed(M)) return;
  const MeshView mv = M->view();
  const int ntr = M->params.ntracers;
  const bool do_vn = !M->params.disable_vn_tendency;
  const int dterms = do_vn ? M->terms : (M->terms & T_TDEL4);
  Launch ld = M->lc; // the diagnostics launch may have its own block height
  if (ld.drows >= 0) ld.rows = ld.drows;
  if (hex_usable(mv, M->lc) && !general_only(M))
    launch_diag_hex(mv, ld, dterms, ntr, h, vn, tr, d, s);
  else launch_diag(mv, ld, M->coef(), dterms, ntr, h, vn, tr, d, s);
  if (do_vn && (M->terms & T_VDEL4)) {
    if (hex_vdel4(mv, M->lc) && !general_only(M)) launch_vdel4_hex(mv, M->lc, d, s);
    else launch_vdel4(mv, M->lc, d, s);
  }
}

void run_diag(omega_model_s *M, const Real *h, const Real *vn, const Real *tr, hipStream_t s) {
  M->aux_h = h;
  M->aux_vn = vn;
  M->aux_tr = tr;
  if (!diag_needed(M)) return;
  if (M->terms & T_VDEL4) ensure_vdel4(M);
  run_diag_into(M, h, vn, tr, M->diag(), s);
}

void run_stage_into(omega_model_s *M, StageArgs a, hipStream_t s, int fields,
                    const DiagOut *dx) {
  const MeshView mv = M->view();
  const ModelCoef mc = M->coef();
  a.terms = M->terms;
  a.do_h = !M->params.disable_h_tendency && (fields & 1);
  a.do_vn = !M->params.disable_vn_tendency && (fields & 2);
  a.ntr = (fields & 4) ? M->params.ntracers : 0;
  if (M->linear) {
    launch_tend_linear(mv, M->lc, M->lin_coef(), a, s);
    return;
  }
  if (dx) {
    // caller-provided diagnostics buffers
    run_diag_into(M, a.h.in, a.vn.in, a.tr.in, *dx, s);
    if (hex_usable(mv, M->lc) && !general_only(M)) launch_tend_hex(mv, M->lc, mc, a, *dx, s);
    else launch_tend(mv, M->lc, mc, a, *dx, s);
    return;
  }
  run_stage(M, a, s, fields, true);
}

// One tendency evaluation (+ optional fused stage update).  with_diag =
// false reads the diagnostics an earlier run_diag left in the model's
// buffers (the per-field tendency calls after compute_aux).
bool stage_writes_pre(omega_model_s *M) {
  const MeshView mv = M->view();
  return (M->forms & 8) && !M->linear && !general_only(M) &&
         !M->params.disable_vn_tendency && hex_usable(mv, M->lc) &&
         hex_writes_pre(mv, M->lc, M->coef());
}

void run_stage(omega_model_s *M, StageArgs a, hipStream_t s, int fields, bool with_diag,
               const PreOut *pre) {
  const MeshView mv = M->view();
  const ModelCoef mc = M->coef();
  a.terms = M->terms;
  a.do_h = !M->params.disable_h_tendency && (fields & 1);
  a.do_vn = !M->params.disable_vn_tendency && (fields & 2);
  a.ntr = (fields & 4) ? M->params.ntracers : 0;
  if (M->linear) {
    launch_tend_linear(mv, M->lc, M->lin_coef(), a, s);
    return;
  }
  if (with_diag) run_diag(M, a.h.in, a.vn.in, a.tr.in, s);
  const DiagOut d = M->diag();
  OMEGA_REQUIRE(!pre || (fields == 7 && stage_writes_pre(M)),
                "stage: this model's stage kernel does not write the VJP rows");
  if (general_only(M)) {
    launch_tend(mv, M->lc, mc, a, d, s);
    return;
  }
  if (hex_usable(mv, M->lc)) launch_tend_hex(mv, M->lc, mc, a, d, s, pre);
  else launch_tend(mv, M->lc, mc, a, d, s);
}

FieldIO fio(const Real *in, Real *tend, Real *acc, const Real *yold, Real *out) {
  return FieldIO{in, tend, acc, yold, out};
}

void call_hook(omega_model_s *M, Real *ht, Real *vt, const Real *h, const Real *vn, double t,
               hipStream_t s) {
  if (!M->add_fn) return;
  const int rc = M->add_fn(M->add_user, ht, vt, h, vn, t, s);
  if (rc != 0) throw Error(OMEGA_ERR_CALLBACK, "additional tendency hook returned " + std::to_string(rc));
}

void check_state(omega_model_s *M, omega_state_s *S, const char *what) {
  OMEGA_REQUIRE(S != nullptr, std::string(what) + ": null state");
  OMEGA_REQUIRE(S->mesh == M->mesh, std::string(what) + ": state belongs to another mesh");
  OMEGA_REQUIRE(S->ntr >= M->params.ntracers,
                std::string(what) + ": state has fewer tracers than the model");
}

void ensure3(DBuf<Real> (&b)[3], omega_state_s *S) {
  for (int f = 0; f < 3; ++f)
    if (b[f].n != S->n[f]) b[f].alloc(S->n[f]);
}

// compute_tendency (shallow_water_model.cpp:313-347), writing tend buffers.
void compute_tendency_ptrs(omega_model_s *M, Real *ht, Real *vt, Real *tt, const Real *h,
                           const Real *vn, const Real *tr, double t, int add_mode,
                           hipStream_t s, int fields) {
  StageArgs a{};
  a.epi = EPI_STORE;
  a.base_from_tend = add_mode == OMEGA_ADD_INCREMENT;
  a.a = 1.0;
  a.h = fio(h, ht, nullptr, nullptr, nullptr);
  a.vn = fio(vn, vt, nullptr, nullptr, nullptr);
  a.tr = fio(tr, tt, nullptr, nullptr, nullptr);
  run_stage(M, a, s, fields);
  // the hook runs after the built-in terms of a full evaluation only
  // (shallow_water_model.cpp:344-345)
  if (fields == 7) call_hook(M, ht, vt, h, vn, t, s);
}

// ---------------------------------------------------------------------------
// Steppers
// ---------------------------------------------------------------------------

void stepper_coefficients(int kind, int nstages, std::vector<Real> &rka, std::vector<Real> &rkb,
                          std::vector<Real> &rkc) {
  if (kind == OMEGA_STEPPER_RK4) {
    // rk4.cpp:14-25
    rka = {1. / 2, 1. / 2, 1};
    rkb = {1. / 6, 1. / 3, 1. / 3, 1. / 6};
    rkc = {1. / 2, 1. / 2, 1};
    return;
  }
  rka.assign(nstages, 0.0);
  rkb.assign(nstages, 0.0);
  rkc.assign(nstages, 0.0);
  if (nstages == 5) {
    // Carpenter-Kennedy (5,4) 2N-storage coefficients (lsrk.cpp:16-28)
    rka = {0., -567301805773. / 1357537059087., -2404267990393. / 2016746695238.,
           -3550918686646. / 2091501179385., -1275806237668. / 842570457699.};
    rkb = {1432997174477. / 9575080441755., 5161836677717. / 13612068292357.,
           1720146321549. / 2090206949498., 3134564353537. / 4481467310338.,
           2277821191437. / 14882151754819.};
    rkc = {0., 1432997174477. / 9575080441755., 2526269341429. / 6820363962896.,
           2006345519317. / 3224310063776., 2802321613138. / 2924317926251.};
  } else if (nstages == 1) { // forward Euler (lsrk.cpp:29-33)
    rka = {0};
    rkb = {1};
    rkc = {0};
  }
}

// RK4 with a fused tendency + update per stage (rk4.cpp:28-104 reorganised:
// the accumulator lives in acc until the last stage writes the state).
void rk4_fused_step(omega_stepper_s *St, omega_state_s *S, double dt, hipStream_t s) {
  omega_model_s *M = St->model;
  ensure3(St->acc, S);
  ensure3(St->p1, S);
  ensure3(St->p2, S);
  Real *Y[3] = {S->ptr(0), S->ptr(1), S->ptr(2)};
  Real *A[3] = {St->acc[0].p, St->acc[1].p, St->acc[2].p};
  Real *P1[3] = {St->p1[0].p, St->p1[1].p, St->p1[2].p};
  Real *P2[3] = {St->p2[0].p, St->p2[1].p, St->p2[2].p};
  for (int stage = 0; stage < 4; ++stage) {
    StageArgs a{};
    a.base_from_tend = 0;
    a.db = dt * St->rkb[stage];
    a.da = stage < 3 ? dt * St->rka[stage] : 0.0;
    Real **in = stage == 0 ? Y : (stage == 2 ? P2 : P1);
    Real **out = stage == 0 ? P1 : (stage == 1 ? P2 : (stage == 2 ? P1 : Y));
    a.epi = stage == 0 ? EPI_RK4_FIRST : (stage == 3 ? EPI_RK4_LAST : EPI_RK4_MID);
    a.h = fio(in[0], nullptr, A[0], Y[0], out[0]);
    a.vn = fio(in[1], nullptr, A[1], Y[1], out[1]);
    a.tr = fio(in[2], nullptr, A[2], Y[2], out[2]);
    run_stage(M, a, s);
  }
}

// RK4 exactly as rk4.cpp:28-104 (used when an additional-tendency hook is set).
void rk4_unfused_step(omega_stepper_s *St, omega_state_s *S, double t, double dt, hipStream_t s) {
  omega_model_s *M = St->model;
  ensure3(St->tend, S);
  ensure3(St->provis, S);
  ensure3(St->old, S);
  for (int f = 0; f < 3; ++f)
    if (S->n[f])
      OMEGA_HIP_CHECK(hipMemcpyAsync(St->old[f].p, S->ptr(f), S->n[f] * sizeof(Real),
                                     hipMemcpyDeviceToDevice, s));
  compute_tendency_ptrs(M, St->tend[0].p, St->tend[1].p, St->tend[2].p, S->ptr(0), S->ptr(1),
                        S->ptr(2), t, OMEGA_ADD_REPLACE, s);
  for (int stage = 0; stage < 4; ++stage) {
    const double db = dt * St->rkb[stage];
    for (int f = 0; f < 3; ++f) launch_axpy(S->ptr(f), db, St->tend[f].p, S->n[f], s);
    if (stage < 3) {
      const double stagetime = t + St->rkc[stage] * dt;
      const double da = dt * St->rka[stage];
      for (int f = 0; f < 3; ++f)
        launch_xpay(St->provis[f].p, St->old[f].p, da, St->tend[f].p, S->n[f], s);
      compute_tendency_ptrs(M, St->tend[0].p, St->tend[1].p, St->tend[2].p, St->provis[0].p,
                            St->provis[1].p, St->provis[2].p, stagetime, OMEGA_ADD_REPLACE, s);
    }
  }
}

// Tracer rows of a state beyond the model's tracer count are not stepped
// (the reference's loops run over the model's ntracers, lsrk.cpp:47-87);
// steppers that alternate the state between two buffers copy those rows to
// the other buffer first, so they read the same from either.
void carry_extra_tracers(omega_model_s *M, omega_state_s *S, hipStream_t s) {
  const size_t per = S->ntr > 0 ? S->n[2] / S->ntr : 0;
  const size_t first = (size_t)M->params.ntracers * per;
  if (S->ntr <= M->params.ntracers || !per) return;
  OMEGA_HIP_CHECK(hipMemcpyAsync(S->alt(2) + first, S->ptr(2) + first,
                                 (S->n[2] - first) * sizeof(Real), hipMemcpyDeviceToDevice, s));
}

// Low-storage RK with the state alternating between two buffers per stage
// (lsrk.cpp:36-89 fused: tend = a*tend + F(y); y' = y + dt*b*tend).
void lsrk_fused_step(omega_stepper_s *St, omega_state_s *S, double dt, hipStream_t s) {
  omega_model_s *M = St->model;
  ensure3(St->tend, S);
  carry_extra_tracers(M, S, s);
  for (int stage = 0; stage < St->nstages; ++stage) {
    StageArgs a{};
    a.epi = EPI_LSRK;
    a.a = St->rka[stage];
    a.base_from_tend = St->rka[stage] != 0.0;
    a.db = dt * St->rkb[stage];
    Real *in[3] = {S->ptr(0), S->ptr(1), S->ptr(2)};
    Real *out[3] = {S->alt(0), S->alt(1), S->alt(2)};
    a.h = fio(in[0], St->tend[0].p, nullptr, nullptr, out[0]);
    a.vn = fio(in[1], St->tend[1].p, nullptr, nullptr, out[1]);
    a.tr = fio(in[2], St->tend[2].p, nullptr, nullptr, out[2]);
    run_stage(M, a, s);
    S->cur = 1 - S->cur;
  }
}

// LSRK exactly as lsrk.cpp:36-89 (used when an additional-tendency hook is set).
void lsrk_unfused_step(omega_stepper_s *St, omega_state_s *S, double t, double dt, hipStream_t s) {
  omega_model_s *M = St->model;
  ensure3(St->tend, S);
  for (int stage = 0; stage < St->nstages; ++stage) {
    for (int f = 0; f < 3; ++f) launch_scale(St->tend[f].p, St->rka[stage], S->n[f], s);
    const double stagetime = t + St->rkc[stage] * dt;
    compute_tendency_ptrs(M, St->tend[0].p, St->tend[1].p, St->tend[2].p, S->ptr(0), S->ptr(1),
                          S->ptr(2), stagetime, OMEGA_ADD_INCREMENT, s);
    const double db = dt * St->rkb[stage];
    for (int f = 0; f < 3; ++f) launch_axpy(S->ptr(f), db, St->tend[f].p, S->n[f], s);
  }
}

void step_once(omega_stepper_s *St, omega_state_s *S, double t, double dt, hipStream_t s) {
  const bool fused = St->model->add_fn == nullptr;
  if (St->kind == OMEGA_STEPPER_RK4) {
    if (fused) rk4_fused_step(St, S, dt, s);
    else rk4_unfused_step(St, S, t, dt, s);
  } else {
    if (fused) lsrk_fused_step(St, S, dt, s);
    else lsrk_unfused_step(St, S, t, dt, s);
  }
}

void do_step_maybe_graph(omega_stepper_s *St, omega_state_s *S, double t, double dt,
                         hipStream_t s) {
  const bool graphable = St->use_graph && St->model->add_fn == nullptr && !stream_capturing(s);
  if (!graphable) {
    step_once(St, S, t, dt, s);
    return;
  }
  if (St->kind == OMEGA_STEPPER_LSRK) {
    // make sure both buffers exist before capture
    for (int f = 0; f < 3; ++f) (void)S->alt(f);
  }
  // every buffer the step touches is allocated before the key is formed
  if (St->kind == OMEGA_STEPPER_RK4) {
    ensure3(St->acc, S);
    ensure3(St->p1, S);
    ensure3(St->p2, S);
  } else {
    ensure3(St->tend, S);
  }
  omega_model_s *M = St->model;
  if (M->terms & T_VDEL4) ensure_vdel4(M);
  GraphKey key{dt, S->cur, M->gen, {}};
  for (int f = 0; f < 3; ++f)
    key.bufs.insert(key.bufs.end(), {S->buf[0][f].p, S->buf[1][f].p, St->acc[f].p,
                                     St->p1[f].p, St->p2[f].p, St->tend[f].p});
  key.bufs.insert(key.bufs.end(), {M->vdel2_e.p, M->trdel2_c.p, M->f_v.p});
  const DiagOut dk = M->diag();
  key.bufs.insert(key.bufs.end(), {dk.rvort_v, dk.inv_thick_v, dk.ke_c, dk.div_c, dk.trdel2_c,
                                   dk.vdel2_e, dk.vdel2_rv_v, dk.vdel2_div_c});
  auto it = St->graphs.find(key);
  if (it == St->graphs.end()) {
    // bound the cache: states created and destroyed in a loop (the adjoint's
    // work states) would otherwise leave a graph each behind
    if (St->graphs.size() >= 16) destroy_graphs(St->graphs);
    // capture on the process's capture stream (the null stream cannot be
    // captured); the graph is then launched on the caller's stream
    hipStream_t cs = capture_stream();
    const int cur0 = S->cur;
    hipGraph_t g;
    OMEGA_HIP_CHECK(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
    try {
      step_once(St, S, t, dt, cs);
    } catch (...) {
      hipGraph_t dummy;
      (void)hipStreamEndCapture(cs, &dummy);
      S->cur = cur0;
      throw;
    }
    OMEGA_HIP_CHECK(hipStreamEndCapture(cs, &g));
    hipGraphExec_t ex;
    OMEGA_HIP_CHECK(hipGraphInstantiate(&ex, g, nullptr, nullptr, 0));
    OMEGA_HIP_CHECK(hipGraphDestroy(g));
    S->cur = cur0; // capture recorded the flips; replay applies them below
    it = St->graphs.emplace(key, ex).first;
  }
  OMEGA_HIP_CHECK(hipGraphLaunch(it->second, s));
  if (St->kind == OMEGA_STEPPER_LSRK && (St->nstages % 2 == 1)) S->cur = 1 - S->cur;
}


} // namespace omega_hip
