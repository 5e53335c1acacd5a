// Library objects behind the opaque C ABI handles, and the host-side
// orchestration shared by the ABI, the steppers and the adjoint.
#pragma once

#include <array>
#include <map>
#include <string>

#include "internal.hpp"

using omega_hip::DBuf;
using omega_hip::DeviceMesh;
using omega_hip::Launch;
using omega_hip::ModelCoef;
using omega_hip::DiagOut;
using omega_hip::Real;

// ---------------------------------------------------------------------------
// Opaque objects
// ---------------------------------------------------------------------------
#include "adjoint.hpp"

namespace omega_hip {
// Destroys cached graph executables after the device has finished every
// launch of them (a launch may be in flight on any stream).
template <class Map> void destroy_graphs(Map &graphs) {
  if (graphs.empty()) return;
  (void)hipDeviceSynchronize();
  for (auto &kv : graphs) (void)hipGraphExecDestroy(kv.second);
  graphs.clear();
}
} // namespace omega_hip

struct omega_mesh_s {
  DeviceMesh dm;
  DBuf<Real> staging;             // device scratch for omega_add_host_field
  omega_hip::AdjLayout adj;       // transposed tables, built on first adjoint use
};

struct omega_state_s {
  omega_mesh_s *mesh = nullptr;
  int ntr = 0;
  size_t n[3] = {0, 0, 0}; // element counts of h, vn, tr
  DBuf<Real> buf[2][3];    // [buffer][field]
  int cur = 0;             // buffer holding the state
  Real *ptr(int f) { return buf[cur][f].p; }
  Real *alt(int f) {
    if (buf[1 - cur][f].n != n[f]) buf[1 - cur][f].alloc(n[f]);
    return buf[1 - cur][f].p;
  }
};

struct omega_model_s {
  omega_mesh_s *mesh = nullptr;
  omega_params params{};
  int terms = 0;
  Launch lc;
  // configuration generation: bumped by every setter that changes what a
  // step's kernels are or read (terms, launch shape, kernel family, forms,
  // f_vertex, hook), so captured graphs and ensemble tables built for an
  // older configuration are never replayed
  uint64_t gen = 0;
  void touch() { ++gen; }
  // arrays the auxiliary state (diagnostics buffers) was last computed from
  // (ShallowWaterAuxiliaryState::compute, shallow_water_auxstate.cpp:12-50);
  // the per-field tendency calls read the thickness and tracers through it
  const Real *aux_h = nullptr, *aux_vn = nullptr, *aux_tr = nullptr;
  DBuf<Real> f_v, cv2, cv4, ct2, ct4;
  DBuf<Real> rvort_v, inv_thick_v, ke_c, div_c, trdel2_c, vdel2_e, vdel2_rv_v, vdel2_div_c;
  DBuf<Real> partial, result;
  omega_additional_tendency_fn add_fn = nullptr;
  void *add_user = nullptr;
  // LinearShallowWaterModel (shallow_water_hierarchical.cpp:676-773)
  bool linear = false;
  Real h0 = 0;
  DBuf<Real> lin_cdv, lin_dc, lin_fe;
  // host copies of the tracer mixing coefficients (for the adjoint tables)
  std::vector<Real> ct2_h, ct4_h;
  // adjoint: transposed tracer coefficients and cotangent scratch, built on
  // first use
  DBuf<Real> ect2, ect4;
  bool ect_uniform = false;
  DBuf<Real> a_se, a_fu, a_px, a_xa;
  DBuf<Real> a_sbar, a_zbar, a_tbar, a_kbar, a_dbar, a_gbar, a_fbar, a_d2bar, a_z2bar, a_e2bar,
      a_partial, a_pgrad, a_mpart;
  int a_nparts = 0;

  omega_hip::LinCoef lin_coef() const {
    return omega_hip::LinCoef{lin_cdv.p, mesh->dm.area_cell.p, lin_dc.p, lin_fe.p, h0, params.grav};
  }
  // f_v, and cv2/cv4/ct2/ct4 at every stencil slot, bit-identical across
  // the mesh (ModelCoef::uniform)
  bool f_uniform = false, mix_uniform = false;
  // kernel forms the model may use (bit 0 shared coefficients, bit 1
  // 32-bit offsets, bit 2 compiled-in term sets, bit 3 VJP rows written by
  // a step adjoint's stage recomputation); clearing bits forces the general
  // forms (tests, tuning)
  int forms = 15;
  // the mesh view this model's kernels get
  omega_hip::MeshView view() const {
    omega_hip::MeshView v = mesh->dm.view();
    if (!(forms & 1)) v.coef_uni = 0;
    if (!(forms & 2)) v.off32 = 0;
    if (!(forms & 4)) v.term_sets = 0;
    return v;
  }
  ModelCoef coef() const {
    return ModelCoef{params.grav, f_v.p, cv2.p, cv4.p, ct2.p, ct4.p,
                     f_uniform && mix_uniform ? 1 : 0, params.drag_coeff};
  }
  // step adjoint: diagnostics of the recomputed stage inputs, one set per
  // stage input (stage 0 uses the buffers above), kept for the reverse
  // sweep; diag_slot selects the set run_diag writes and the kernels read
  struct DiagSet {
    DBuf<Real> rvort_v, inv_thick_v, ke_c, div_c, trdel2_c, vdel2_e, vdel2_rv_v, vdel2_div_c;
  };
  DiagSet diag_sets[4];
  int diag_slot = 0;
  // step adjoint: the tendency VJP's rows at stage inputs 0..3, written by
  // the stage recomputation where its kernel can (stage_writes_pre)
  struct PreSet {
    DBuf<Real> se, fu, own;
  };
  PreSet pre_sets[4];
  DiagOut diag() {
    if (diag_slot > 0) {
      DiagSet &d = diag_sets[diag_slot - 1];
      return DiagOut{d.rvort_v.p, d.inv_thick_v.p, d.ke_c.p, d.div_c.p, d.trdel2_c.p,
                     d.vdel2_e.p, d.vdel2_rv_v.p, d.vdel2_div_c.p};
    }
    return DiagOut{rvort_v.p, inv_thick_v.p, ke_c.p, div_c.p, trdel2_c.p,
                   vdel2_e.p, vdel2_rv_v.p, vdel2_div_c.p};
  }
};

// A captured step is valid only for the exact device buffers and model
// configuration it recorded: the key holds dt, the buffer parity, the
// model's configuration generation and every buffer address the step
// touches, so a reallocated buffer or a changed setting never replays a
// stale graph.
struct GraphKey {
  double dt;
  int cur;
  uint64_t gen;
  std::vector<const void *> bufs;
  bool operator<(const GraphKey &o) const {
    if (dt != o.dt) return dt < o.dt;
    if (cur != o.cur) return cur < o.cur;
    if (gen != o.gen) return gen < o.gen;
    return bufs < o.bufs;
  }
};

struct omega_stepper_s {
  omega_model_s *model = nullptr;
  int kind = OMEGA_STEPPER_RK4;
  int nstages = 4;
  std::vector<Real> rka, rkb, rkc;
  // fused RK4: accumulator and two provisional states; unfused RK4: tend,
  // provisional and old states; LSRK: tend
  DBuf<Real> acc[3], p1[3], p2[3], tend[3], provis[3], old[3];
  bool use_graph = false;
  std::map<GraphKey, hipGraphExec_t> graphs;
  // step adjoint: recomputed stage states (RK4 stage inputs 2..4, LSRK stage
  // inputs 2..S), their tendency, the incoming cotangent and stage cotangents
  DBuf<Real> adj_y[4][3], adj_t[3], adj_gy[3], adj_pb[3], adj_kb[3];
  ~omega_stepper_s() { omega_hip::destroy_graphs(graphs); }
};

// ---------------------------------------------------------------------------

namespace omega_hip {

extern thread_local std::string g_last_error;

// Runs f, converting exceptions to status codes and the thread's last error.
template <class F> int guarded(F &&f) {
  try {
    f();
    return OMEGA_OK;
  } catch (const Error &e) {
    g_last_error = e.what();
    return e.code;
  } catch (const std::exception &e) {
    g_last_error = e.what();
    return OMEGA_ERR_STATE;
  } catch (...) {
    g_last_error = "unknown error";
    return OMEGA_ERR_STATE;
  }
}

inline hipStream_t as_stream(omega_stream_t s) { return reinterpret_cast<hipStream_t>(s); }
bool stream_capturing(hipStream_t s);
// The stream steps are captured on: one per device for the whole process
// (the null stream cannot be captured), created on first use and kept.
hipStream_t capture_stream();

void model_build_coefficients(omega_model_s *M);
void ensure_vdel4(omega_model_s *M);
void model_build_linear(omega_model_s *M);
void run_diag(omega_model_s *M, const Real *h, const Real *vn, const Real *tr, hipStream_t s);
// fields: bit 0 h, bit 1 vn, bit 2 tracers (all by default); with_diag =
// false reuses the diagnostics of the last run_diag
// pre: also write the tendency VJP's rows at the stage input (only where
// stage_writes_pre says the model's stage kernel can)
void run_stage(omega_model_s *M, StageArgs a, hipStream_t s, int fields = 7,
               bool with_diag = true, const omega_hip::PreOut *pre = nullptr);
bool stage_writes_pre(omega_model_s *M);
// true when the model's terms need the general kernels (bottom drag)
bool general_only(const omega_model_s *M);
// true when the enabled terms read diagnostics (vn tendency, tracer del4)
bool diag_needed(const omega_model_s *M);
// diagnostics into caller-provided buffers (no model state is touched)
void run_diag_into(omega_model_s *M, const Real *h, const Real *vn, const Real *tr,
                   const DiagOut &d, hipStream_t s);
// run_stage with the diagnostics in dx (the default kernels), or, with
// dx == nullptr, exactly run_stage
void run_stage_into(omega_model_s *M, StageArgs a, hipStream_t s, int fields,
                    const DiagOut *dx);
void call_hook(omega_model_s *M, Real *ht, Real *vt, const Real *h, const Real *vn, double t,
               hipStream_t s);
void carry_extra_tracers(omega_model_s *M, omega_state_s *S, hipStream_t s);
FieldIO fio(const Real *in, Real *tend, Real *acc, const Real *yold, Real *out);
void check_state(omega_model_s *M, omega_state_s *S, const char *what);
void ensure3(DBuf<Real> (&b)[3], omega_state_s *S);
void compute_tendency_ptrs(omega_model_s *M, Real *ht, Real *vt, Real *tt, const Real *h,
                           const Real *vn, const Real *tr, double t, int add_mode,
                           hipStream_t s, int fields = 7);
// RK4 tableau (rk4.cpp:14-25) or the low-storage coefficients (lsrk.cpp:16-33)
void stepper_coefficients(int kind, int nstages, std::vector<Real> &rka, std::vector<Real> &rkb,
                          std::vector<Real> &rkc);
void step_once(omega_stepper_s *St, omega_state_s *S, double t, double dt, hipStream_t s);
void do_step_maybe_graph(omega_stepper_s *St, omega_state_s *S, double t, double dt,
                         hipStream_t s);

} // namespace omega_hip
