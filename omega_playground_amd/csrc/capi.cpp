// C ABI of the shallow-water library: object lifetimes, host copies,
// diagnostics and timing.  Stage orchestration lives in stepper.cpp.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <fstream>
#include <mutex>
#include <sstream>

#include "objects.hpp"

using namespace omega_hip;

// ---------------------------------------------------------------------------
// Error plumbing
// ---------------------------------------------------------------------------
namespace omega_hip {
thread_local std::string g_last_error;
} // namespace omega_hip

namespace {

// ---------------------------------------------------------------------------
// Timing
// ---------------------------------------------------------------------------
struct TimingTable {
  std::mutex mu;
  bool enabled = false;
  std::vector<std::string> names;
  std::vector<double> total_ms;
  std::vector<int64_t> count;
  struct Pending {
    int slot;
    hipEvent_t a, b;
  };
  std::vector<Pending> pending;
  std::vector<hipEvent_t> pool;
  int slot_of(const char *name) {
    for (size_t i = 0; i < names.size(); ++i)
      if (names[i] == name) return (int)i;
    names.emplace_back(name);
    total_ms.push_back(0);
    count.push_back(0);
    return (int)names.size() - 1;
  }
  hipEvent_t get() {
    if (!pool.empty()) {
      hipEvent_t e = pool.back();
      pool.pop_back();
      return e;
    }
    hipEvent_t e;
    OMEGA_HIP_CHECK(hipEventCreate(&e));
    return e;
  }
  void drain() {
    for (auto &p : pending) {
      OMEGA_HIP_CHECK(hipEventSynchronize(p.b));
      float ms = 0;
      OMEGA_HIP_CHECK(hipEventElapsedTime(&ms, p.a, p.b));
      total_ms[p.slot] += ms;
      count[p.slot] += 1;
      pool.push_back(p.a);
      pool.push_back(p.b);
    }
    pending.clear();
  }
};
TimingTable &timing() {
  static TimingTable t;
  return t;
}

} // namespace

namespace omega_hip {
bool stream_capturing(hipStream_t s) {
  hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(s, &st) != hipSuccess) return false;
  return st == hipStreamCaptureStatusActive;
}
} // namespace omega_hip

namespace omega_hip {
TimerScope::TimerScope(const char *name, hipStream_t st) : slot(-1), s(st) {
  TimingTable &T = timing();
  if (!T.enabled || stream_capturing(st)) return;
  std::lock_guard<std::mutex> g(T.mu);
  slot = T.slot_of(name);
  hipEvent_t a = T.get(), b = T.get();
  OMEGA_HIP_CHECK(hipEventRecord(a, st));
  T.pending.push_back({slot, a, b});
}
TimerScope::~TimerScope() {
  if (slot < 0) return;
  TimingTable &T = timing();
  std::lock_guard<std::mutex> g(T.mu);
  for (auto it = T.pending.rbegin(); it != T.pending.rend(); ++it)
    if (it->slot == slot) {
      (void)hipEventRecord(it->b, s);
      break;
    }
  if (T.pending.size() > 4096) {
    try {
      T.drain();
    } catch (...) {
    }
  }
}
} // namespace omega_hip

// ---------------------------------------------------------------------------
// Model helpers
// ---------------------------------------------------------------------------

// ---------------------------------------------------------------------------
// ABI
// ---------------------------------------------------------------------------
extern "C" {

const char *omega_last_error(void) { return g_last_error.c_str(); }
const char *omega_version(void) { return "omega-mi355x 0.1 (gfx950, fp64)"; }

int omega_mesh_create_planar_hex(int nx, int ny, double dc, int nlayers, int ordering,
                                 omega_mesh_t *out) {
  return guarded([&] {
    OMEGA_REQUIRE(out, "out is null");
    auto m = std::make_unique<omega_mesh_s>();
    build_planar_hex(m->dm.host, nx, ny, dc, nlayers);
    m->dm.build_device(ordering);
    *out = m.release();
  });
}

int omega_mesh_create_mpas(const omega_mpas_mesh_desc *d, int ordering, omega_mesh_t *out) {
  return guarded([&] {
    OMEGA_REQUIRE(d && out, "null argument");
    OMEGA_REQUIRE(d->ncells > 0 && d->nedges > 0 && d->nvertices > 0 && d->nlayers > 0,
                  "mesh sizes must be positive");
    OMEGA_REQUIRE(d->max_edges >= 3 && d->max_edges2 >= 1, "bad row lengths");
    OMEGA_REQUIRE(d->nedges_on_cell && d->edges_on_cell && d->vertices_on_cell &&
                      d->nedges_on_edge && d->edges_on_edge && d->cells_on_edge &&
                      d->vertices_on_edge && d->edges_on_vertex && d->cells_on_vertex &&
                      d->area_cell && d->dc_edge && d->dv_edge && d->weights_on_edge &&
                      d->area_triangle && d->kiteareas_on_vertex,
                  "required mesh array missing");
    auto m = std::make_unique<omega_mesh_s>();
    HostMesh &h = m->dm.host;
    h.ncells = d->ncells;
    h.nedges = d->nedges;
    h.nvertices = d->nvertices;
    h.nlayers = d->nlayers;
    h.max_edges = d->max_edges;
    h.max_edges2 = d->max_edges2;
    h.resize();
    const size_t NC = h.ncells, NE = h.nedges, NV = h.nvertices;
    auto cpi = [](std::vector<int> &dst, const int *src, size_t n) { dst.assign(src, src + n); };
    auto cpd = [](std::vector<double> &dst, const double *src, size_t n) {
      if (src) dst.assign(src, src + n);
    };
    cpi(h.nedges_on_cell, d->nedges_on_cell, NC);
    cpi(h.edges_on_cell, d->edges_on_cell, NC * h.max_edges);
    cpi(h.vertices_on_cell, d->vertices_on_cell, NC * h.max_edges);
    cpi(h.nedges_on_edge, d->nedges_on_edge, NE);
    cpi(h.edges_on_edge, d->edges_on_edge, NE * h.max_edges2);
    cpi(h.cells_on_edge, d->cells_on_edge, NE * 2);
    cpi(h.vertices_on_edge, d->vertices_on_edge, NE * 2);
    cpi(h.edges_on_vertex, d->edges_on_vertex, NV * 3);
    cpi(h.cells_on_vertex, d->cells_on_vertex, NV * 3);
    for (size_t c = 0; c < NC; ++c) {
      OMEGA_REQUIRE(h.nedges_on_cell[c] >= 1 && h.nedges_on_cell[c] <= h.max_edges,
                    "nedges_on_cell out of range");
    }
    for (size_t e = 0; e < NE; ++e) {
      OMEGA_REQUIRE(h.nedges_on_edge[e] >= 0 && h.nedges_on_edge[e] <= h.max_edges2,
                    "nedges_on_edge out of range");
      for (int s = 0; s < 2; ++s) {
        OMEGA_REQUIRE(h.cells_on_edge[2 * e + s] >= 0 && h.cells_on_edge[2 * e + s] < (int)NC,
                      "cells_on_edge out of range");
        OMEGA_REQUIRE(h.vertices_on_edge[2 * e + s] >= 0 &&
                          h.vertices_on_edge[2 * e + s] < (int)NV,
                      "vertices_on_edge out of range");
      }
      for (int j = 0; j < h.nedges_on_edge[e]; ++j)
        OMEGA_REQUIRE(h.edges_on_edge[e * h.max_edges2 + j] >= 0 &&
                          h.edges_on_edge[e * h.max_edges2 + j] < (int)NE,
                      "edges_on_edge out of range");
    }
    for (size_t v = 0; v < NV * 3; ++v) {
      OMEGA_REQUIRE(h.edges_on_vertex[v] >= 0 && h.edges_on_vertex[v] < (int)NE,
                    "edges_on_vertex out of range");
      OMEGA_REQUIRE(h.cells_on_vertex[v] >= 0 && h.cells_on_vertex[v] < (int)NC,
                    "cells_on_vertex out of range");
    }
    cpd(h.area_cell, d->area_cell, NC);
    cpd(h.dc_edge, d->dc_edge, NE);
    cpd(h.dv_edge, d->dv_edge, NE);
    cpd(h.weights_on_edge, d->weights_on_edge, NE * h.max_edges2);
    cpd(h.area_triangle, d->area_triangle, NV);
    cpd(h.kiteareas_on_vertex, d->kiteareas_on_vertex, NV * 3);
    std::fill(h.mesh_scaling_del2.begin(), h.mesh_scaling_del2.end(), 1.0);
    std::fill(h.mesh_scaling_del4.begin(), h.mesh_scaling_del4.end(), 1.0);
    cpd(h.mesh_scaling_del2, d->mesh_scaling_del2, NE);
    cpd(h.mesh_scaling_del4, d->mesh_scaling_del4, NE);
    std::fill(h.max_level_edge_top.begin(), h.max_level_edge_top.end(), h.nlayers);
    if (d->max_level_edge_top) cpi(h.max_level_edge_top, d->max_level_edge_top, NE);
    cpd(h.x_cell, d->x_cell, NC);
    cpd(h.y_cell, d->y_cell, NC);
    cpd(h.x_edge, d->x_edge, NE);
    cpd(h.y_edge, d->y_edge, NE);
    cpd(h.angle_edge, d->angle_edge, NE);
    cpd(h.x_vertex, d->x_vertex, NV);
    cpd(h.y_vertex, d->y_vertex, NV);
    h.derive_signs_and_levels();
    m->dm.build_device(ordering);
    *out = m.release();
  });
}

int omega_mesh_destroy(omega_mesh_t mesh) {
  return guarded([&] { delete mesh; });
}

int omega_mesh_info(omega_mesh_t mesh, int *counts, double *reals) {
  return guarded([&] {
    OMEGA_REQUIRE(mesh, "null mesh");
    const HostMesh &h = mesh->dm.host;
    if (counts) {
      counts[0] = h.ncells; counts[1] = h.nedges; counts[2] = h.nvertices;
      counts[3] = h.nlayers; counts[4] = h.nx; counts[5] = h.ny;
    }
    if (reals) {
      reals[0] = h.dc; reals[1] = h.period_x; reals[2] = h.period_y;
    }
  });
}

int omega_mesh_host_array(omega_mesh_t mesh, const char *name, const void **ptr, int *dtype,
                          int64_t *n0, int64_t *n1) {
  return guarded([&] {
    OMEGA_REQUIRE(mesh && name && ptr && dtype && n0 && n1, "null argument");
    HostMesh &h = mesh->dm.host;
    const std::string nm(name);
    const int64_t NC = h.ncells, NE = h.nedges, NV = h.nvertices, ME = h.max_edges,
                  ME2 = h.max_edges2;
    bool found = false;
    auto setI = [&](const char *nmx, std::vector<int> &v, int64_t a, int64_t b) {
      if (!found && nm == nmx) { *ptr = v.data(); *dtype = 0; *n0 = a; *n1 = b; found = true; }
    };
    auto setD = [&](const char *nmx, std::vector<double> &v, int64_t a, int64_t b) {
      if (!found && nm == nmx) { *ptr = v.data(); *dtype = 1; *n0 = a; *n1 = b; found = true; }
    };
    setI("nedges_on_cell", h.nedges_on_cell, NC, 1);
    setI("edges_on_cell", h.edges_on_cell, NC, ME);
    setI("cells_on_cell", h.cells_on_cell, NC, ME);
    setI("vertices_on_cell", h.vertices_on_cell, NC, ME);
    setI("kite_index_on_cell", h.kite_index_on_cell, NC, ME);
    setI("max_level_cell", h.max_level_cell, NC, 1);
    setD("area_cell", h.area_cell, NC, 1);
    setD("lat_cell", h.lat_cell, NC, 1);
    setD("lon_cell", h.lon_cell, NC, 1);
    setD("x_cell", h.x_cell, NC, 1);
    setD("y_cell", h.y_cell, NC, 1);
    setD("z_cell", h.z_cell, NC, 1);
    setD("mesh_density", h.mesh_density, NC, 1);
    setD("edge_sign_on_cell", h.edge_sign_on_cell, NC, ME);
    setI("nedges_on_edge", h.nedges_on_edge, NE, 1);
    setI("edges_on_edge", h.edges_on_edge, NE, ME2);
    setI("cells_on_edge", h.cells_on_edge, NE, 2);
    setI("vertices_on_edge", h.vertices_on_edge, NE, 2);
    setI("max_level_edge_bot", h.max_level_edge_bot, NE, 1);
    setI("max_level_edge_top", h.max_level_edge_top, NE, 1);
    setD("dc_edge", h.dc_edge, NE, 1);
    setD("dv_edge", h.dv_edge, NE, 1);
    setD("angle_edge", h.angle_edge, NE, 1);
    setD("lat_edge", h.lat_edge, NE, 1);
    setD("lon_edge", h.lon_edge, NE, 1);
    setD("x_edge", h.x_edge, NE, 1);
    setD("y_edge", h.y_edge, NE, 1);
    setD("z_edge", h.z_edge, NE, 1);
    setD("weights_on_edge", h.weights_on_edge, NE, ME2);
    setD("mesh_scaling_del2", h.mesh_scaling_del2, NE, 1);
    setD("mesh_scaling_del4", h.mesh_scaling_del4, NE, 1);
    setI("edges_on_vertex", h.edges_on_vertex, NV, 3);
    setI("cells_on_vertex", h.cells_on_vertex, NV, 3);
    setI("max_level_vertex_bot", h.max_level_vertex_bot, NV, 1);
    setI("max_level_vertex_top", h.max_level_vertex_top, NV, 1);
    setD("area_triangle", h.area_triangle, NV, 1);
    setD("lat_vertex", h.lat_vertex, NV, 1);
    setD("lon_vertex", h.lon_vertex, NV, 1);
    setD("x_vertex", h.x_vertex, NV, 1);
    setD("y_vertex", h.y_vertex, NV, 1);
    setD("z_vertex", h.z_vertex, NV, 1);
    setD("kiteareas_on_vertex", h.kiteareas_on_vertex, NV, 3);
    setD("edge_sign_on_vertex", h.edge_sign_on_vertex, NV, 3);
    if (!found) throw Error(OMEGA_ERR_NOTFOUND, "unknown mesh array: " + nm);
  });
}

int omega_mesh_device_array(omega_mesh_t mesh, const char *name, const double **dptr,
                            int64_t *n) {
  return guarded([&] {
    OMEGA_REQUIRE(mesh && name && dptr && n, "null argument");
    DeviceMesh &dm = mesh->dm;
    dm.upload(); // the device copies are made on first use
    const std::string nm(name);
    DBuf<Real> *b = nullptr;
    if (nm == "x_cell") b = &dm.x_cell;
    else if (nm == "y_cell") b = &dm.y_cell;
    else if (nm == "x_edge") b = &dm.x_edge;
    else if (nm == "y_edge") b = &dm.y_edge;
    else if (nm == "angle_edge") b = &dm.angle_edge;
    else if (nm == "x_vertex") b = &dm.x_vertex;
    else if (nm == "y_vertex") b = &dm.y_vertex;
    else if (nm == "area_cell") b = &dm.area_cell;
    if (!b) throw Error(OMEGA_ERR_NOTFOUND, "unknown device array: " + nm);
    *dptr = b->p;
    *n = (int64_t)b->n;
  });
}

int omega_mesh_permutation(omega_mesh_t mesh, int kind, int *new2old) {
  return guarded([&] {
    OMEGA_REQUIRE(mesh && new2old, "null argument");
    const DeviceMesh &dm = mesh->dm;
    const std::vector<int> *p = kind == 0 ? &dm.cell_n2o : kind == 1 ? &dm.edge_n2o
                                : kind == 2 ? &dm.vert_n2o : nullptr;
    OMEGA_REQUIRE(p, "kind must be 0, 1 or 2");
    std::copy(p->begin(), p->end(), new2old);
  });
}

int omega_mesh_regular_hex(omega_mesh_t mesh, int *is_regular) {
  return guarded([&] {
    OMEGA_REQUIRE(mesh && is_regular, "null argument");
    *is_regular = mesh->dm.hex_ok ? 1 : 0;
  });
}

// --------------------------------------------------------------------------
int omega_state_create(omega_mesh_t mesh, int ntracers, omega_state_t *out) {
  return guarded([&] {
    OMEGA_REQUIRE(mesh && out, "null argument");
    OMEGA_REQUIRE(ntracers >= 0, "ntracers must be >= 0");
    mesh->dm.upload();
    auto s = std::make_unique<omega_state_s>();
    const HostMesh &h = mesh->dm.host;
    s->mesh = mesh;
    s->ntr = ntracers;
    s->n[0] = (size_t)h.ncells * h.nlayers;
    s->n[1] = (size_t)h.nedges * h.nlayers;
    s->n[2] = (size_t)ntracers * h.ncells * h.nlayers;
    for (int f = 0; f < 3; ++f) s->buf[0][f].alloc(s->n[f]);
    *out = s.release();
  });
}

int omega_state_destroy(omega_state_t state) {
  return guarded([&] { delete state; });
}

namespace {
// host (original numbering) <-> device (device numbering) row permutation
void permute_rows(const double *src, double *dst, const std::vector<int> &n2o, size_t L,
                  bool to_device) {
  const size_t n = n2o.size();
  for (size_t i = 0; i < n; ++i) {
    const size_t o = (size_t)n2o[i];
    if (to_device) std::memcpy(dst + i * L, src + o * L, L * sizeof(double));
    else std::memcpy(dst + o * L, src + i * L, L * sizeof(double));
  }
}
} // namespace

int omega_state_upload(omega_state_t S, const double *h, const double *vn, const double *tr,
                       omega_stream_t stream) {
  return guarded([&] {
    OMEGA_REQUIRE(S, "null state");
    const DeviceMesh &dm = S->mesh->dm;
    const size_t L = dm.host.nlayers;
    hipStream_t s = as_stream(stream);
    std::vector<double> tmp;
    auto put = [&](int f, const double *src, const std::vector<int> &n2o, int reps) {
      if (!src || !S->n[f]) return;
      tmp.resize(S->n[f]);
      const size_t per = n2o.size() * L;
      for (int r = 0; r < reps; ++r) permute_rows(src + r * per, tmp.data() + r * per, n2o, L, true);
      OMEGA_HIP_CHECK(hipMemcpyAsync(S->ptr(f), tmp.data(), S->n[f] * sizeof(double),
                                     hipMemcpyHostToDevice, s));
      OMEGA_HIP_CHECK(hipStreamSynchronize(s));
    };
    put(0, h, dm.cell_n2o, 1);
    put(1, vn, dm.edge_n2o, 1);
    put(2, tr, dm.cell_n2o, S->ntr);
  });
}

int omega_state_download(omega_state_t S, double *h, double *vn, double *tr,
                         omega_stream_t stream) {
  return guarded([&] {
    OMEGA_REQUIRE(S, "null state");
    const DeviceMesh &dm = S->mesh->dm;
    const size_t L = dm.host.nlayers;
    hipStream_t s = as_stream(stream);
    std::vector<double> tmp;
    auto get = [&](int f, double *dst, const std::vector<int> &n2o, int reps) {
      if (!dst || !S->n[f]) return;
      tmp.resize(S->n[f]);
      OMEGA_HIP_CHECK(hipMemcpyAsync(tmp.data(), S->ptr(f), S->n[f] * sizeof(double),
                                     hipMemcpyDeviceToHost, s));
      OMEGA_HIP_CHECK(hipStreamSynchronize(s));
      const size_t per = n2o.size() * L;
      for (int r = 0; r < reps; ++r) permute_rows(tmp.data() + r * per, dst + r * per, n2o, L, false);
    };
    get(0, h, dm.cell_n2o, 1);
    get(1, vn, dm.edge_n2o, 1);
    get(2, tr, dm.cell_n2o, S->ntr);
  });
}

int omega_state_device_ptrs(omega_state_t S, double **h, double **vn, double **tr) {
  return guarded([&] {
    OMEGA_REQUIRE(S, "null state");
    if (h) *h = S->ptr(0);
    if (vn) *vn = S->ptr(1);
    if (tr) *tr = S->ptr(2);
  });
}

int omega_state_copy(omega_state_t dst, omega_state_t src, omega_stream_t stream) {
  return guarded([&] {
    OMEGA_REQUIRE(dst && src, "null state");
    OMEGA_REQUIRE(dst->mesh == src->mesh && dst->ntr == src->ntr, "state shapes differ");
    for (int f = 0; f < 3; ++f)
      if (src->n[f])
        OMEGA_HIP_CHECK(hipMemcpyAsync(dst->ptr(f), src->ptr(f), src->n[f] * sizeof(double),
                                       hipMemcpyDeviceToDevice, as_stream(stream)));
  });
}

// --------------------------------------------------------------------------
int omega_model_create(omega_mesh_t mesh, const omega_params *p, omega_model_t *out) {
  return guarded([&] {
    OMEGA_REQUIRE(mesh && p && out, "null argument");
    OMEGA_REQUIRE(p->ntracers >= 0, "ntracers must be >= 0");
    mesh->dm.upload();
    auto M = std::make_unique<omega_model_s>();
    M->mesh = mesh;
    M->params = *p;
    // shallow_water_model.cpp:17-35
    M->terms = T_PV | T_KE | T_SSH;
    if (p->visc_del2 > 0) M->terms |= T_VDEL2;
    if (p->visc_del4 > 0) M->terms |= T_VDEL4;
    if (p->eddy_diff2 > 0) M->terms |= T_TDEL2;
    if (p->eddy_diff4 > 0) M->terms |= T_TDEL4;
    const int L = mesh->dm.host.nlayers;
    M->lc.lpt = (L % 2 == 0 && L >= 2) ? 2 : 1;
    M->lc.rows = 0;
    M->lc.swizzle = 1;
    model_build_coefficients(M.get());
    if (M->terms & T_VDEL4) ensure_vdel4(M.get());
    *out = M.release();
  });
}

int omega_model_create_linear(omega_mesh_t mesh, const omega_params *p, double h0,
                              omega_model_t *out) {
  return guarded([&] {
    OMEGA_REQUIRE(mesh && p && out, "null argument");
    OMEGA_REQUIRE(p->ntracers == 0, "the linear model carries no tracers (ntracers must be 0)");
    OMEGA_REQUIRE(std::isfinite(h0) && h0 >= 0, "h0 must be finite and >= 0");
    mesh->dm.upload();
    auto M = std::make_unique<omega_model_s>();
    M->mesh = mesh;
    M->params = *p;
    M->linear = true;
    M->h0 = h0;
    M->terms = 0;
    const int L = mesh->dm.host.nlayers;
    M->lc.lpt = (L % 2 == 0 && L >= 2) ? 2 : 1;
    M->lc.rows = 0;
    M->lc.swizzle = 1;
    model_build_coefficients(M.get());
    model_build_linear(M.get());
    *out = M.release();
  });
}

int omega_model_destroy(omega_model_t model) {
  return guarded([&] { delete model; });
}

int omega_model_set_f_vertex(omega_model_t M, const double *f) {
  return guarded([&] {
    OMEGA_REQUIRE(M && f, "null argument");
    const DeviceMesh &dm = M->mesh->dm;
    std::vector<Real> tmp(dm.host.nvertices);
    for (int i = 0; i < dm.host.nvertices; ++i) tmp[i] = f[dm.vert_n2o[i]];
    M->f_v.upload(tmp);
    M->f_uniform = rows_equal(tmp, 1);
    M->touch();
  });
}

int omega_model_set_additional_tendency(omega_model_t M, omega_additional_tendency_fn fn,
                                        void *user) {
  return guarded([&] {
    OMEGA_REQUIRE(M, "null model");
    M->add_fn = fn;
    M->add_user = user;
    M->touch();
  });
}

int omega_model_compute_tendency(omega_model_t M, omega_state_t tend, omega_state_t state,
                                 double t, int add_mode, omega_stream_t stream) {
  return guarded([&] {
    OMEGA_REQUIRE(M, "null model");
    check_state(M, state, "compute_tendency(state)");
    check_state(M, tend, "compute_tendency(tend)");
    OMEGA_REQUIRE(add_mode == OMEGA_ADD_REPLACE || add_mode == OMEGA_ADD_INCREMENT,
                  "bad add_mode");
    compute_tendency_ptrs(M, tend->ptr(0), tend->ptr(1), tend->ptr(2), state->ptr(0),
                          state->ptr(1), state->ptr(2), t, add_mode, as_stream(stream));
  });
}

int omega_model_compute_field_tendency(omega_model_t M, int field, omega_state_t tend,
                                       omega_state_t state, int add_mode,
                                       omega_stream_t stream) {
  return guarded([&] {
    OMEGA_REQUIRE(M, "null model");
    OMEGA_REQUIRE(field >= 0 && field <= 2, "field must be 0 (h), 1 (vn) or 2 (tracers)");
    check_state(M, state, "compute_field_tendency(state)");
    check_state(M, tend, "compute_field_tendency(tend)");
    OMEGA_REQUIRE(add_mode == OMEGA_ADD_REPLACE || add_mode == OMEGA_ADD_INCREMENT,
                  "bad add_mode");
    compute_tendency_ptrs(M, tend->ptr(0), tend->ptr(1), tend->ptr(2), state->ptr(0),
                          state->ptr(1), state->ptr(2), 0.0, add_mode, as_stream(stream),
                          1 << field);
  });
}

int omega_model_compute_aux(omega_model_t M, const double *h, const double *vn,
                            const double *tr, omega_stream_t stream) {
  return guarded([&] {
    OMEGA_REQUIRE(M, "null model");
    OMEGA_REQUIRE(h && vn, "compute_aux: h and vn are required");
    OMEGA_REQUIRE(tr || M->params.ntracers == 0, "compute_aux: tr is required with tracers");
    if (M->linear) {
      // the linear model has no auxiliary variables; only the source is kept
      M->aux_h = h;
      M->aux_vn = vn;
      M->aux_tr = tr;
      return;
    }
    run_diag(M, h, vn, tr, as_stream(stream));
  });
}

namespace {
// One field of compute_tendency from device pointers; the thickness and
// tracers come from the auxiliary state's source, as the reference reads
// flux_h_edge, mean_h_edge and norm_tr_cell from m_aux_state.
void field_ptrs(omega_model_t M, int field, double *tend, const double *h, const double *vn,
                const double *tr, int add_mode, omega_stream_t stream) {
  OMEGA_REQUIRE(M, "null model");
  OMEGA_REQUIRE(tend && vn, "null argument");
  OMEGA_REQUIRE(add_mode == OMEGA_ADD_REPLACE || add_mode == OMEGA_ADD_INCREMENT,
                "bad add_mode");
  const char *name[3] = {"compute_h_tendency", "compute_vn_tendency", "compute_tr_tendency"};
  const Real *hh = M->aux_h, *tt = M->aux_tr;
  if (M->linear) {
    // the linear tendencies read their arguments directly
    // (shallow_water_hierarchical.cpp:680-741)
    OMEGA_REQUIRE(field != 2, "the linear model has no tracers");
    hh = field == 1 ? h : (M->aux_h ? M->aux_h : h);
  } else {
    OMEGA_REQUIRE(M->aux_h != nullptr,
                  std::string(name[field]) + ": no auxiliary state (call compute_aux or "
                                             "compute_tendency first)");
    if (field == 1)
      OMEGA_REQUIRE(h == M->aux_h,
                    "compute_vn_tendency: h must be the thickness the auxiliary state was "
                    "computed from (its SSH gradient and thickness flux read the same array)");
    if (field == 2 && M->params.ntracers > 0)
      OMEGA_REQUIRE(tt != nullptr, "compute_tr_tendency: the auxiliary state has no tracers");
  }
  Real *ht = field == 0 ? tend : nullptr;
  Real *vt = field == 1 ? tend : nullptr;
  Real *trt = field == 2 ? tend : nullptr;
  (void)tr;
  StageArgs a{};
  a.epi = EPI_STORE;
  a.base_from_tend = add_mode == OMEGA_ADD_INCREMENT;
  a.a = 1.0;
  a.h = fio(hh, ht, nullptr, nullptr, nullptr);
  a.vn = fio(vn, vt, nullptr, nullptr, nullptr);
  a.tr = fio(tt, trt, nullptr, nullptr, nullptr);
  run_stage(M, a, as_stream(stream), 1 << field, false);
}
} // namespace

int omega_model_compute_h_tendency(omega_model_t M, double *h_tend, const double *h,
                                   const double *vn, int add_mode, omega_stream_t stream) {
  return guarded([&] { field_ptrs(M, 0, h_tend, h, vn, nullptr, add_mode, stream); });
}

int omega_model_compute_vn_tendency(omega_model_t M, double *vn_tend, const double *h,
                                    const double *vn, int add_mode, omega_stream_t stream) {
  return guarded([&] {
    OMEGA_REQUIRE(h, "compute_vn_tendency: null h");
    field_ptrs(M, 1, vn_tend, h, vn, nullptr, add_mode, stream);
  });
}

int omega_model_compute_tr_tendency(omega_model_t M, double *tr_tend, const double *tr,
                                    const double *vn, int add_mode, omega_stream_t stream) {
  return guarded([&] { field_ptrs(M, 2, tr_tend, nullptr, vn, tr, add_mode, stream); });
}

int omega_model_aux_download(omega_model_t M, const char *name, double *out,
                             omega_stream_t stream) {
  return guarded([&] {
    OMEGA_REQUIRE(M && name && out, "null argument");
    OMEGA_REQUIRE(!M->linear, "the linear model has no auxiliary variables");
    const DeviceMesh &dm = M->mesh->dm;
    const size_t L = dm.host.nlayers;
    const std::string nm(name);
    const DiagOut d = M->diag();
    const Real *src = nullptr;
    const std::vector<int> *n2o = nullptr;
    int reps = 1;
    if (nm == "rvort_vertex") { src = d.rvort_v; n2o = &dm.vert_n2o; }
    else if (nm == "inv_thick_vertex") { src = d.inv_thick_v; n2o = &dm.vert_n2o; }
    else if (nm == "ke_cell") { src = d.ke_c; n2o = &dm.cell_n2o; }
    else if (nm == "vel_div_cell") { src = d.div_c; n2o = &dm.cell_n2o; }
    else if (nm == "tracer_del2_cell") {
      src = d.trdel2_c;
      n2o = &dm.cell_n2o;
      reps = M->params.ntracers;
    } else throw Error(OMEGA_ERR_NOTFOUND, "unknown auxiliary array: " + nm);
    OMEGA_REQUIRE(src != nullptr && reps > 0,
                  "auxiliary array " + nm + " is not computed for this model's terms");
    const size_t n = n2o->size() * L * reps;
    std::vector<double> tmp(n);
    hipStream_t s = as_stream(stream);
    OMEGA_HIP_CHECK(hipMemcpyAsync(tmp.data(), src, n * sizeof(double), hipMemcpyDeviceToHost, s));
    OMEGA_HIP_CHECK(hipStreamSynchronize(s));
    const size_t per = n2o->size() * L;
    for (int r = 0; r < reps; ++r)
      permute_rows(tmp.data() + r * per, out + r * per, *n2o, L, false);
  });
}

int omega_model_set_terms(omega_model_t M, int mask) {
  return guarded([&] {
    OMEGA_REQUIRE(M, "null model");
    OMEGA_REQUIRE((mask & ~T_ALL) == 0, "unknown term bits");
    OMEGA_REQUIRE(!(mask & T_DRAG) || !M->linear,
                  "bottom drag is a term of the nonlinear model only");
    M->terms = mask;
    M->touch();
    if (mask & T_VDEL4) ensure_vdel4(M);
    if ((mask & T_TDEL4) && M->params.ntracers > 0 && !M->trdel2_c.p)
      M->trdel2_c.alloc((size_t)M->params.ntracers * M->mesh->dm.host.ncells *
                        M->mesh->dm.host.nlayers);
  });
}

int omega_model_get_terms(omega_model_t M, int *mask) {
  return guarded([&] {
    OMEGA_REQUIRE(M && mask, "null argument");
    *mask = M->terms;
  });
}

int omega_model_set_launch(omega_model_t M, int rows, int lpt, int swizzle, int chunk) {
  return guarded([&] {
    OMEGA_REQUIRE(M, "null model");
    if (rows > 0) {
      OMEGA_REQUIRE(rows <= 256, "rows_per_block must be <= 256");
      M->lc.rows = rows;
    }
    if (lpt > 0) {
      OMEGA_REQUIRE(lpt == 1 || lpt == 2, "levels_per_thread must be 1 or 2");
      M->lc.lpt = lpt;
    }
    if (swizzle >= 0) M->lc.swizzle = swizzle ? 1 : 0;
    if (chunk >= 0) M->lc.chunk = chunk;
    M->lc.drows = -1; // the diagnostics launch follows
    M->touch();
  });
}

// Times reps tendency evaluations of `state` (diagnostics + tendency
// launches) for each launch shape -- cells per block (the default fill, 4,
// 2), levels per thread (2 when the layer count is even, 1) and the XCD
// swizzle (on, off) -- and keeps the fastest.  OMEGA_TUNE_CACHE works as for
// omega_batch_autotune, keyed by the model's cells, layers, tracers, terms
// and kernel family.
int omega_model_autotune(omega_model_t M, omega_state_t state, int reps, double *best_ms,
                         omega_stream_t stream) {
  return guarded([&] {
    OMEGA_REQUIRE(M, "null model");
    check_state(M, state, "autotune");
    reps = std::max(1, reps);
    hipStream_t s = as_stream(stream);
    const HostMesh &hm = M->mesh->dm.host;
    const MeshView mv = M->view();
    const std::string sig = "model;" + std::to_string(hm.ncells) + ":" +
                            std::to_string(hm.nlayers) + ":" +
                            std::to_string(M->params.ntracers) + ":" + std::to_string(M->terms) +
                            ":" + (M->linear ? "l" : hex_usable(mv, M->lc) ? "h" : "g");
    const char *cache = std::getenv("OMEGA_TUNE_CACHE");
    if (cache && *cache) {
      std::ifstream in(cache);
      std::string line;
      while (std::getline(in, line)) {
        std::istringstream ls(line);
        std::string key;
        int r = 0, l = 0, w = 0, dr = -1;
        double ms = -1;
        if (!(ls >> key >> r >> l >> w >> ms) || key != sig) continue;
        if (!(ls >> dr) || dr < -1 || dr > 256) dr = -1;
        if (r >= 0 && r <= 256 && (l == 1 || (l == 2 && hm.nlayers % 2 == 0))) {
          M->lc.rows = r;
          M->lc.lpt = l;
          M->lc.swizzle = w ? 1 : 0;
          M->lc.drows = dr;
          M->touch();
          if (best_ms) *best_ms = ms;
          return;
        }
      }
    }
    DBuf<Real> t0(state->n[0]), t1(state->n[1]), t2(state->n[2]);
    hipEvent_t a, b;
    OMEGA_HIP_CHECK(hipEventCreate(&a));
    OMEGA_HIP_CHECK(hipEventCreate(&b));
    const Launch keep = M->lc;
    double best = -1;
    Launch pick = keep;
    for (int lpt : {2, 1}) {
      if (lpt == 2 && hm.nlayers % 2 != 0) continue;
      for (int rows : {0, 4, 2})
        for (int swz : {1, 0}) {
          M->lc = keep;
          M->lc.rows = rows;
          M->lc.lpt = lpt;
          M->lc.swizzle = swz;
          M->lc.drows = -1;
          compute_tendency_ptrs(M, t0.p, t1.p, t2.p, state->ptr(0), state->ptr(1),
                                state->ptr(2), 0.0, OMEGA_ADD_REPLACE, s, 7);
          OMEGA_HIP_CHECK(hipEventRecord(a, s));
          for (int i = 0; i < reps; ++i)
            compute_tendency_ptrs(M, t0.p, t1.p, t2.p, state->ptr(0), state->ptr(1),
                                  state->ptr(2), 0.0, OMEGA_ADD_REPLACE, s, 7);
          OMEGA_HIP_CHECK(hipEventRecord(b, s));
          OMEGA_HIP_CHECK(hipEventSynchronize(b));
          float ms = 0;
          OMEGA_HIP_CHECK(hipEventElapsedTime(&ms, a, b));
          const double per = ms / reps;
          if (best < 0 || per < best) {
            best = per;
            pick = M->lc;
          }
        }
    }
    // then the diagnostics launch's own block height, timed alone
    {
      double dbest = -1;
      int dpick = -1;
      for (int dr : {-1, 0, 4, 2}) {
        if (dr >= 0 && dr == pick.rows) continue;
        M->lc = pick;
        M->lc.drows = dr;
        run_diag(M, state->ptr(0), state->ptr(1), state->ptr(2), s);
        OMEGA_HIP_CHECK(hipEventRecord(a, s));
        for (int i = 0; i < reps; ++i) run_diag(M, state->ptr(0), state->ptr(1), state->ptr(2), s);
        OMEGA_HIP_CHECK(hipEventRecord(b, s));
        OMEGA_HIP_CHECK(hipEventSynchronize(b));
        float ms = 0;
        OMEGA_HIP_CHECK(hipEventElapsedTime(&ms, a, b));
        if (dbest < 0 || ms < dbest) {
          dbest = ms;
          dpick = dr;
        }
      }
      pick.drows = dpick;
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    M->lc = pick;
    M->touch();
    if (cache && *cache) {
      std::ofstream out(cache, std::ios::app);
      if (out)
        out << sig << " " << pick.rows << " " << pick.lpt << " " << pick.swizzle << " " << best
            << " " << pick.drows << "\n";
    }
    if (best_ms) *best_ms = best;
  });
}

int omega_model_set_variant(omega_model_t M, int variant) {
  return guarded([&] {
    OMEGA_REQUIRE(M, "null model");
    OMEGA_REQUIRE(variant == 0 || variant == 1, "variant must be 0 or 1");
    M->lc.variant = variant;
    M->touch();
  });
}

int omega_model_kernel_variant(omega_model_t M, int *in_use) {
  return guarded([&] {
    OMEGA_REQUIRE(M && in_use, "null argument");
    const MeshView mv = M->view();
    *in_use = M->linear ? 6 : general_only(M) ? 1 : hex_usable(mv, M->lc) ? 2 : 1;
  });
}

int omega_model_set_kernel_forms(omega_model_t M, int forms) {
  return guarded([&] {
    OMEGA_REQUIRE(M, "null model");
    OMEGA_REQUIRE(forms >= 0 && forms <= 15, "forms must be 0..15");
    M->forms = forms;
    M->touch();
  });
}

int omega_model_kernel_flags(omega_model_t M, int *flags) {
  return guarded([&] {
    OMEGA_REQUIRE(M && flags, "null argument");
    const MeshView mv = M->view();
    const ModelCoef mc = M->coef();
    const AdjLayout &al = M->mesh->adj;
    // the VJP's slot-uniform form (known once the adjoint tables exist)
    const bool adj_uni = al.built && al.fixed && al.uni && M->ect_uniform && mv.coef_uni &&
                         mc.uniform && mv.L % 2 == 0;
    // the pair tendency kernel's compiled-in term set runs for this model
    const bool tset = mv.term_sets && mv.coef_uni && mc.uniform && mv.off32 && mv.hexu_ok &&
                      !M->linear && !general_only(M) && hex_usable(mv, M->lc) &&
                      M->lc.lpt != 1 && mv.L % 64 == 0 && hex_term_set(M->terms) &&
                      !M->params.disable_h_tendency && !M->params.disable_vn_tendency;
    *flags = (mv.coef_uni && mc.uniform ? 1 : 0) | (mv.off32 ? 2 : 0) | (adj_uni ? 4 : 0) |
             (tset ? 8 : 0) | (stage_writes_pre(M) ? 16 : 0);
  });
}

int omega_model_get_launch(omega_model_t M, int *rows, int *lpt, int *swizzle, int *chunk) {
  return guarded([&] {
    OMEGA_REQUIRE(M, "null model");
    if (rows) *rows = M->lc.rows;
    if (lpt) *lpt = M->lc.lpt;
    if (swizzle) *swizzle = M->lc.swizzle;
    if (chunk) *chunk = M->lc.chunk;
  });
}

// --------------------------------------------------------------------------
int omega_stepper_create(omega_model_t M, int kind, int nstages, omega_stepper_t *out) {
  return guarded([&] {
    OMEGA_REQUIRE(M && out, "null argument");
    OMEGA_REQUIRE(kind == OMEGA_STEPPER_RK4 || kind == OMEGA_STEPPER_LSRK, "bad stepper kind");
    auto S = std::make_unique<omega_stepper_s>();
    S->model = M;
    S->kind = kind;
    const HostMesh &h = M->mesh->dm.host;
    const size_t L = h.nlayers, ntr = M->params.ntracers;
    OMEGA_REQUIRE(kind == OMEGA_STEPPER_RK4 || nstages >= 1, "nstages must be >= 1");
    S->nstages = kind == OMEGA_STEPPER_RK4 ? 4 : nstages;
    stepper_coefficients(kind, S->nstages, S->rka, S->rkb, S->rkc);
    if (kind == OMEGA_STEPPER_LSRK) {
      // tendency storage starts at zero (lsrk.cpp:10-14)
      S->tend[0].alloc((size_t)h.ncells * L);
      S->tend[1].alloc((size_t)h.nedges * L);
      S->tend[2].alloc(ntr * h.ncells * L);
    }
    *out = S.release();
  });
}

int omega_stepper_destroy(omega_stepper_t S) {
  return guarded([&] { delete S; });
}

int omega_stepper_set_graph(omega_stepper_t S, int enable) {
  return guarded([&] {
    OMEGA_REQUIRE(S, "null stepper");
    S->use_graph = enable != 0;
  });
}

int omega_stepper_coeffs(omega_stepper_t St, int *nstages, double *a, double *b, double *c) {
  return guarded([&] {
    OMEGA_REQUIRE(St, "null stepper");
    if (nstages) *nstages = St->nstages;
    auto put = [&](double *dst, const std::vector<Real> &src) {
      if (!dst) return;
      for (int i = 0; i < St->nstages; ++i) dst[i] = i < (int)src.size() ? src[i] : 0.0;
    };
    put(a, St->rka);
    put(b, St->rkb);
    put(c, St->rkc);
  });
}

int omega_stepper_do_step(omega_stepper_t St, double t, double dt, omega_state_t S,
                          omega_stream_t stream) {
  return guarded([&] {
    OMEGA_REQUIRE(St, "null stepper");
    check_state(St->model, S, "do_step");
    do_step_maybe_graph(St, S, t, dt, as_stream(stream));
  });
}

int omega_stepper_run(omega_stepper_t St, double t0, double dt, int nsteps, int t_offset,
                      omega_state_t S, omega_stream_t stream) {
  return guarded([&] {
    OMEGA_REQUIRE(St, "null stepper");
    OMEGA_REQUIRE(nsteps >= 0, "nsteps must be >= 0");
    check_state(St->model, S, "run");
    for (int i = 0; i < nsteps; ++i)
      do_step_maybe_graph(St, S, t0 + (i + t_offset) * dt, dt, as_stream(stream));
  });
}

// --------------------------------------------------------------------------
namespace {
int diag_integral(omega_model_t M, omega_state_t S, double *out, omega_stream_t stream,
                  int which) {
  return guarded([&] {
    OMEGA_REQUIRE(M && out, "null argument");
    check_state(M, S, "integral");
    const DeviceMesh &dm = M->mesh->dm;
    const HostMesh &h = dm.host;
    hipStream_t s = as_stream(stream);
    launch_reduce(which, h.ncells, h.nvertices, h.max_edges, h.nlayers, dm.noc.p, dm.eoc.p,
                  dm.eov.p, dm.ae_c.p, dm.area_cell.p, dm.dcs_v.p, dm.area_tri.p, M->f_v.p,
                  M->params.grav, S->ptr(0), S->ptr(1), M->partial.p, M->result.p, s,
                  M->linear ? M->h0 : -1.0);
    OMEGA_HIP_CHECK(hipMemcpyAsync(out, M->result.p, sizeof(double), hipMemcpyDeviceToHost, s));
    OMEGA_HIP_CHECK(hipStreamSynchronize(s));
  });
}
} // namespace

int omega_energy_integral(omega_model_t M, omega_state_t S, double *out, omega_stream_t s) {
  return diag_integral(M, S, out, s, 0);
}
int omega_mass_integral(omega_model_t M, omega_state_t S, double *out, omega_stream_t s) {
  return diag_integral(M, S, out, s, 1);
}
int omega_circulation_integral(omega_model_t M, omega_state_t S, double *out, omega_stream_t s) {
  return diag_integral(M, S, out, s, 2);
}

// --------------------------------------------------------------------------
int omega_add_host_field(omega_mesh_t mesh, int kind, double *dst, const double *src, int ncols,
                         double alpha, omega_stream_t stream) {
  return guarded([&] {
    OMEGA_REQUIRE(mesh && dst && src, "null argument");
    DeviceMesh &dm = mesh->dm;
    dm.upload();
    const std::vector<int> *p = kind == 0 ? &dm.cell_n2o : kind == 1 ? &dm.edge_n2o
                                : kind == 2 ? &dm.vert_n2o : nullptr;
    OMEGA_REQUIRE(p, "kind must be 0, 1 or 2");
    const int L = dm.host.nlayers;
    OMEGA_REQUIRE(ncols == 1 || ncols == L, "ncols must be 1 or nlayers");
    const size_t n = p->size();
    std::vector<double> tmp(n * ncols);
    permute_rows(src, tmp.data(), *p, ncols, true);
    if (mesh->staging.n < tmp.size()) mesh->staging.alloc(tmp.size());
    hipStream_t s = as_stream(stream);
    OMEGA_HIP_CHECK(hipMemcpyAsync(mesh->staging.p, tmp.data(), tmp.size() * sizeof(double),
                                   hipMemcpyHostToDevice, s));
    launch_add_field(dst, mesh->staging.p, (int)n, L, ncols, alpha, s);
    OMEGA_HIP_CHECK(hipStreamSynchronize(s));
  });
}

int omega_synchronize(omega_stream_t stream) {
  return guarded([&] {
    if (stream) OMEGA_HIP_CHECK(hipStreamSynchronize(as_stream(stream)));
    else OMEGA_HIP_CHECK(hipDeviceSynchronize());
  });
}

int omega_timing_enable(int enable) {
  return guarded([&] {
    TimingTable &T = timing();
    std::lock_guard<std::mutex> g(T.mu);
    T.enabled = enable != 0;
  });
}

int omega_timing_reset(void) {
  return guarded([&] {
    TimingTable &T = timing();
    std::lock_guard<std::mutex> g(T.mu);
    T.drain();
    std::fill(T.total_ms.begin(), T.total_ms.end(), 0.0);
    std::fill(T.count.begin(), T.count.end(), 0);
  });
}

int omega_timing_read(const char *name, double *total_ms, int64_t *launches) {
  return guarded([&] {
    OMEGA_REQUIRE(name && total_ms && launches, "null argument");
    TimingTable &T = timing();
    std::lock_guard<std::mutex> g(T.mu);
    T.drain();
    const int slot = T.slot_of(name);
    *total_ms = T.total_ms[slot];
    *launches = T.count[slot];
  });
}

} // extern "C"
