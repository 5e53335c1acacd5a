// Internal declarations of the MI355X shallow-water library.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/omega_swe.h"

namespace omega_hip {

using Real = double;

// ---------------------------------------------------------------------------
// Errors: thrown inside the library, converted to status codes at the ABI.
// ---------------------------------------------------------------------------
struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

#define OMEGA_HIP_CHECK(expr)                                                          \
  do {                                                                                 \
    hipError_t _e = (expr);                                                            \
    if (_e != hipSuccess)                                                              \
      throw ::omega_hip::Error(OMEGA_ERR_HIP, std::string(#expr " failed: ") +         \
                                                  hipGetErrorString(_e) + " at " +     \
                                                  __FILE__ + ":" + std::to_string(__LINE__)); \
  } while (0)

#define OMEGA_REQUIRE(cond, msg)                                                       \
  do {                                                                                 \
    if (!(cond)) throw ::omega_hip::Error(OMEGA_ERR_ARG, msg);                         \
  } while (0)

// Device buffer with RAII ownership.
template <class T> struct DBuf {
  T *p = nullptr;
  size_t n = 0;
  DBuf() = default;
  explicit DBuf(size_t count) { alloc(count); }
  DBuf(const DBuf &) = delete;
  DBuf &operator=(const DBuf &) = delete;
  DBuf(DBuf &&o) noexcept : p(o.p), n(o.n) { o.p = nullptr; o.n = 0; }
  DBuf &operator=(DBuf &&o) noexcept {
    if (this != &o) { release(); p = o.p; n = o.n; o.p = nullptr; o.n = 0; }
    return *this;
  }
  ~DBuf() { release(); }
  void alloc(size_t count) {
    release();
    n = count;
    if (count) {
      OMEGA_HIP_CHECK(hipMalloc(&p, count * sizeof(T)));
      OMEGA_HIP_CHECK(hipMemset(p, 0, count * sizeof(T)));
    }
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  void upload(const std::vector<T> &h) {
    if (n != h.size()) alloc(h.size());
    if (n) OMEGA_HIP_CHECK(hipMemcpy(p, h.data(), n * sizeof(T), hipMemcpyHostToDevice));
  }
};

// ---------------------------------------------------------------------------
// Host mesh in the reference's numbering (MPASMesh fields, mpas_mesh.hpp:7-71)
// ---------------------------------------------------------------------------
struct HostMesh {
  int nx = 0, ny = 0;
  double dc = 0, period_x = 0, period_y = 0;
  int nlayers = 0, ncells = 0, nedges = 0, nvertices = 0;
  int max_edges = 6, max_edges2 = 12;

  std::vector<int> nedges_on_cell, edges_on_cell, cells_on_cell, vertices_on_cell,
      kite_index_on_cell, max_level_cell;
  std::vector<double> area_cell, lat_cell, lon_cell, x_cell, y_cell, z_cell, mesh_density,
      edge_sign_on_cell;
  std::vector<int> nedges_on_edge, edges_on_edge, cells_on_edge, vertices_on_edge,
      max_level_edge_bot, max_level_edge_top;
  std::vector<double> dc_edge, dv_edge, angle_edge, lat_edge, lon_edge, x_edge, y_edge, z_edge,
      weights_on_edge, mesh_scaling_del2, mesh_scaling_del4;
  std::vector<int> edges_on_vertex, cells_on_vertex, max_level_vertex_bot, max_level_vertex_top;
  std::vector<double> area_triangle, lat_vertex, lon_vertex, x_vertex, y_vertex, z_vertex,
      kiteareas_on_vertex, edge_sign_on_vertex;

  void resize();
  void derive_signs_and_levels(); // MPASMesh::finalize_mesh
};

void build_planar_hex(HostMesh &m, int nx, int ny, double dc, int nlayers);

// ---------------------------------------------------------------------------
// Device mesh: reordered connectivity plus host-precomputed coefficients.
// ---------------------------------------------------------------------------
struct MeshView {
  int nc, ne, nv, L;
  int me;   // row length of per-cell tables
  int me2;  // row length of per-edge neighbour tables
  const int *noc, *eoc;         // nedges_on_cell, edges_on_cell (device ids)
  const int *coe, *voe;         // cells_on_edge, vertices_on_edge
  const int *noe, *eoe;         // nedges_on_edge, edges_on_edge
  const int *eov, *cov;         // edges_on_vertex, cells_on_vertex
  const int *eofs, *vofs;       // owned edge / vertex ranges per cell [nc+1]
  const int *maxlev_e;          // max_level_edge_top
  const Real *cdiv;             // (dv*inv_area_cell)*sign  [nc][me]
  const Real *cke;              // inv_area_cell*(dv*dc)     [nc][me]
  const Real *w;                // weights_on_edge           [ne][me2]
  const Real *inv_dc, *inv_dv, *inv_dv4; // 1/dc, 1/dv, 1/max(dv, dc/4)
  const Real *s2, *s4;          // mesh_scaling_del2/4
  const Real *crv;              // (dc*inv_area_triangle)*sign [nv][3]
  const Real *cth;              // kite*inv_area_triangle     [nv][3]
  // regular-hexagon records (valid when hex_ok, see hex_pattern.hpp)
  int hex_ok, hexu_ok, hexd_ok;
  // coef_uni: the mesh's coefficient tables (inv_dc, inv_dv, maxlev_e, s2,
  // s4, w, cdiv, cke, crv, cth) are equal at every stencil slot, bit for
  // bit; off32: every [rows][L] array has < 2^24 rows and < 4 GiB
  int coef_uni, off32;
  // the pair tendency kernel may use its compiled-in term sets (cleared by
  // a model's kernel forms for tests)
  int term_sets;
  const int *hexrec;            // [ne][30]
  const int *hexu;              // [nc][49] per-cell union stencil (hex_pattern.hpp)
  const int *coc;               // cells_on_cell [nc][6]
};

struct DeviceMesh {
  HostMesh host;
  int ordering = OMEGA_ORDER_HILBERT;
  std::vector<int> cell_n2o, cell_o2n, edge_n2o, edge_o2n, vert_n2o, vert_o2n;
  std::vector<int> eofs_h, vofs_h;
  int me2 = 10;
  DBuf<int> noc, eoc, coe, voe, noe, eoe, eov, cov, eofs, vofs, maxlev_e;
  DBuf<Real> cdiv, cke, w, inv_dc, inv_dv, inv_dv4, s2, s4, crv, cth;
  // geometry in device numbering (for hooks and diagnostics)
  DBuf<Real> x_cell, y_cell, x_edge, y_edge, angle_edge, x_vertex, y_vertex, area_cell;
  DBuf<Real> area_tri, ae_c, dcs_v; // area_triangle; dv*dc per cell edge; dc*sign per vertex edge
  // host copies of the device-numbered coefficient tables (for the adjoint
  // and batch builders)
  std::vector<int> noc_h, eoc_h, coe_h, voe_h, noe_h, eoe_h, eov_h, cov_h, maxlev_h;
  std::vector<Real> cdiv_h, cke_h, w_h, inv_dc_h, inv_dv_h, inv_dv4_h, s2_h, s4_h, crv_h, cth_h;
  // regular-hexagon records for the pattern-specialised kernels
  bool hex_ok = false;
  bool hexu_ok = false; // per-cell union records valid (hexu)
  bool coef_uni = false; // coefficient tables equal at every stencil slot
  bool off32 = false;    // 32-bit byte offsets suffice for every row array
  bool hexd_ok = false; // owned vertices follow the diagnostics maps (hexd)
  std::vector<int> hexrec_h, coc_h, hexu_h;
  DBuf<int> hexrec, coc, hexu;

  bool uploaded = false;

  MeshView view() const;
  void build_device(int ordering); // host layout (permutations, coefficient tables)
  void upload();                   // device copies, on first use
};

// Cell ordering along a space-filling curve (host).
std::vector<int> order_cells(const HostMesh &m, int ordering);
// Builds and verifies the regular-hexagon records (hex_layout.cpp).
bool build_hex_records(DeviceMesh &dm);

// ---------------------------------------------------------------------------
// Launch configuration and kernel parameter blocks
// ---------------------------------------------------------------------------
struct Launch {
  int rows = 0;      // entities per block (blockDim.y); 0 fills 256 threads
  int lpt = 1;       // levels per thread (1 or 2)
  int swizzle = 1;   // XCD-contiguous block order
  int chunk = 0;     // levels per block (0: all levels, up to 64 lanes)
  int variant = 0;   // 0: pattern-specialised kernels when the mesh allows, 1: general
  int drows = -1;    // entities per block of the diagnostics launch (-1: rows)
};

// Enabled-term bits (same meaning as omega_model_set_terms).
enum : int {
  T_PV = 1, T_KE = 2, T_SSH = 4, T_VDEL2 = 8, T_VDEL4 = 16, T_TDEL2 = 32, T_TDEL4 = 64,
  T_DRAG = 128, // bottom drag, opt-in (no reference kernel applies it)
  T_ALL = 255
};

// Stage epilogue applied to a freshly computed tendency value.
enum Epilogue : int {
  EPI_STORE = 0,     // tend = base + F
  EPI_RK4_FIRST = 1, // acc = y + db*T; pout = y + da*T (y = stage input)
  EPI_RK4_MID = 2,   // acc += db*T; pout = yold + da*T
  EPI_RK4_LAST = 3,  // y = acc + db*T
  EPI_LSRK = 4,      // tend = base + F; yout = yin + db*T
  EPI_PROVIS = 5,    // pout = yold + da*T (stage recomputation for the adjoint)
};

// Buffers of one prognostic field (h, vn, or the tracer block) for a stage.
struct FieldIO {
  const Real *in;   // stage input (read at neighbours)
  Real *tend;       // tendency buffer (EPI_STORE / EPI_LSRK)
  Real *acc;        // RK4 accumulator
  const Real *yold; // RK4 y^n
  Real *out;        // stage output (provisional or new state)
};

struct DiagOut {
  Real *rvort_v, *inv_thick_v; // [nv][L]
  Real *ke_c, *div_c;          // [nc][L]
  Real *trdel2_c;              // [ntr][nc][L]
  Real *vdel2_e;               // [ne][L]  (momentum del4 pre-pass)
  Real *vdel2_rv_v, *vdel2_div_c;
};

struct ModelCoef {
  Real grav;
  const Real *f_v;   // [nv]
  const Real *cv2;   // visc_del2*mesh_scaling_del2 [ne]
  const Real *cv4;   // visc_del4*mesh_scaling_del4 [ne]
  const Real *ct2;   // ((eddy_diff2*inv_area)*dv)*sign [nc][me]
  const Real *ct4;   // (((eddy_diff4*inv_area)*dv)*sign)*scaling4 [nc][me]
  // f_v, cv2, cv4, ct2, ct4 equal at every stencil slot (bit for bit) on a
  // regular hexagonal mesh, so a pair wavefront's two cells share them
  int uniform;
  Real drag; // bottom drag coefficient (term T_DRAG, opt-in)
};

struct StageArgs {
  int terms;         // T_* bits
  int do_h, do_vn;   // !disable_*_tendency
  int ntr;
  int epi;           // Epilogue
  int base_from_tend;// T starts from a*tend (increment) instead of 0
  Real a;            // increment scale
  Real db, da;       // dt*b_s, dt*a_s
  FieldIO h, vn, tr;
};

// x[i*stride + j] == x[j] bit for bit for every row i (mesh_host.cpp)
template <class T> bool rows_equal(const std::vector<T> &x, size_t stride);

// Kernel launchers (kernels_fwd.hip)
void launch_diag(const MeshView &m, const Launch &lc, const ModelCoef &mc, int terms, int ntr,
                 const Real *h, const Real *vn, const Real *tr, const DiagOut &d,
                 hipStream_t s);
void launch_vdel4(const MeshView &m, const Launch &lc, const DiagOut &d, hipStream_t s);
void launch_tend(const MeshView &m, const Launch &lc, const ModelCoef &mc, const StageArgs &a,
                 const DiagOut &d, hipStream_t s);
// Linear model (LinearShallowWaterModel): h_t = -h0 div(vn),
// vn_t = f_e * (TRiSK tangential velocity) - g grad(h), with the stage update
struct LinCoef {
  const Real *cdv;  // dv*sign of each cell edge [nc][me]
  const Real *area; // area_cell [nc]
  const Real *dc;   // dc_edge [ne]
  const Real *f_e;  // f_edge [ne]
  Real h0, grav;
};
void launch_tend_linear(const MeshView &m, const Launch &lc, const LinCoef &lin,
                        const StageArgs &a, hipStream_t s);
// Rows a tendency VJP reads at its linearisation state ([ne][L] each): the
// edge PV sum Q_e + P_e, the edge thickness flux F_e u_e and the flux-weighted
// edges-on-edge sum sum_j w_ej F_j u_j.  A stage kernel that evaluates the
// tendency at that state can write them (the adjoint's stage recomputation).
struct PreOut {
  Real *se, *fu, *own;
};
// Pattern-specialised kernels for the regular hexagonal mesh (kernels_hex.hip)
bool hex_usable(const MeshView &m, const Launch &lc);
// launch_tend_hex writes PreOut rows for this mesh, launch shape and
// coefficients (the one-cell-per-row tendency kernel, or the pair kernel's
// shared-coefficient 32-bit vertex-union form; its vn tendency must be on)
bool hex_writes_pre(const MeshView &m, const Launch &lc, const ModelCoef &mc);
// the pair tendency kernel has a compiled-in form for this term set
bool hex_term_set(int terms);
// momentum del4 pre-pass in the pair layout (rows of 64 levels, two levels
// per lane on the pattern-specialised mesh)
bool hex_vdel4(const MeshView &m, const Launch &lc);
void launch_vdel4_hex(const MeshView &m, const Launch &lc, const DiagOut &d, hipStream_t s);
void launch_diag_hex(const MeshView &m, const Launch &lc, int terms, int ntr, const Real *h,
                     const Real *vn, const Real *tr, const DiagOut &d, hipStream_t s);
void launch_tend_hex(const MeshView &m, const Launch &lc, const ModelCoef &mc, const StageArgs &a,
                     const DiagOut &d, hipStream_t s, const PreOut *pre = nullptr);
// y += alpha*x elementwise
void launch_axpy(Real *y, Real alpha, const Real *x, size_t n, hipStream_t s);
// y = x + alpha*z elementwise
void launch_xpay(Real *y, const Real *x, Real alpha, const Real *z, size_t n, hipStream_t s);
// y *= alpha
void launch_scale(Real *y, Real alpha, size_t n, hipStream_t s);
// dst[i][k] += alpha * src[i][k or 0]
void launch_add_field(Real *dst, const Real *src, int n, int L, int ncols, Real alpha,
                      hipStream_t s);
// deterministic diagnostics reductions (kernels_reduce.hip); which: 0 energy,
// 1 mass, 2 circulation
int reduce_blocks(int n);
void launch_reduce(int which, int nc, int nv, int me, int L, const int *noc, const int *eoc,
                   const int *eov, const double *ae_c, const double *area_c, const double *dcs_v,
                   const double *area_t, const double *f_v, double grav, const double *h,
                   const double *vn, double *partial, double *out, hipStream_t s,
                   double linear_h0 = -1.0);

// Timing hooks
struct TimerScope {
  int slot;
  hipStream_t s;
  TimerScope(const char *name, hipStream_t s);
  ~TimerScope();
};

} // namespace omega_hip
