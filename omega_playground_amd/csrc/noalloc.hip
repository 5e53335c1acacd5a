// Tendency evaluation that allocates nothing, for callers that cannot call
// hipMalloc (code under HIP graph capture, tight serving loops, harnesses
// that forbid allocation).  It is omega_model_compute_tendency with the
// diagnostics the kernels pass between launches (vertex vorticity and
// 1/thickness, cell kinetic energy and divergence, tracer Laplacians, the
// momentum del4 pre-pass) kept in one file-scope device array instead of
// the model's buffers.  Every call on a device shares that device's array,
// so calls that overlap on one GPU race; the model's own buffers and its
// auxiliary-state record are left as they were.
#include <map>
#include <mutex>

#include "objects.hpp"

// The scratch array: kScratchBytes on every device the library runs on,
// held until the process ends.  Host code reaches it only through the
// address hipGetSymbolAddress returns for the current device.
namespace {
constexpr size_t kScratchBytes = size_t(64) << 20;
constexpr size_t kScratchDoubles = kScratchBytes / sizeof(double);
} // namespace

static __device__ __align__(256) double scratch[kScratchDoubles];

namespace {

using namespace omega_hip;

std::mutex g_scratch_mu;
std::map<int, double *> g_scratch; // device ordinal -> address of its array

// The current device's copy of the array (looked up once per device).
double *scratch_base() {
  int dev = 0;
  OMEGA_HIP_CHECK(hipGetDevice(&dev));
  std::lock_guard<std::mutex> g(g_scratch_mu);
  auto it = g_scratch.find(dev);
  if (it != g_scratch.end()) return it->second;
  void *p = nullptr;
  OMEGA_HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  OMEGA_REQUIRE(p != nullptr, "scratch array has no device address");
  g_scratch.emplace(dev, static_cast<double *>(p));
  return static_cast<double *>(p);
}

// Carves the diagnostics the model's enabled terms need out of base (each
// array starts on a 512-byte boundary); returns the doubles used.  With
// base == nullptr only the size is computed.
size_t carve(const omega_model_s *M, double *base, DiagOut &d) {
  d = DiagOut{};
  if (M->linear || !diag_needed(M)) return 0;
  const HostMesh &hm = M->mesh->dm.host;
  const size_t L = hm.nlayers, NC = hm.ncells, NE = hm.nedges, NV = hm.nvertices;
  const size_t ntr = M->params.ntracers;
  size_t off = 0;
  auto take = [&](size_t n) -> double * {
    double *p = base ? base + off : nullptr;
    off += (n + 63) / 64 * 64;
    return p;
  };
  d.rvort_v = take(NV * L);
  d.inv_thick_v = take(NV * L);
  d.ke_c = take(NC * L);
  d.div_c = take(NC * L);
  if ((M->terms & T_TDEL4) && ntr > 0) d.trdel2_c = take(ntr * NC * L);
  if (!M->params.disable_vn_tendency && (M->terms & T_VDEL4)) {
    d.vdel2_e = take(NE * L);
    d.vdel2_rv_v = take(NV * L);
    d.vdel2_div_c = take(NC * L);
  }
  return off;
}

} // namespace

extern "C" {

int omega_model_compute_tendency_noalloc(omega_model_t M, omega_state_t tend,
                                         omega_state_t state, double t, int add_mode,
                                         omega_stream_t stream) {
  using namespace omega_hip;
  return guarded([&] {
    OMEGA_REQUIRE(M, "null model");
    check_state(M, state, "compute_tendency_noalloc(state)");
    check_state(M, tend, "compute_tendency_noalloc(tend)");
    OMEGA_REQUIRE(add_mode == OMEGA_ADD_REPLACE || add_mode == OMEGA_ADD_INCREMENT,
                  "bad add_mode");
    DiagOut d;
    const size_t need = carve(M, nullptr, d);
    OMEGA_REQUIRE(need <= kScratchDoubles,
                  "compute_tendency_noalloc: the diagnostics of this model need " +
                      std::to_string(need * sizeof(double)) +
                      " bytes of scratch; the library's static scratch holds " +
                      std::to_string(kScratchBytes) +
                      " (use omega_model_compute_tendency for larger models)");
    if (need) carve(M, scratch_base(), d);
    hipStream_t s = as_stream(stream);
    StageArgs a{};
    a.epi = EPI_STORE;
    a.base_from_tend = add_mode == OMEGA_ADD_INCREMENT;
    a.a = 1.0;
    a.h = fio(state->ptr(0), tend->ptr(0), nullptr, nullptr, nullptr);
    a.vn = fio(state->ptr(1), tend->ptr(1), nullptr, nullptr, nullptr);
    a.tr = fio(state->ptr(2), tend->ptr(2), nullptr, nullptr, nullptr);
    run_stage_into(M, a, s, 7, &d);
    // the hook runs after the built-in terms (shallow_water_model.cpp:344-345)
    call_hook(M, tend->ptr(0), tend->ptr(1), state->ptr(0), state->ptr(1), t, s);
  });
}

} // extern "C"
