// Ensembles (no reference counterpart): independent members -- each its own
// mesh, layer count, tracer count, coefficients and time step -- advanced
// together.  Every stage phase (diagnostics, momentum del4 pre-pass,
// tendency + update) is one launch over all members' blocks, driven by a
// table of per-member descriptors built once per (step parity, stage); the
// step can be captured in a HIP graph.  The stage updates are the fused
// epilogues of the single-model steppers (rk4.cpp:28-104, lsrk.cpp:36-89),
// so each member advances bit for bit as it would alone.
#include <algorithm>
#include <cstdlib>
#include <fstream>
#include <map>
#include <sstream>

#include "batch.hpp"
#include "objects.hpp"

using namespace omega_hip;

struct omega_batch_s {
  int kind = OMEGA_STEPPER_LSRK;
  int nstages = 5;
  std::vector<Real> rka, rkb, rkc;
  struct Member {
    omega_model_s *M;
    omega_state_s *S;
    double dt;
    int cur0 = 0;                    // state buffer at parity 0
    DBuf<Real> acc[3], p1[3], p2[3]; // RK4
    DBuf<Real> tend[3];              // LSRK
    uint64_t gen = 0;                // model configuration the tables hold
  };
  std::vector<std::unique_ptr<Member>> mem;
  int N = 2;         // levels per lane
  int threads = 256; // threads per block
  int per_cu = 0;    // persistent grid: blocks per CU (0: one block per entry)
  bool built = false;
  int parity = 0;
  // one table over all members; member kinds share each launch
  int nblocks = 0;
  int kinds = 0; // bit 0 general-kernel members, bit 1 pattern-specialised
  int tset = 0;  // term set compiled into the pair tendency launch (0: none)
  bool any_vdel4 = false;
  DBuf<int> blk;
  DBuf<BatchMember> desc; // [parity][stage][member]
  bool use_graph = false;
  std::map<int, hipGraphExec_t> graphs;

  ~omega_batch_s() { drop_graphs(); }
  void drop_graphs() { destroy_graphs(graphs); }
};

namespace {

// Blocks of the persistent grid (0: one block per table entry).
int persistent_blocks(const omega_batch_s *B) {
  if (B->per_cu <= 0) return 0;
  static int cus = 0;
  if (!cus) {
    int dev = 0;
    OMEGA_HIP_CHECK(hipGetDevice(&dev));
    OMEGA_HIP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  }
  return B->per_cu * cus;
}

bool odd_flip(const omega_batch_s *B) {
  return B->kind == OMEGA_STEPPER_LSRK && (B->nstages % 2 == 1);
}

// Buffers of state field f at buffer index c (allocating the second buffer).
Real *state_buf(omega_state_s *S, int c, int f) {
  if (c != S->cur) return S->alt(f);
  return S->ptr(f);
}

void build_tables(omega_batch_s *B) {
  OMEGA_REQUIRE(!B->mem.empty(), "batch: no members");
  const int nm = (int)B->mem.size(), ns = B->nstages, N = B->N, T = B->threads;
  std::vector<int> blk;
  std::vector<BatchMember> geo(nm);
  B->any_vdel4 = false;
  for (int i = 0; i < nm; ++i) {
    auto &mb = *B->mem[i];
    omega_model_s *M = mb.M;
    omega_state_s *S = mb.S;
    const HostMesh &hm = M->mesh->dm.host;
    OMEGA_REQUIRE(N == 1 || hm.nlayers % 2 == 0,
                  "batch: two levels per lane need an even layer count in every member");
    M->mesh->dm.upload();
    if (M->terms & T_VDEL4) ensure_vdel4(M);
    for (int f = 0; f < 3; ++f) (void)S->alt(f);
    if (B->kind == OMEGA_STEPPER_RK4) {
      ensure3(mb.acc, S);
      ensure3(mb.p1, S);
      ensure3(mb.p2, S);
    } else {
      ensure3(mb.tend, S);
    }
    BatchMember &g = geo[i];
    // 16 or 32 layers run the pattern kernels with several cells per
    // wavefront, which needs two levels per lane
    g.hex = hex_usable(M->mesh->dm.view(), M->lc) && !general_only(M) &&
                    (hm.nlayers % 64 == 0 || N == 2)
                ? 1
                : 0;
    if (g.hex) {
      // one row of lanes covers 64 levels (or all 16 / 32); the pair form
      // (64-level rows, N = 2) needs an even row count, which every
      // admissible block size gives
      const int span = hm.nlayers % 64 == 0 ? 64 : hm.nlayers;
      g.lx = span / N;
      g.rows = std::max(1, T / g.lx);
      g.chunks = hm.nlayers / span;
    } else {
      int lanes = (hm.nlayers + N - 1) / N;
      if (lanes > 64) lanes = 64;
      int lx = 1;
      while (lx < lanes) lx <<= 1;
      g.lx = lx;
      g.rows = std::max(1, T / lx);
      g.chunks = (hm.nlayers + lx * N - 1) / (lx * N);
    }
    g.block0 = (int)blk.size();
    const int nb = ((hm.ncells + g.rows - 1) / g.rows) * g.chunks;
    blk.insert(blk.end(), nb, i);
    const bool do_vn = !M->params.disable_vn_tendency;
    const int ntr = M->params.ntracers;
    const bool need_diag = do_vn || ((M->terms & T_TDEL4) && ntr > 0);
    g.dterms = need_diag ? (do_vn ? M->terms : (M->terms & T_TDEL4)) : 0;
    g.vdel4 = do_vn && (M->terms & T_VDEL4);
    B->any_vdel4 = B->any_vdel4 || g.vdel4;
    mb.cur0 = S->cur;
    mb.gen = M->gen;
  }
  B->nblocks = (int)blk.size();
  B->kinds = 0;
  for (const BatchMember &g : geo) B->kinds |= g.hex ? 2 : 1;
  // the members' term set, compiled into the tendency launch when they all
  // share it (the 64k-layer members' pair operators also need the
  // shared-coefficient, 32-bit-offset forms and the vertex-union records)
  B->tset = 0;
  if (B->kinds == 2 && N == 2) {
    int ts = 0;
    bool ok = true;
    for (int i = 0; i < nm && ok; ++i) {
      omega_model_s *M = B->mem[i]->M;
      const MeshView v = M->view();
      const ModelCoef c = M->coef();
      const bool pair = M->mesh->dm.host.nlayers % 64 == 0;
      ok = v.term_sets && (!pair || (v.coef_uni && c.uniform && v.off32 && v.hexu_ok)) &&
           !M->params.disable_h_tendency && !M->params.disable_vn_tendency &&
           term_set_batch(M->terms) && (ts == 0 || ts == M->terms);
      ts = M->terms;
    }
    if (ok) B->tset = ts;
  }
  std::vector<BatchMember> tab((size_t)2 * ns * nm);
  for (int p = 0; p < 2; ++p)
    for (int s = 0; s < ns; ++s)
      for (int i = 0; i < nm; ++i) {
        auto &mb = *B->mem[i];
        omega_model_s *M = mb.M;
        omega_state_s *S = mb.S;
        BatchMember g = geo[i];
        g.m = M->view();
        g.mc = M->coef();
        g.d = M->diag();
        StageArgs a{};
        a.terms = M->terms;
        a.do_h = !M->params.disable_h_tendency;
        a.do_vn = !M->params.disable_vn_tendency;
        a.ntr = M->params.ntracers;
        const double dt = mb.dt;
        if (B->kind == OMEGA_STEPPER_RK4) {
          // rk4_fused_step's buffer rotation
          const int c = mb.cur0;
          Real *Y[3], *A[3], *P1[3], *P2[3];
          for (int f = 0; f < 3; ++f) {
            Y[f] = state_buf(S, c, f);
            A[f] = mb.acc[f].p;
            P1[f] = mb.p1[f].p;
            P2[f] = mb.p2[f].p;
          }
          a.base_from_tend = 0;
          a.db = dt * B->rkb[s];
          a.da = s < 3 ? dt * B->rka[s] : 0.0;
          Real **in = s == 0 ? Y : (s == 2 ? P2 : P1);
          Real **out = s == 0 ? P1 : (s == 1 ? P2 : (s == 2 ? P1 : Y));
          a.epi = s == 0 ? EPI_RK4_FIRST : (s == 3 ? EPI_RK4_LAST : EPI_RK4_MID);
          a.h = fio(in[0], nullptr, A[0], Y[0], out[0]);
          a.vn = fio(in[1], nullptr, A[1], Y[1], out[1]);
          a.tr = fio(in[2], nullptr, A[2], Y[2], out[2]);
        } else {
          // lsrk_fused_step: stage s reads buffer (cur0 ^ parity ^ s) and
          // writes the other one
          const int steps_parity = odd_flip(B) ? p : 0;
          const int cin = (mb.cur0 ^ steps_parity ^ (s & 1)) & 1;
          a.epi = EPI_LSRK;
          a.a = B->rka[s];
          a.base_from_tend = B->rka[s] != 0.0;
          a.db = dt * B->rkb[s];
          Real *in[3], *out[3];
          for (int f = 0; f < 3; ++f) {
            in[f] = state_buf(S, cin, f);
            out[f] = state_buf(S, 1 - cin, f);
          }
          a.h = fio(in[0], mb.tend[0].p, nullptr, nullptr, out[0]);
          a.vn = fio(in[1], mb.tend[1].p, nullptr, nullptr, out[1]);
          a.tr = fio(in[2], mb.tend[2].p, nullptr, nullptr, out[2]);
        }
        g.a = a;
        tab[((size_t)p * ns + s) * nm + i] = g;
      }
  B->blk.upload(blk);
  if (B->desc.n != tab.size()) B->desc.alloc(tab.size());
  OMEGA_HIP_CHECK(hipMemcpy(B->desc.p, tab.data(), tab.size() * sizeof(BatchMember),
                            hipMemcpyHostToDevice));
  B->parity = 0;
  B->drop_graphs();
  B->built = true;
}

// The tables hold buffer addresses for the states' current buffer parity
// and each member's model configuration; a member stepped on its own in
// between, or a model setting changed since (terms, f_vertex, launch shape,
// kernel forms: the model's configuration generation), forces a rebuild.
void ensure_tables(omega_batch_s *B) {
  bool ok = B->built;
  if (ok)
    for (auto &mbp : B->mem) {
      const auto &mb = *mbp;
      const int expect = odd_flip(B) ? (mb.cur0 ^ B->parity) : mb.cur0;
      if (mb.S->cur != expect || mb.gen != mb.M->gen) ok = false;
    }
  if (!ok) build_tables(B);
}

void step_launches(omega_batch_s *B, int parity, hipStream_t s) {
  const int nm = (int)B->mem.size();
  if (B->kind == OMEGA_STEPPER_LSRK)
    for (auto &mb : B->mem) {
      // the stage-0 input buffer of this parity is the state's current one
      omega_state_s *S = mb->S;
      const int cin = (mb->cur0 ^ (odd_flip(B) ? parity : 0)) & 1;
      const int keep = S->cur;
      S->cur = cin;
      carry_extra_tracers(mb->M, S, s);
      S->cur = keep;
    }
  const int pers = persistent_blocks(B);
  for (int st = 0; st < B->nstages; ++st) {
    const BatchMember *d = B->desc.p + ((size_t)parity * B->nstages + st) * nm;
    const int *g = B->blk.p;
    const int n = B->nblocks;
    launch_batch_diag(d, g, n, B->threads, B->N, pers, B->kinds, s);
    if (B->any_vdel4) launch_batch_vdel4(d, g, n, B->threads, B->N, pers, s);
    launch_batch_tend(d, g, n, B->threads, B->N, pers, B->kinds, B->tset, s);
  }
}

void batch_step(omega_batch_s *B, hipStream_t s) {
  ensure_tables(B);
  const int p = B->parity;
  if (B->use_graph && !stream_capturing(s)) {
    auto it = B->graphs.find(p);
    if (it == B->graphs.end()) {
      hipStream_t cs = capture_stream();
      hipGraph_t g;
      OMEGA_HIP_CHECK(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
      try {
        step_launches(B, p, cs);
      } catch (...) {
        hipGraph_t dummy;
        (void)hipStreamEndCapture(cs, &dummy);
        throw;
      }
      OMEGA_HIP_CHECK(hipStreamEndCapture(cs, &g));
      hipGraphExec_t ex;
      OMEGA_HIP_CHECK(hipGraphInstantiate(&ex, g, nullptr, nullptr, 0));
      OMEGA_HIP_CHECK(hipGraphDestroy(g));
      it = B->graphs.emplace(p, ex).first;
    }
    OMEGA_HIP_CHECK(hipGraphLaunch(it->second, s));
  } else {
    step_launches(B, p, s);
  }
  if (odd_flip(B)) {
    for (auto &mb : B->mem) mb->S->cur = 1 - mb->S->cur;
    B->parity ^= 1;
  }
}

double time_steps(omega_batch_s *B, int reps, hipStream_t s) {
  hipEvent_t a, b;
  OMEGA_HIP_CHECK(hipEventCreate(&a));
  OMEGA_HIP_CHECK(hipEventCreate(&b));
  batch_step(B, s); // warm-up (and graph capture)
  OMEGA_HIP_CHECK(hipEventRecord(a, s));
  for (int i = 0; i < reps; ++i) batch_step(B, s);
  OMEGA_HIP_CHECK(hipEventRecord(b, s));
  OMEGA_HIP_CHECK(hipEventSynchronize(b));
  float ms = 0;
  OMEGA_HIP_CHECK(hipEventElapsedTime(&ms, a, b));
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  return ms / std::max(reps, 1);
}

} // namespace

extern "C" {

int omega_batch_create(int kind, int nstages, omega_batch_t *out) {
  return guarded([&] {
    OMEGA_REQUIRE(out, "null argument");
    OMEGA_REQUIRE(kind == OMEGA_STEPPER_RK4 || kind == OMEGA_STEPPER_LSRK, "bad stepper kind");
    OMEGA_REQUIRE(kind == OMEGA_STEPPER_RK4 || nstages >= 1, "nstages must be >= 1");
    auto B = std::make_unique<omega_batch_s>();
    B->kind = kind;
    B->nstages = kind == OMEGA_STEPPER_RK4 ? 4 : nstages;
    stepper_coefficients(kind, B->nstages, B->rka, B->rkb, B->rkc);
    *out = B.release();
  });
}

int omega_batch_destroy(omega_batch_t B) {
  return guarded([&] { delete B; });
}

int omega_batch_add(omega_batch_t B, omega_model_t M, omega_state_t S, double dt, int *index) {
  return guarded([&] {
    OMEGA_REQUIRE(B && M, "null argument");
    check_state(M, S, "batch member");
    OMEGA_REQUIRE(M->add_fn == nullptr,
                  "batch: members cannot have an additional-tendency hook");
    OMEGA_REQUIRE(!M->linear, "batch: linear models step on their own (omega_stepper_*)");
    for (auto &mb : B->mem)
      OMEGA_REQUIRE(mb->S != S, "batch: a state can belong to one member only");
    // two levels per lane need an even layer count in every member
    if (M->mesh->dm.host.nlayers % 2 != 0) B->N = 1;
    auto mb = std::make_unique<omega_batch_s::Member>();
    mb->M = M;
    mb->S = S;
    mb->dt = dt;
    B->mem.push_back(std::move(mb));
    B->built = false;
    if (index) *index = (int)B->mem.size() - 1;
  });
}

int omega_batch_step(omega_batch_t B, omega_stream_t stream) {
  return guarded([&] {
    OMEGA_REQUIRE(B, "null batch");
    batch_step(B, as_stream(stream));
  });
}

int omega_batch_run(omega_batch_t B, int nsteps, omega_stream_t stream) {
  return guarded([&] {
    OMEGA_REQUIRE(B, "null batch");
    OMEGA_REQUIRE(nsteps >= 0, "nsteps must be >= 0");
    for (int i = 0; i < nsteps; ++i) batch_step(B, as_stream(stream));
  });
}

int omega_batch_set_graph(omega_batch_t B, int enable) {
  return guarded([&] {
    OMEGA_REQUIRE(B, "null batch");
    B->use_graph = enable != 0;
  });
}

int omega_batch_set_launch(omega_batch_t B, int threads, int levels_per_thread) {
  return guarded([&] {
    OMEGA_REQUIRE(B, "null batch");
    OMEGA_REQUIRE(threads == 64 || threads == 128 || threads == 256,
                  "threads must be 64, 128 or 256");
    OMEGA_REQUIRE(levels_per_thread == 1 || levels_per_thread == 2,
                  "levels_per_thread must be 1 or 2");
    B->threads = threads;
    B->N = levels_per_thread;
    B->built = false;
  });
}

int omega_batch_get_launch(omega_batch_t B, int *threads, int *levels_per_thread) {
  return guarded([&] {
    OMEGA_REQUIRE(B, "null batch");
    if (threads) *threads = B->threads;
    if (levels_per_thread) *levels_per_thread = B->N;
  });
}

// Times `reps` steps for every launch shape the ensemble admits -- levels
// per lane (2, 1), threads per block (256, 128, 64: the cells per block of
// every member) and the grid (one block per table entry, or a persistent
// grid of 4, 8 or 16 blocks per CU) -- and keeps the fastest.  The members'
// states are restored afterwards, so the ensemble continues from where it
// was.  With OMEGA_TUNE_CACHE naming a file, a choice recorded there for
// the same ensemble signature (stepper, and per member: cells, layers,
// tracers, terms, kernel kind) is applied without timing, and a new choice
// is appended to it.
int omega_batch_autotune(omega_batch_t B, int reps, double *best_ms, omega_stream_t stream) {
  return guarded([&] {
    OMEGA_REQUIRE(B && !B->mem.empty(), "batch: no members");
    hipStream_t s = as_stream(stream);
    reps = std::max(1, reps);
    bool even_layers = true;
    for (auto &mb : B->mem) even_layers = even_layers && mb->M->mesh->dm.host.nlayers % 2 == 0;
    // the ensemble's signature for the tuning cache
    std::string sig = std::to_string(B->kind) + "/" + std::to_string(B->nstages);
    for (auto &mb : B->mem) {
      const HostMesh &hm = mb->M->mesh->dm.host;
      const bool hx = hex_usable(mb->M->mesh->dm.view(), mb->M->lc) && !general_only(mb->M);
      sig += ";" + std::to_string(hm.ncells) + ":" + std::to_string(hm.nlayers) + ":" +
             std::to_string(mb->M->params.ntracers) + ":" + std::to_string(mb->M->terms) + ":" +
             (hx ? "h" : "g");
    }
    const char *cache = std::getenv("OMEGA_TUNE_CACHE");
    if (cache && *cache) {
      std::ifstream in(cache);
      std::string line;
      while (std::getline(in, line)) {
        std::istringstream ls(line);
        std::string key;
        int n = 0, t = 0, pc = -1;
        double ms = -1;
        if (!(ls >> key >> n >> t >> pc >> ms) || key != sig) continue;
        if ((n == 1 || (n == 2 && even_layers)) && (t == 64 || t == 128 || t == 256) && pc >= 0) {
          B->N = n;
          B->threads = t;
          B->per_cu = pc;
          B->built = false;
          if (best_ms) *best_ms = ms;
          return;
        }
      }
    }
    // save every member's state
    std::vector<std::array<DBuf<Real>, 3>> saved(B->mem.size());
    std::vector<int> cur(B->mem.size());
    for (size_t i = 0; i < B->mem.size(); ++i) {
      omega_state_s *S = B->mem[i]->S;
      cur[i] = S->cur;
      for (int f = 0; f < 3; ++f) {
        saved[i][f].alloc(S->n[f]);
        if (S->n[f])
          OMEGA_HIP_CHECK(hipMemcpyAsync(saved[i][f].p, S->ptr(f), S->n[f] * sizeof(Real),
                                         hipMemcpyDeviceToDevice, s));
      }
    }
    double best = -1;
    int bestN = B->N, bestT = B->threads, bestP = B->per_cu;
    for (int N : {2, 1}) {
      if (N == 2 && !even_layers) continue;
      for (int T : {256, 128, 64})
        for (int P : {0, 4, 8, 16}) {
          B->N = N;
          B->threads = T;
          B->per_cu = P;
          B->built = false;
          const double ms = time_steps(B, reps, s);
          if (best < 0 || ms < best) {
            best = ms;
            bestN = N;
            bestT = T;
            bestP = P;
          }
        }
    }
    B->N = bestN;
    B->threads = bestT;
    B->per_cu = bestP;
    B->built = false;
    // restore the states
    for (size_t i = 0; i < B->mem.size(); ++i) {
      omega_state_s *S = B->mem[i]->S;
      S->cur = cur[i];
      for (int f = 0; f < 3; ++f)
        if (S->n[f])
          OMEGA_HIP_CHECK(hipMemcpyAsync(S->ptr(f), saved[i][f].p, S->n[f] * sizeof(Real),
                                         hipMemcpyDeviceToDevice, s));
    }
    OMEGA_HIP_CHECK(hipStreamSynchronize(s));
    if (cache && *cache) {
      std::ofstream out(cache, std::ios::app);
      if (out) out << sig << " " << bestN << " " << bestT << " " << bestP << " " << best << "\n";
    }
    if (best_ms) *best_ms = best;
  });
}

int omega_batch_set_persistent(omega_batch_t B, int blocks_per_cu) {
  return guarded([&] {
    OMEGA_REQUIRE(B, "null batch");
    OMEGA_REQUIRE(blocks_per_cu >= 0 && blocks_per_cu <= 64, "blocks_per_cu must be in 0..64");
    B->per_cu = blocks_per_cu;
  });
}

int omega_batch_get_persistent(omega_batch_t B, int *blocks_per_cu) {
  return guarded([&] {
    OMEGA_REQUIRE(B && blocks_per_cu, "null argument");
    *blocks_per_cu = B->per_cu;
  });
}

int omega_batch_info(omega_batch_t B, int *nmembers, int64_t *blocks_per_launch,
                     int64_t *cell_layers) {
  return guarded([&] {
    OMEGA_REQUIRE(B, "null batch");
    if (!B->built && !B->mem.empty()) build_tables(B);
    if (nmembers) *nmembers = (int)B->mem.size();
    if (blocks_per_launch) *blocks_per_launch = B->nblocks;
    if (cell_layers) {
      int64_t n = 0;
      for (auto &mb : B->mem)
        n += (int64_t)mb->M->mesh->dm.host.ncells * mb->M->mesh->dm.host.nlayers;
      *cell_layers = n;
    }
  });
}

} // extern "C"

// ---------------------------------------------------------------------------
// Ensemble adjoint.  No operator couples one layer to another, so members
// that share a mesh, layer and tracer counts and coefficients are the
// layers of one model: member m occupies layers [m*L, (m+1)*L) of a
// stacked model with M*L layers.  The stacked model runs the checkpointed
// adjoint once, so every stage of the forward, recompute and reverse sweeps
// is one launch per kernel for the whole ensemble, the per-level arithmetic
// is that of each member alone, and the gradients of the shared
// coefficients are summed over all members by the same reductions.
// ---------------------------------------------------------------------------
namespace {

void ok(int rc) {
  if (rc != OMEGA_OK) throw Error(rc, g_last_error);
}

bool same_params(const omega_params &a, const omega_params &b) {
  return a.f0 == b.f0 && a.grav == b.grav && a.drag_coeff == b.drag_coeff &&
         a.visc_del2 == b.visc_del2 && a.visc_del4 == b.visc_del4 &&
         a.eddy_diff2 == b.eddy_diff2 && a.eddy_diff4 == b.eddy_diff4 &&
         a.ntracers == b.ntracers && a.disable_h_tendency == b.disable_h_tendency &&
         a.disable_vn_tendency == b.disable_vn_tendency;
}

// dst rows of width M*L (member m at column m*L) <-> src rows of width L
void stack_fields(omega_state_s *big, omega_state_s *mem, int m, int L, int M, bool to_big,
                  hipStream_t s) {
  for (int f = 0; f < 3; ++f) {
    const size_t rows = mem->n[f] / (size_t)L;
    if (!rows) continue;
    const size_t wb = (size_t)L * sizeof(Real), pb = (size_t)M * L * sizeof(Real);
    if (to_big)
      OMEGA_HIP_CHECK(hipMemcpy2DAsync(big->ptr(f) + (size_t)m * L, pb, mem->ptr(f), wb, wb, rows,
                                       hipMemcpyDeviceToDevice, s));
    else
      OMEGA_HIP_CHECK(hipMemcpy2DAsync(mem->ptr(f), wb, big->ptr(f) + (size_t)m * L, pb, wb, rows,
                                       hipMemcpyDeviceToDevice, s));
  }
}

struct Owned {
  omega_mesh_s *mesh = nullptr;
  omega_model_s *model = nullptr;
  omega_stepper_s *st = nullptr;
  omega_state_s *y = nullptr, *obs = nullptr, *grad = nullptr;
  ~Owned() {
    (void)omega_state_destroy(grad);
    (void)omega_state_destroy(obs);
    (void)omega_state_destroy(y);
    (void)omega_stepper_destroy(st);
    (void)omega_model_destroy(model);
    delete mesh;
  }
};

} // namespace

extern "C" {

int omega_batch_adjoint_run(omega_batch_t B, double t0, int nsteps, int n_snapshots,
                            int64_t byte_cap, const omega_state_t *obs, const double *weights,
                            double *J, omega_state_t *grads, double *param_grad,
                            omega_stream_t stream) {
  return guarded([&] {
    OMEGA_REQUIRE(B && obs && weights && J && grads, "null argument");
    const int M = (int)B->mem.size();
    OMEGA_REQUIRE(M >= 1, "batch adjoint: the batch has no members");
    const omega_model_s *M0 = B->mem[0]->M;
    const omega_mesh_s *mesh0 = M0->mesh;
    const HostMesh &h0 = mesh0->dm.host;
    const int L = h0.nlayers, ntr = M0->params.ntracers;
    const double dt = B->mem[0]->dt;
    for (int m = 0; m < M; ++m) {
      const auto &mb = *B->mem[m];
      OMEGA_REQUIRE(mb.M->mesh == mesh0,
                    "batch adjoint: all members must share one mesh (layers are stacked)");
      OMEGA_REQUIRE(same_params(mb.M->params, M0->params) && mb.M->terms == M0->terms,
                    "batch adjoint: members must share the model parameters and terms");
      OMEGA_REQUIRE(mb.dt == dt, "batch adjoint: members must share dt");
      OMEGA_REQUIRE(!mb.M->add_fn, "batch adjoint: members must not have an additional tendency");
      OMEGA_REQUIRE(mb.S->ntr == ntr, "batch adjoint: member state tracer count differs");
      OMEGA_REQUIRE(obs[m] && obs[m]->mesh == mesh0 && obs[m]->ntr == ntr,
                    "batch adjoint: observation state does not match its member");
      OMEGA_REQUIRE(grads[m] && grads[m]->mesh == mesh0 && grads[m]->ntr == ntr,
                    "batch adjoint: gradient state does not match its member");
    }
    for (int e = 0; e < h0.nedges; ++e)
      OMEGA_REQUIRE(h0.max_level_edge_top[e] >= L,
                    "batch adjoint: stacking needs every layer active on every edge");
    OMEGA_REQUIRE((int64_t)M * L <= (1 << 20), "batch adjoint: too many stacked layers");
    hipStream_t s = as_stream(stream);

    // stacked mesh, model and stepper
    Owned o;
    o.mesh = new omega_mesh_s();
    o.mesh->dm.host = h0;
    HostMesh &hb = o.mesh->dm.host;
    hb.nlayers = M * L;
    for (auto *v : {&hb.max_level_cell, &hb.max_level_edge_bot, &hb.max_level_edge_top,
                    &hb.max_level_vertex_bot, &hb.max_level_vertex_top})
      for (int &x : *v)
        if (x >= L) x = M * L;
    o.mesh->dm.build_device(mesh0->dm.ordering);
    ok(omega_model_create(o.mesh, &M0->params, &o.model));
    o.model->terms = M0->terms;
    ok(omega_stepper_create(o.model, B->kind, B->nstages, &o.st));
    ok(omega_state_create(o.mesh, ntr, &o.y));
    ok(omega_state_create(o.mesh, ntr, &o.obs));
    ok(omega_state_create(o.mesh, ntr, &o.grad));
    for (int m = 0; m < M; ++m) {
      stack_fields(o.y, B->mem[m]->S, m, L, M, true, s);
      stack_fields(o.obs, obs[m], m, L, M, true, s);
    }
    // ensemble-mean misfit
    const double w[3] = {weights[0] / M, weights[1] / M, weights[2] / M};
    ok(omega_adjoint_run(o.st, t0, dt, nsteps, n_snapshots, byte_cap, o.y, o.obs, w, J, o.grad,
                         param_grad, stream));
    for (int m = 0; m < M; ++m) stack_fields(o.grad, grads[m], m, L, M, false, s);
    OMEGA_HIP_CHECK(hipStreamSynchronize(s));
  });
}

} // extern "C"
