// Discrete adjoint of the tendency and of the time steppers.
#pragma once

#include "internal.hpp"

namespace omega_hip {

// Transposed connectivity and coefficients in device numbering, built on
// the host the first time an adjoint is requested.
struct AdjMeshView {
  const Real *ecd;    // [ne][2] cdiv of the edge's slot in cells_on_edge[0..1]
  const Real *ecke;   // [ne][2] cke of the same slots
  const Real *ecrv;   // [ne][2] crv of the edge's slot in vertices_on_edge[0..1]
  const int *voc;     // [nc][me] vertices_on_cell
  const Real *ccth;   // [nc][me] cth of the cell's slot in each of its vertices
  const int *tofs;    // [ne+1] edges that list e among their edges_on_edge
  const int *tedge;   // [nnz] source edge
  const Real *tw;     // [nnz] its weight for e
  // every cell has 6 edges, every edge 10 edges-on-edge and appears in 10
  // such lists (the regular hexagonal mesh): loops get constant trip counts
  int fixed;
  // slot-uniform form (regular hexagonal numbering, ecd/ecke/ecrv repeating
  // every 3 edges, ccth the same for every cell, the mesh's coefficients
  // equal at each stencil slot; lists of 10 transposed entries per edge)
  int uni;
  int ufirst[6]; // cell slot j: the cell is cells_on_edge[0] of its edge
  int ukslot[6]; // cell slot j: the edge's class (edge index mod 3)
  int uvsg[6];   // vertex 2c + x, slot j: the vertex is vertices_on_edge[1]
};

struct AdjLayout {
  DBuf<Real> ecd, ecke, ecrv, ccth, tw;
  DBuf<int> voc, tofs, tedge;
  DBuf<Real> area_e;          // [ne] dc*dv/2, the area an edge accounts for
  std::vector<int> cslot_h;   // [ne][2] slot of the edge in cells_on_edge[0..1]
  bool built = false;
  bool fixed = false;
  bool uni = false;
  int ufirst[6] = {0}, ukslot[6] = {0}, uvsg[6] = {0};
  AdjMeshView view() const {
    AdjMeshView v{ecd.p, ecke.p, ecrv.p, voc.p, ccth.p, tofs.p, tedge.p, tw.p,
                  fixed ? 1 : 0, uni ? 1 : 0, {0}, {0}, {0}};
    for (int j = 0; j < 6; ++j) {
      v.ufirst[j] = ufirst[j];
      v.ukslot[j] = ukslot[j];
      v.uvsg[j] = uvsg[j];
    }
    return v;
  }
};

void build_adj_layout(const DeviceMesh &dm, AdjLayout &a);

// Model-dependent transposed tracer coefficients.
struct AdjCoef {
  const Real *ect2;   // [ne][2] ct2 of the edge's slot in cells_on_edge[0..1]
  const Real *ect4;   // [ne][2] ct4 likewise
  int uniform;        // both repeat every 3 edges, bit for bit
};

// Cotangent buffers of one prognostic state (device numbering).
struct CotIO {
  const Real *h, *vn, *tr; // cotangent of the tendency (input)
  Real *oh, *ovn, *otr;    // cotangent of the state (accumulated output)
};

// Adjoint intermediates (cotangents of diagnostics).
struct AdjScratch {
  Real *sbar_e;   // [ne][L] cotangent of the edge PV sum Q_e + P_e
  Real *zbar_v;   // [nv][L] rvort
  Real *tbar_v;   // [nv][L] thickness at vertices
  Real *kbar_c;   // [nc][L] kinetic energy
  Real *dbar_c;   // [nc][L] divergence
  Real *gbar_c;   // [ntr][nc][L] tracer Laplacian (tracer del4)
  Real *fbar_e;   // [ne][L] edge thickness
  Real *d2bar_c;  // [nc][L] momentum del4: del2 divergence
  Real *z2bar_v;  // [nv][L] momentum del4: del2 vorticity
  Real *e2bar_e;  // [ne][L] momentum del4: del2 edge
  Real *partial;  // [4][nparts] parameter-gradient block partials
  Real *pgrad;    // [4] accumulated parameter gradient
  Real *se_e;     // [ne][L] edge PV sum Q_e + P_e at the linearisation state
  Real *fu_e;     // [ne][L] edge thickness flux F_e u_e at the linearisation state
  Real *px_e;     // [ne][L] (A_e + S_e B_e)/2 with A_e = sum_t w_te kbar_t S_t and
                  // B_e = sum_t w_te kbar_t over the edges t listing e
  Real *xa_e;     // [ntr][ne][L] tracer advection share of each cell of the edge,
                  // -(cd0 b0 + cd1 b1) F_e u_e / 2 (no tracer diffusion only)
  Real *own_e;    // [ne][L] sum_j w_ej F_j u_j over the edge's edges-on-edge (given
                  // with pre rows only; otherwise the PV pass forms it)
};

// What the VJP does with its result v = vs * J^T k for each state element:
// out = base + v (base null: out += v); with rk4 set also kbo = db*gy + da*v,
// the next tendency cotangent of the RK4 reverse sweep; with lsrk set
// kbo = da*k + db*out, that of the low-storage RK reverse sweep (kbo must not
// alias the input k).  vs scales the result (and the coefficient gradient),
// so a scaled cotangent need not be formed.
struct AdjEpi {
  int rk4 = 0, lsrk = 0;
  Real db = 0, da = 0;
  Real vs = 1;
  const Real *gy[3] = {nullptr, nullptr, nullptr};
  Real *kbo[3] = {nullptr, nullptr, nullptr};
  const Real *base[3] = {nullptr, nullptr, nullptr};
};

struct AdjArgs {
  int terms, do_h, do_vn, ntr;
  Real grav;
  const Real *h, *vn, *tr;  // state at which the tendency is linearised
  CotIO c;
  int nparts;               // partial slots available per parameter
  AdjEpi epi;
  // coefficient-gradient partials shared by the VJPs of one step adjoint:
  // part_add adds vs * this VJP's block sums to the slots an earlier VJP
  // filled instead of storing them, part_keep leaves them for a later VJP
  // to reduce (no reduction launch); with either set the block sums carry
  // the vs scaling and the reduction adds them unscaled
  int part_add, part_keep;
};

// pre: se_e, fu_e and own_e of s already hold the rows at the linearisation
// state (written by the stage kernel there), so the pre-pass is skipped and
// the PV pass reads own_e
void launch_tendency_vjp(const MeshView &m, const AdjMeshView &am, const Launch &lc,
                         const ModelCoef &mc, const AdjCoef &ac, const AdjArgs &a,
                         const DiagOut &d, const AdjScratch &s, hipStream_t st,
                         bool pre = false);
int adj_partial_slots(const MeshView &m, const Launch &lc);

// y[i] = v
void launch_fill(Real *y, Real v, size_t n, hipStream_t s);
// y = alpha*x + beta*z (y may alias x or z)
void launch_lincomb(Real *y, Real alpha, const Real *x, Real beta, const Real *z, size_t n,
                    hipStream_t s);
// Weighted least-squares misfit of one field: for entity i, level k,
// r = y - obs, cot = w*area[i]*r, partial sums of w*area[i]*r^2/2 into
// partial[0..misfit_blocks(n*L)); returns nothing (fold with launch_sum).
int misfit_blocks(size_t n);
void launch_misfit(const Real *y, const Real *obs, const Real *area, int nent, int L, int nrep,
                   Real w, Real *cot, Real *partial, hipStream_t s);
// out[0] (+)= fixed-order sum of partial[0..n)
void launch_sum(const Real *partial, int n, Real *out, int accumulate, hipStream_t s);

} // namespace omega_hip
