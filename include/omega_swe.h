/*
 * omega_swe.h -- C ABI of the MI355X shallow-water hot path.
 *
 * Drop-in boundary for mwarusz/omega_playground's mesh / state / tendency /
 * time-stepper API.  Every entry point names the reference interface it
 * replaces (file:line in omega_playground).  The header-only C++ classes
 * under omega_playground_amd/cpp (namespace omega, headers laid out as the
 * reference's src/) call these functions; Python reaches them through
 * ctypes.
 *
 * Conventions
 *  - Every function returns OMEGA_OK (0) or an error code; the message of the
 *    last failure on the calling thread is omega_last_error().  Nothing
 *    throws across this boundary.
 *  - Host arrays handed in or out use the reference's original entity
 *    numbering and LayoutRight shapes (level fastest):
 *      h[ncells][nlayers], vn[nedges][nlayers], tr[ntracers][ncells][nlayers].
 *  - Device arrays (omega_state_device_ptrs, the additional-tendency hook,
 *    omega_mesh_device_array) use the library's device numbering: cells in
 *    space-filling-curve order, edges and vertices grouped by owning cell.
 *    omega_mesh_permutation() maps between the two.
 *  - Launching calls take a hipStream_t (passed as void*; NULL is the null
 *    stream) and are asynchronous with respect to the host unless stated.
 *  - All arithmetic is IEEE binary64; no kernel uses atomics, so results are
 *    bit-identical from run to run.
 */
#ifndef OMEGA_SWE_H
#define OMEGA_SWE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OMEGA_OK 0
#define OMEGA_ERR_ARG 1      /* invalid argument or size */
#define OMEGA_ERR_HIP 2      /* HIP runtime failure */
#define OMEGA_ERR_STATE 3    /* object used in an invalid state */
#define OMEGA_ERR_NOTFOUND 4 /* unknown array name */
#define OMEGA_ERR_CALLBACK 5 /* user hook returned non-zero */

typedef struct omega_mesh_s *omega_mesh_t;
typedef struct omega_model_s *omega_model_t;
typedef struct omega_state_s *omega_state_t;
typedef struct omega_stepper_s *omega_stepper_t;
typedef void *omega_stream_t; /* hipStream_t */

/* Message of the last error on this thread ("" if none). */
const char *omega_last_error(void);
/* Library version string. */
const char *omega_version(void);

/* ----------------------------------------------------------------------
 * Mesh
 * -------------------------------------------------------------------- */

/* Cell ordering on the device. */
#define OMEGA_ORDER_ORIGINAL 0 /* reference numbering */
#define OMEGA_ORDER_HILBERT 1  /* cells along a Hilbert curve (default) */
#define OMEGA_ORDER_MORTON 2   /* cells along a Morton (Z) curve */

/* Replaces PlanarHexagonalMesh(nx, ny, dc, nlayers) + MPASMesh::finalize_mesh
 * (src/mesh/planar_hexagonal_mesh.hpp:17-18,
 *  src/mesh/planar_hexagonal_mesh.cpp:8-398, src/mesh/mpas_mesh.cpp:5-57).
 * Builds the doubly periodic hexagonal mesh on the host in the reference's
 * numbering and uploads it reordered.  nlayers is used as given (no padding
 * to a SIMD width). */
int omega_mesh_create_planar_hex(int nx, int ny, double dc, int nlayers, int ordering,
                                 omega_mesh_t *out);

/* Replaces constructing an MPASMesh from its arrays followed by
 * MPASMesh::finalize_mesh (src/mesh/mpas_mesh.hpp:7-71,
 * src/mesh/mpas_mesh.cpp:5-57).  Arrays are host, original numbering.
 * edge_sign_on_cell / kite_index_on_cell / edge_sign_on_vertex are derived as
 * finalize_mesh does.  edge_mask is 1 below max_level_edge_top (NULL: all
 * levels); mesh_scaling_del2/4 NULL means 1. */
typedef struct {
  int ncells, nedges, nvertices, nlayers;
  int max_edges;                 /* row length of *_on_cell arrays (>= 3) */
  int max_edges2;                /* row length of edges_on_edge / weights_on_edge */
  const int *nedges_on_cell;     /* [ncells] */
  const int *edges_on_cell;      /* [ncells][max_edges] */
  const int *vertices_on_cell;   /* [ncells][max_edges] */
  const int *nedges_on_edge;     /* [nedges] */
  const int *edges_on_edge;      /* [nedges][max_edges2] */
  const int *cells_on_edge;      /* [nedges][2] */
  const int *vertices_on_edge;   /* [nedges][2] */
  const int *edges_on_vertex;    /* [nvertices][3] */
  const int *cells_on_vertex;    /* [nvertices][3] */
  const double *area_cell;       /* [ncells] */
  const double *dc_edge;         /* [nedges] */
  const double *dv_edge;         /* [nedges] */
  const double *weights_on_edge; /* [nedges][max_edges2] */
  const double *area_triangle;   /* [nvertices] */
  const double *kiteareas_on_vertex; /* [nvertices][3] */
  const double *mesh_scaling_del2;   /* [nedges] or NULL */
  const double *mesh_scaling_del4;   /* [nedges] or NULL */
  const int *max_level_edge_top;     /* [nedges] or NULL */
  /* optional geometry (kept for host queries; may be NULL) */
  const double *x_cell, *y_cell, *x_edge, *y_edge, *angle_edge, *x_vertex, *y_vertex;
} omega_mpas_mesh_desc;

int omega_mesh_create_mpas(const omega_mpas_mesh_desc *desc, int ordering, omega_mesh_t *out);

int omega_mesh_destroy(omega_mesh_t mesh);

/* counts[0..5] = ncells, nedges, nvertices, nlayers, nx, ny (nx = ny = 0 for
 * a general MPAS mesh); reals[0..2] = dc, period_x, period_y. */
int omega_mesh_info(omega_mesh_t mesh, int *counts, double *reals);

/* Host copy of an MPASMesh field in the original numbering, by the
 * reference's field name without the m_ prefix (e.g. "edges_on_cell",
 * "x_cell", "weights_on_edge"; mpas_mesh.hpp:7-71).  dtype: 0 int32,
 * 1 float64.  The pointer stays valid for the mesh's lifetime. */
int omega_mesh_host_array(omega_mesh_t mesh, const char *name, const void **ptr, int *dtype,
                          int64_t *n0, int64_t *n1);

/* Device copy of a geometry field in device numbering ("x_cell", "y_cell",
 * "x_edge", "y_edge", "angle_edge", "x_vertex", "y_vertex", "area_cell"). */
int omega_mesh_device_array(omega_mesh_t mesh, const char *name, const double **dptr,
                            int64_t *n);

/* new2old[i] = original index of device entity i.  kind: 0 cells,
 * 1 edges, 2 vertices. */
int omega_mesh_permutation(omega_mesh_t mesh, int kind, int *new2old);

/* 1 if every edge of the mesh follows the regular-hexagon stencil pattern
 * the specialised kernels are compiled for, else 0 (host-side check; no
 * reference counterpart). */
int omega_mesh_regular_hex(omega_mesh_t mesh, int *is_regular);

/* ----------------------------------------------------------------------
 * State (ShallowWaterState, src/model/shallow_water_state.hpp:9-16)
 * -------------------------------------------------------------------- */

/* Replaces ShallowWaterState(const MPASMesh*, Int ntracers)
 * (src/model/shallow_water_state.cpp:7-14).  Device arrays are zeroed, as
 * Kokkos views are. */
int omega_state_create(omega_mesh_t mesh, int ntracers, omega_state_t *out);
int omega_state_destroy(omega_state_t state);

/* Host <-> device copies in the original numbering (the drop-in for
 * Kokkos::deep_copy between host mirrors and the state views).  NULL skips a
 * field.  Download synchronises the stream before returning. */
int omega_state_upload(omega_state_t state, const double *h, const double *vn, const double *tr,
                       omega_stream_t stream);
int omega_state_download(omega_state_t state, double *h, double *vn, double *tr,
                         omega_stream_t stream);

/* Current device buffers (device numbering).  Low-storage RK alternates two
 * buffers per field, so re-query after stepping. */
int omega_state_device_ptrs(omega_state_t state, double **h, double **vn, double **tr);

/* deep_copy(dst, src) of all fields (same mesh and tracer count). */
int omega_state_copy(omega_state_t dst, omega_state_t src, omega_stream_t stream);

/* ----------------------------------------------------------------------
 * Model (ShallowWaterModel, src/model/shallow_water_model.hpp:12-51)
 * -------------------------------------------------------------------- */

/* Field order of ShallowWaterParams (src/model/shallow_water_params.hpp:7-18). */
typedef struct {
  double f0;
  double grav;
  double drag_coeff; /* carried, unused as in the reference unless the
                        bottom-drag term (omega_model_set_terms bit 7) is on */
  double visc_del2;
  double visc_del4;
  double eddy_diff2;
  double eddy_diff4;
  int ntracers;
  int disable_h_tendency;
  int disable_vn_tendency;
} omega_params;

/* Replaces ShallowWaterModel(MPASMesh*, const ShallowWaterParams&)
 * (src/model/shallow_water_model.cpp:5-41): terms are enabled from the
 * parameters exactly as there; f_vertex = f_edge = f0. */
int omega_model_create(omega_mesh_t mesh, const omega_params *params, omega_model_t *out);
/* Replaces LinearShallowWaterModel(MPASMesh*, const LinearShallowWaterParams&)
 * (src/shallow_water_hierarchical.cpp:676-773, not compiled in the
 * reference): the shallow-water equations linearised about a resting layer
 * of thickness h0, h_t = -h0 div(vn) (:680-707) and vn_t = f_edge * (TRiSK
 * tangential velocity) - g grad(h) (:709-741), with f_edge = f0 and no
 * tracers (params->ntracers must be 0; the mixing coefficients are unused).
 * Steppers, hooks, increment mode, field tendencies and the integrals
 * (energy with h0*K, :743-773) work as for omega_model_create; the adjoint
 * and ensembles do not take linear models. */
int omega_model_create_linear(omega_mesh_t mesh, const omega_params *params, double h0,
                              omega_model_t *out);
int omega_model_destroy(omega_model_t model);

/* Overwrites m_f_vertex (host array, original numbering). */
int omega_model_set_f_vertex(omega_model_t model, const double *f_vertex);

/* Replaces the virtual ShallowWaterModel::additional_tendency
 * (src/model/shallow_water_model.hpp:42-44, called at
 * shallow_water_model.cpp:344-345).  The hook runs on the same stream after
 * the built-in terms of every tendency evaluation and must only enqueue
 * work on that stream.  Pointers are device buffers in device numbering.
 * A non-zero return aborts the call with OMEGA_ERR_CALLBACK.  fn = NULL
 * removes the hook. */
typedef int (*omega_additional_tendency_fn)(void *user, double *h_tend, double *vn_tend,
                                            const double *h, const double *vn, double t,
                                            omega_stream_t stream);
int omega_model_set_additional_tendency(omega_model_t model, omega_additional_tendency_fn fn,
                                        void *user);

#define OMEGA_ADD_REPLACE 0   /* AddMode::replace */
#define OMEGA_ADD_INCREMENT 1 /* AddMode::increment */

/* Replaces ShallowWaterModel::compute_tendency(tend, state, t, add_mode)
 * (src/model/shallow_water_model.cpp:313-347). */
int omega_model_compute_tendency(omega_model_t model, omega_state_t tend, omega_state_t state,
                                 double t, int add_mode, omega_stream_t stream);

/* Replaces ShallowWaterModel::compute_h_tendency, compute_vn_tendency and
 * compute_tr_tendency (src/model/shallow_water_model.hpp:35-40,
 * shallow_water_model.cpp:43-311): field 0 h, 1 vn, 2 tracers.  Only that
 * field of tend is written.  The diagnostics are computed from state first
 * (the reference expects compute_tendency's aux pass to have run); the
 * additional-tendency hook is not called, as there. */
int omega_model_compute_field_tendency(omega_model_t model, int field, omega_state_t tend,
                                       omega_state_t state, int add_mode,
                                       omega_stream_t stream);

/* omega_model_compute_tendency without any allocation.  No counterpart in
 * the reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  The diagnostics the tendency kernels pass between launches
 * live in one static device array of 64 MiB per GPU (held until the process
 * ends) instead of the model's buffers; models whose enabled terms need more
 * (about (12 + ntracers) * ncells * nlayers doubles with every mixing term
 * on) get OMEGA_ERR_ARG.  The results are the bits of
 * omega_model_compute_tendency; the model's buffers and auxiliary-state
 * record are not touched.  The array keeps the last call's contents and is
 * shared by every call on the GPU, so calls that overlap on one GPU race.
 * The additional-tendency hook, when set, runs as in compute_tendency. */
int omega_model_compute_tendency_noalloc(omega_model_t model, omega_state_t tend,
                                         omega_state_t state, double t, int add_mode,
                                         omega_stream_t stream);

/* Replaces ShallowWaterAuxiliaryState::compute(h, vn, tr, f_vertex, mesh)
 * (src/model/shallow_water_auxstate.cpp:12-50), the aux pass that
 * compute_tendency runs first (shallow_water_model.cpp:317-319): vertex
 * vorticity and thickness, cell kinetic energy and divergence, and what the
 * enabled mixing terms read (tracer Laplacian, momentum del4 pre-pass), into
 * the model's buffers.  h, vn, tr are device arrays (device numbering; tr may
 * be NULL without tracers).  The model keeps the three pointers as the
 * auxiliary state's source: the edge thickness and normalised tracers the
 * field tendencies below read are formed from them, so they must stay
 * unchanged until those calls have run.  Every tendency evaluation and step
 * recomputes the auxiliary state from its own input, as the reference's
 * compute_tendency does. */
int omega_model_compute_aux(omega_model_t model, const double *h, const double *vn,
                            const double *tr, omega_stream_t stream);

/* Replace ShallowWaterModel::compute_h_tendency(Real2d h_tend_cell,
 * RealConst2d h_cell, RealConst2d vn_edge, AddMode),
 * compute_vn_tendency(Real2d vn_tend_edge, RealConst2d h_cell,
 * RealConst2d vn_edge, AddMode) and compute_tr_tendency(Real3d tr_tend_cell,
 * RealConst3d tr_cell, RealConst2d vn_edge, AddMode)
 * (src/model/shallow_water_model.hpp:35-40, shallow_water_model.cpp:43-311)
 * on device arrays in device numbering.  As in the reference they read the
 * auxiliary state of the last compute_aux / compute_tendency / step:
 * the thickness flux and the normalised tracers come from its source arrays
 * (h_cell is not read by compute_h_tendency, tr_cell not by
 * compute_tr_tendency), and compute_vn_tendency's h must be the auxiliary
 * state's own thickness array (OMEGA_ERR_ARG otherwise).  Without an
 * auxiliary state they fail with OMEGA_ERR_ARG.  The linear model reads its
 * arguments directly. */
int omega_model_compute_h_tendency(omega_model_t model, double *h_tend, const double *h,
                                   const double *vn, int add_mode, omega_stream_t stream);
int omega_model_compute_vn_tendency(omega_model_t model, double *vn_tend, const double *h,
                                    const double *vn, int add_mode, omega_stream_t stream);
int omega_model_compute_tr_tendency(omega_model_t model, double *tr_tend, const double *tr,
                                    const double *vn, int add_mode, omega_stream_t stream);

/* Host copy, in the original numbering, of an array of the auxiliary state
 * (m_aux_state fields, src/model/shallow_water_auxstate.hpp:10-22 and
 * the auxiliary_vars headers): "rvort_vertex" [nvertices][L], "inv_thick_vertex"
 * (1 / the vertex thickness of vorticity_aux.hpp:103-104) [nvertices][L],
 * "ke_cell", "vel_div_cell" [ncells][L], "tracer_del2_cell"
 * [ntracers][ncells][L].  Only arrays the enabled terms need are computed;
 * the others fail with OMEGA_ERR_ARG or hold stale values.  Synchronises. */
int omega_model_aux_download(omega_model_t model, const char *name, double *out,
                             omega_stream_t stream);

/* Term switches (bit 0 PV flux, 1 KE gradient, 2 SSH gradient, 3 momentum
 * del2, 4 momentum del4, 5 tracer del2, 6 tracer del4, 7 bottom drag); the
 * default is what the constructor enabled.  The reference toggles the same
 * m_enabled flags.  Bit 7 is opt-in only: the reference's active model
 * carries m_drag_coeff but applies no drag (shallow_water_params.hpp:10);
 * its formula survives only as a comment at
 * src/shallow_water_hierarchical.cpp:479-485, which this term follows:
 * on an edge's deepest active level k = max_level_edge_top - 1,
 * vn_tend += -drag_coeff * sqrt(ke_cell0 + ke_cell1) * vn * (1 / h_edge)
 * with h_edge = 0.5 * (h_cell0 + h_cell1), added after the SSH gradient and
 * before the mixing terms.  Models with it run the general kernels; the
 * adjoint rejects them. */
int omega_model_set_terms(omega_model_t model, int mask);
int omega_model_get_terms(omega_model_t model, int *mask);

/* Launch shape: cells per block (0: fill 256 threads), levels per thread
 * (1 or 2), whether blocks are swizzled so each XCD walks a contiguous
 * stretch of the cell curve (-1 keeps it), and levels per block (0: all).
 * No reference counterpart (Kokkos chose its tiles, src/common.hpp:78-104);
 * these are the autotuner's knobs.  Non-positive values keep the current
 * setting unless stated. */
int omega_model_set_launch(omega_model_t model, int rows_per_block, int levels_per_thread,
                           int xcd_swizzle, int levels_per_block);
int omega_model_get_launch(omega_model_t model, int *rows_per_block, int *levels_per_thread,
                           int *xcd_swizzle, int *levels_per_block);
/* Times reps tendency evaluations of state for each launch shape (cells per
 * block: the default fill, 4 or 2; levels per thread 2 or 1; XCD swizzle on
 * or off), keeps the fastest; *best_ms = its time per evaluation; then
 * times the diagnostics launch alone for its own cells per block (as the
 * tendency launch, the default fill, 4 or 2).  With the environment
 * variable OMEGA_TUNE_CACHE naming a file, a choice recorded there for the
 * same model shape (cells, layers, tracers, terms, kernel family) is
 * applied without timing, and a new choice is appended.
 * omega_model_set_launch makes the diagnostics follow the tendency launch
 * again.  No reference counterpart (Kokkos picks its tiles,
 * src/common.hpp:78-104); results are the same bits for every shape. */
int omega_model_autotune(omega_model_t model, omega_state_t state, int reps, double *best_ms,
                         omega_stream_t stream);

/* Kernel family: 0 (default) uses the pattern-specialised kernels when the
 * mesh is a regular hexagonal mesh verified against the built-in stencil
 * pattern and nlayers is a multiple of 64 (or 16 / 32 with two levels per
 * thread), 1 forces the general kernels.  omega_model_kernel_variant
 * reports what runs: 6 linear model, 2 specialised, 1 general.  No reference
 * counterpart (performance control only; both give the same bits). */
int omega_model_set_variant(omega_model_t model, int variant);
int omega_model_kernel_variant(omega_model_t model, int *in_use);
/* Forms of the two-cells-per-wavefront kernels this model runs: bit 0 set
 * when every coefficient table (mesh and model) is bit-identical at each
 * stencil slot, so a wavefront's two cells share one scalar read of each;
 * bit 1 set when all [rows][levels] arrays have < 2^24 rows and < 4 GiB,
 * so row addresses are 32-bit offsets from the array base; bit 2 set when
 * the tendency VJP runs its slot-uniform form (regular hexagonal
 * numbering, an even layer count, every adjoint table bit-identical at
 * each stencil slot; known after the model's first adjoint call); bit 3
 * set when the pair tendency kernel (64k layers) runs a form compiled for
 * the model's term set (7, 15, 79 or 127, both tendencies on); bit 4 set
 * when a step adjoint's stage recomputation also writes the rows its
 * tendency VJPs linearise about (edge PV sums, thickness fluxes and their
 * edges-on-edge sums; the one-cell-per-row kernels at 16 or 32 layers, and
 * the pair kernel with bits 0 and 1), so those VJPs skip their pre-pass.  No reference counterpart (introspection
 * for tests and tuning; the forward results are the same bits either way,
 * and so are the VJP's state cotangents). */
int omega_model_kernel_flags(omega_model_t model, int *flags);
/* Which of those forms the model may use (bit 0 shared coefficients, bit 1
 * 32-bit offsets, bit 2 the compiled-in term sets, bit 3 the VJP rows
 * written by the stage recomputation; default 15); clearing a bit forces
 * the per-cell, 64-bit, run-time-terms or separate pre-pass form (tests,
 * tuning). */
int omega_model_set_kernel_forms(omega_model_t model, int forms);
/* ----------------------------------------------------------------------
 * Time steppers (src/time_integration/time_stepper.hpp:8-38)
 * -------------------------------------------------------------------- */

#define OMEGA_STEPPER_RK4 0
#define OMEGA_STEPPER_LSRK 1

/* Replaces RK4Stepper(model) (src/time_integration/rk4.cpp:8-25) and
 * LSRKStepper(model, nstages) (src/time_integration/lsrk.cpp:5-34). */
int omega_stepper_create(omega_model_t model, int kind, int nstages, omega_stepper_t *out);
int omega_stepper_destroy(omega_stepper_t stepper);

/* m_rka, m_rkb, m_rkc of RK4Stepper / LSRKStepper (rk4.cpp:14-25,
 * lsrk.cpp:16-33); each non-NULL array receives nstages values (RK4's a and
 * c have 3 entries; the 4th is 0). */
int omega_stepper_coeffs(omega_stepper_t stepper, int *nstages, double *a, double *b, double *c);

/* Replaces TimeStepper::do_step(t, dt, state)
 * (src/time_integration/rk4.cpp:28-104, src/time_integration/lsrk.cpp:36-89).
 * Without an additional-tendency hook each stage is one diagnostics launch
 * and one fused tendency + stage-update launch. */
int omega_stepper_do_step(omega_stepper_t stepper, double t, double dt, omega_state_t state,
                          omega_stream_t stream);

/* nsteps calls of do_step with t = t0 + (i + t_offset_steps) * dt.  With
 * graphs enabled the step is captured once per (dt, state, buffer parity)
 * and replayed. */
int omega_stepper_run(omega_stepper_t stepper, double t0, double dt, int nsteps,
                      int t_offset_steps, omega_state_t state, omega_stream_t stream);
int omega_stepper_set_graph(omega_stepper_t stepper, int enable);

/* ----------------------------------------------------------------------
 * Diagnostics (src/model/shallow_water_diagnostics.cpp:5-89)
 * Deterministic two-level reductions; results are written to *out after
 * the stream is synchronised.
 * -------------------------------------------------------------------- */
int omega_energy_integral(omega_model_t model, omega_state_t state, double *out,
                          omega_stream_t stream);
int omega_mass_integral(omega_model_t model, omega_state_t state, double *out,
                        omega_stream_t stream);
int omega_circulation_integral(omega_model_t model, omega_state_t state, double *out,
                               omega_stream_t stream);

/* ----------------------------------------------------------------------
 * Discrete adjoint.  No reference counterpart: the reference has no
 * adjoint; these differentiate its forward operators
 * (src/model/shallow_water_model.cpp:43-347) and steppers
 * (src/time_integration/rk4.cpp:28-104, lsrk.cpp:36-89) exactly, in the
 * same gather-only, atomic-free style as the forward kernels.  Cotangents
 * are omega_state_t objects of the model's mesh.  param_grad, when not NULL,
 * points to 4 host doubles to which the gradient with respect to
 * (visc_del2, visc_del4, eddy_diff2, eddy_diff4) is ADDED; such calls
 * synchronise the stream.  A term contributes only while it is enabled
 * (omega_model_set_terms can enable a term whose coefficient is zero to
 * differentiate with respect to that coefficient).  Models with an
 * additional-tendency hook are rejected (OMEGA_ERR_ARG).
 * -------------------------------------------------------------------- */
typedef struct omega_tape_s *omega_tape_t;

/* state_cot += J(state)^T tend_cot, J the Jacobian of compute_tendency. */
int omega_model_tendency_vjp(omega_model_t model, omega_state_t state, omega_state_t tend_cot,
                             omega_state_t state_cot, double *param_grad,
                             omega_stream_t stream);

/* cot <- (d y_{n+1} / d y_n)^T cot for one do_step(t, dt) from y_n = state.
 * The stage states are recomputed from state with the forward kernels.
 * param_grad[0..3] += the mixing-coefficient gradient (the call waits for
 * the stream to read it); param_grad = NULL: not read back, no wait. */
int omega_stepper_step_vjp(omega_stepper_t stepper, double t, double dt, omega_state_t state,
                           omega_state_t cot, double *param_grad, omega_stream_t stream);

/* Binomial checkpointing over nsteps steps with at most n_snapshots stored
 * states (the initial state included) and, when byte_cap > 0, at most
 * byte_cap bytes of snapshots.  The forward sweep advances state from y_0 to
 * y_nsteps; the reverse sweep maps dJ/dy_nsteps (cot on entry) to dJ/dy_0
 * (cot on exit), recomputing intermediate states from the snapshots with the
 * minimal number of forward steps (Griewank's binomial schedule). */
int omega_tape_create(omega_stepper_t stepper, int nsteps, int n_snapshots, int64_t byte_cap,
                      omega_tape_t *out);
int omega_tape_destroy(omega_tape_t tape);
int omega_tape_forward(omega_tape_t tape, double t0, double dt, omega_state_t state,
                       omega_stream_t stream);
int omega_tape_reverse(omega_tape_t tape, omega_state_t cot, double *param_grad,
                       omega_stream_t stream);
/* Snapshots in use, forward steps run by the last forward + reverse sweeps
 * (the step adjoints' own stage recomputation not counted), snapshot bytes. */
int omega_tape_info(omega_tape_t tape, int *n_snapshots, int64_t *forward_steps,
                    int64_t *snapshot_bytes);
/* Minimal forward steps to reverse nsteps steps with n_snapshots snapshots. */
int64_t omega_revolve_cost(int64_t nsteps, int n_snapshots);

/* Least-squares misfit J = sum_f w_f/2 sum_i a_i sum_k (y - obs)^2 over
 * f = h (cell areas), vn (edge areas dc*dv/2), tracers (cell areas);
 * cot = dJ/dy.  weights[3] = (w_h, w_vn, w_tr).  Synchronises. */
int omega_misfit(omega_model_t model, omega_state_t state, omega_state_t obs,
                 const double *weights, omega_state_t cot, double *J, omega_stream_t stream);

/* J(y_nsteps) of omega_misfit after nsteps steps from state (left unchanged),
 * grad = dJ/dy_0 (overwritten), param_grad += dJ/d(coefficients). */
int omega_adjoint_run(omega_stepper_t stepper, double t0, double dt, int nsteps,
                      int n_snapshots, int64_t byte_cap, omega_state_t state, omega_state_t obs,
                      const double *weights, double *J, omega_state_t grad, double *param_grad,
                      omega_stream_t stream);

/* ----------------------------------------------------------------------
 * Ensembles.  No reference counterpart: the reference steps one model at a
 * time.  Members are (model, state, dt) triples with their own meshes,
 * layer counts, tracer counts and coefficients; one batch step advances
 * every member by one do_step of the batch's stepper kind, with each stage
 * phase (diagnostics, momentum del4 pre-pass, tendency + update) a single
 * launch over all members.  Each member's state ends bit-identical to
 * stepping it alone.  Members must not have an additional-tendency hook.
 * -------------------------------------------------------------------- */
typedef struct omega_batch_s *omega_batch_t;

int omega_batch_create(int kind, int nstages, omega_batch_t *out);
int omega_batch_destroy(omega_batch_t batch);
/* Adds a member; *index (may be NULL) receives its position. */
int omega_batch_add(omega_batch_t batch, omega_model_t model, omega_state_t state, double dt,
                    int *index);
int omega_batch_step(omega_batch_t batch, omega_stream_t stream);
int omega_batch_run(omega_batch_t batch, int nsteps, omega_stream_t stream);
/* Capture the ensemble step in a HIP graph (one per buffer parity). */
int omega_batch_set_graph(omega_batch_t batch, int enable);
/* Launch shape of the ensemble kernels: threads per block (64, 128, 256)
 * and levels per lane (1 or 2).  Each stage phase (diagnostics, momentum
 * del4 pre-pass, tendency + update) is one launch over every member; members
 * on the regular hexagonal mesh run the pattern-specialised operators and
 * the others the general ones inside it. */
int omega_batch_set_launch(omega_batch_t batch, int threads, int levels_per_thread);
int omega_batch_get_launch(omega_batch_t batch, int *threads, int *levels_per_thread);
/* Persistent grid: 0 (default) launches one block per table entry (a
 * block covers the cells-per-block of one member); n > 0 launches n blocks
 * per CU, each walking a contiguous run of entries (used when that is fewer
 * blocks than entries). */
int omega_batch_set_persistent(omega_batch_t batch, int blocks_per_cu);
int omega_batch_get_persistent(omega_batch_t batch, int *blocks_per_cu);
/* Times reps steps per admissible launch shape (levels per lane 2/1,
 * threads per block 256/128/64, which set every member's cells per block,
 * and a grid of one block per entry or 4/8/16 persistent blocks per CU),
 * keeps the fastest and restores the members' states; *best_ms = that
 * shape's time per step.  When the environment variable OMEGA_TUNE_CACHE
 * names a file, a choice recorded there for the same ensemble (stepper and
 * each member's cells, layers, tracers, terms and kernel kind) is applied
 * without timing, and a new choice is appended to the file. */
int omega_batch_autotune(omega_batch_t batch, int reps, double *best_ms, omega_stream_t stream);
int omega_batch_info(omega_batch_t batch, int *nmembers, int64_t *blocks_per_launch,
                     int64_t *cell_layers);
/* Ensemble adjoint with checkpointing under a memory cap.  The members
 * (all on one mesh, with one layer count, tracer count, parameter set and
 * dt; the batch's stepper kind) are stacked along the layer axis, which no
 * operator couples, into one model of nmembers*nlayers layers, so each
 * stage of the forward, recompute and reverse sweeps is one launch per
 * kernel for the whole ensemble.  J = (1/M) sum_m J_m, J_m the misfit of
 * omega_misfit of member m against obs[m] after nsteps steps; grads[m] =
 * dJ/d(member m's initial state) (overwritten); param_grad[0..4) (may be NULL)
 * += dJ/d(shared coefficients).  At most n_snapshots ensemble states and,
 * when byte_cap > 0, byte_cap bytes of them are stored; the members' states
 * are left unchanged.  Synchronises. */
int omega_batch_adjoint_run(omega_batch_t batch, double t0, int nsteps, int n_snapshots,
                            int64_t byte_cap, const omega_state_t *obs, const double *weights,
                            double *J, omega_state_t *grads, double *param_grad,
                            omega_stream_t stream);

/* ----------------------------------------------------------------------
 * Utilities
 * -------------------------------------------------------------------- */

/* dst[i][k] += alpha * src[i][k] for a device array dst[n][ncols] and a host
 * array src given in the ORIGINAL numbering of entity kind (0 cells, 1
 * edges, 2 vertices) of the mesh; used by host-side additional-tendency
 * hooks.  ncols may be 1 (broadcast over the levels) or nlayers. */
int omega_add_host_field(omega_mesh_t mesh, int kind, double *dst, const double *src, int ncols,
                         double alpha, omega_stream_t stream);

/* Waits for all work on the stream (NULL: the whole device). */
int omega_synchronize(omega_stream_t stream);

/* Per-kernel timing with HIP events recorded on the launch stream around
 * every launch of the library's kernels.  Off by default. */
int omega_timing_enable(int enable);
int omega_timing_reset(void);
/* name: "diag", "tend", "vdel4", "update", "reduce" (and, for the
 * hexagonal tiled variant, "tend_edges" / "tend_cells", the two launches
 * inside "tend"); returns summed milliseconds and launch count since the
 * last reset (synchronises). */
int omega_timing_read(const char *name, double *total_ms, int64_t *launches);

#ifdef __cplusplus
}
#endif

#endif /* OMEGA_SWE_H */
