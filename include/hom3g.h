/*
 * hom3g — C ABI of the MI355X (gfx950) implementation of hom3's
 * finite-volume residual-and-update hot path.
 *
 * The reference (gnzlbg/hom3) is a header-only C++14 library with no FFI;
 * its plug-in boundary is the template pair Solver<PhysicsTT, TimeIntegration>
 * (src/solver/fv/solver.hpp:60-61) with bc::Interface conditions
 * (src/solver/fv/boundary_condition.hpp:22-58), the grid
 * grid::CartesianHSP (src/grid/grid.hpp:147) and the coupling free functions
 * (src/solver/fv/coupling.hpp:24-73).  Each entry point below names the
 * reference member it replaces.  The C++ facade in include/hom3g/fv_solver.hpp
 * and the Python package hom3_amd are the callers.
 *
 * Conventions
 *  - Every call returns HOM3G_OK (0) or a negative status; the message is
 *    available from hom3g_last_error() (replaces TERMINATE/exit(1),
 *    src/misc/error.hpp:27-38).
 *  - Host buffers are owned by the caller and copied; device buffers and the
 *    HIP stream are owned by the context.  Calls on one context are ordered
 *    on that context's stream; one host thread per context.
 *  - Cell-indexed arrays use the reference's column-major SoA layout
 *    (container.hpp:58-66, Matrix<.., ColMajor>): element (cell c, column k)
 *    at [k*ld + c].  Invalid ids are INT64_MAX, invalid reals are NaN
 *    (src/misc/constants.hpp:42-63).
 */
#ifndef HOM3G_H_
#define HOM3G_H_
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HOM3G_OK 0
#define HOM3G_ERR_ARG -1
#define HOM3G_ERR_HIP -2
#define HOM3G_ERR_STATE -3
#define HOM3G_ERR_DIVERGED -4
#define HOM3G_ERR_CAPACITY -5

#define HOM3G_INVALID_ID INT64_MAX

/* physics (solver/fv/{euler,cns,heat}/physics.hpp) */
enum { HOM3G_EULER = 0, HOM3G_CNS = 1, HOM3G_HEAT = 2 };
/* numerical flux tags (euler/tags.hpp:13-16, cns/tags.hpp:15-23, heat/tags.hpp:13-15) */
enum { HOM3G_FLUX_AUSM = 0, HOM3G_FLUX_LLF = 1, HOM3G_FLUX_THREE_POINT = 2 };
/* time integration tags (solver/fv/tags.hpp:20-28) + SSP-RK3 extension */
enum { HOM3G_TI_EULER_FWD = 0, HOM3G_TI_HEUN_RK2 = 1, HOM3G_TI_SSP_RK3 = 2 };
/* reconstruction (extension; NONE = the reference's first-order scheme) */
enum { HOM3G_RECON_NONE = 0, HOM3G_RECON_MUSCL = 1 };
enum { HOM3G_LIM_ZERO = 0, HOM3G_LIM_MINMOD = 1, HOM3G_LIM_VANLEER = 2 };
/* state precision */
enum { HOM3G_FP64 = 0, HOM3G_FP32_STATE = 1 };
/* kernel family: AUTO picks the brick kernel when the topology allows it */
enum { HOM3G_KERNEL_AUTO = 0, HOM3G_KERNEL_GENERIC = 1, HOM3G_KERNEL_BRICK = 2 };

/* implicit geometries (src/geometry/implicit.hpp:39-103, Inverted :119-132) */
enum { HOM3G_GEOM_NONE = 0, HOM3G_GEOM_EDGE = 1, HOM3G_GEOM_SQUARE = 2, HOM3G_GEOM_SPHERE = 3 };

/* boundary conditions: one kind per reference Condition class */
enum {
  HOM3G_BC_NEUMANN = 0,            /* euler/boundary_conditions.hpp:17-38, cns::bc::Neumann */
  HOM3G_BC_DIRICHLET = 1,          /* euler/boundary_conditions.hpp:42-60 */
  HOM3G_BC_ISENTROPIC_VORTEX = 2,  /* euler/boundary_conditions.hpp:65-89 */
  HOM3G_BC_INFLOW = 3,             /* cns/boundary_conditions.hpp:28-65 */
  HOM3G_BC_OUTFLOW = 4,            /* cns/boundary_conditions.hpp:72-105 */
  HOM3G_BC_ADIABATIC_NOSLIP = 5,   /* cns/boundary_conditions.hpp:116-150 */
  HOM3G_BC_ADIABATIC_SLIP = 6,     /* cns/boundary_conditions.hpp:156-193 */
  HOM3G_BC_ISOTHERMAL_NOSLIP = 7,  /* cns/boundary_conditions.hpp:199-238 */
  HOM3G_BC_HEAT_DIRICHLET = 8,     /* heat/boundary_conditions.hpp:15-37 */
  HOM3G_BC_HEAT_NEUMANN = 9,       /* heat/boundary_conditions.hpp:40-63 */
  HOM3G_BC_PERIODIC = 10,          /* extension: ghost = periodic image */
  HOM3G_BC_CHT_FLUID = 11,         /* cns/coupling_conditions.hpp:21-43 */
  HOM3G_BC_CHT_SOLID = 12,         /* heat/coupling_conditions.hpp:18-38 */
  HOM3G_BC_EULER_COUPLING = 13,    /* euler/coupling_conditions.hpp:17-35 */
  HOM3G_BC_HOST_CALLBACK = 14      /* any other Condition: hom3g_solver_append_bc_host */
};

/* slope hook of a host-callback condition (the Condition's slope<>,
 * boundary_condition.hpp:63-67): the boundary cell's slope, or the wall
 * form that negates the velocity slopes (cns/boundary_conditions.hpp:143-147) */
enum { HOM3G_SLOPE_DEFAULT = 0, HOM3G_SLOPE_WALL = 1 };

/* A Condition evaluated on the host (bc::Interface with an arbitrary
 * Condition, boundary_condition.hpp:22-67; its operator()(tag, ghosts)).
 * Called with the ghosts the condition fills, in the apply-range order of
 * solver.hpp:561-586: buffer 0 = lhs (lhs_tag), 1 = rhs (rhs_tag), 2 = the
 * SSP-RK3 stage buffer; the solver's time; the ghosts' and their boundary
 * cells' ids (ghost g's boundary cell is bndry[g]) and centres (nd x n,
 * column-major); their states (nvars x n, column-major): q_ghost is to be
 * filled, q_bndry may be overwritten (as IsentropicVortex does; every ghost
 * of one boundary cell must then write the same value). */
typedef void (*hom3g_bc_callback)(void* user, int buffer, double time, int64_t n, const int64_t* ghosts,
                                  const int64_t* bndry, const double* x_ghost, const double* x_bndry,
                                  double* q_ghost, double* q_bndry);
/* An initial condition evaluated on the host (the "initialCondition"
 * property, solver.hpp:153-156): q = conserved variables at point x. */
typedef void (*hom3g_ic_callback)(void* user, const double* x, double* q);

/* built-in initial conditions evaluated on the device with the reference's
 * 2^nd-point Gauss average (quadrature/quadrature.hpp:55-74) */
enum {
  HOM3G_IC_CONSTANT_PV = 0, HOM3G_IC_SHOCK_TUBE = 1, HOM3G_IC_ISENTROPIC_VORTEX = 2,
  HOM3G_IC_SPHERE = 3, HOM3G_IC_TGV = 4, HOM3G_IC_FOUR_QUADRANT = 5,
  HOM3G_IC_HEAT_CONSTANT = 6, HOM3G_IC_HEAT_SINE = 7, HOM3G_IC_LINEAR_PV = 8
};

typedef struct hom3g_geometry {
  int32_t kind;
  int32_t invert;  /* 1: sign flipped (geometry::implicit::adaptors::Inverted) */
  double a[3];     /* EDGE: point;  SQUARE: centre;  SPHERE: centre */
  double b[3];     /* EDGE: normal; SQUARE: lengths; SPHERE: b[0] = radius */
} hom3g_geometry;

/* Physics/solver properties: the io::Properties keys the hot path reads
 * (solver.hpp:88-95, euler/physics.hpp:39-42, euler/quantities.hpp:23-29,
 * cns/physics.hpp:57-61, cns/quantities.hpp:23-30, heat/physics.hpp:36-39,
 * heat/quantities.hpp:14-17).  The leading fields match the oracle's
 * PhysDesc one-to-one. */
typedef struct hom3g_physics {
  int32_t physics, nd, flux, ti, recon, limiter;
  double gamma, cfl, cfl_viscous;
  double Rspecific, T0, rho0, Minfty;
  double ReInfty, Pr0, Sutherland0;
  double diffusivity, T_ref;
  double time_end;        /* "timeEnd" */
  int64_t max_cells;      /* "maxNoCells" (0: sized automatically) */
  int32_t precision;      /* HOM3G_FP64 | HOM3G_FP32_STATE */
  int32_t kernel;         /* HOM3G_KERNEL_* */
} hom3g_physics;

typedef struct hom3g_ctx hom3g_ctx;

/* ---- context -------------------------------------------------------------- */
const char* hom3g_last_error(void);
int hom3g_create(hom3g_ctx** ctx, int device);
int hom3g_destroy(hom3g_ctx* ctx);
int hom3g_synchronize(hom3g_ctx* ctx);
int hom3g_version(void);

/* ---- grid: grid::CartesianHSP + generation::MinLevel ----------------------
 * grid/grid.hpp:172-196 (constructor + initialize), grid/generation.hpp:39-60,
 * containers/hierarchical/implementation.hpp:320-364 (refine_node).  The
 * tree is generated on the device in the reference's breadth-first order. */
int hom3g_grid_create(hom3g_ctx* ctx, int nd, const double* xmin, const double* xmax,
                      int min_level, int max_solvers, int* grid_id);
/* Replace the tree with a caller-built one (parent/first_child per node, as
 * containers/hierarchical/implementation.hpp:783-784). */
int hom3g_grid_set_tree(hom3g_ctx* ctx, int grid_id, int64_t nnodes, const int64_t* parent,
                        const int64_t* first_child);
/* 2:1 refinement extension (the reference has no refinement criterion; it
 * only offers refine_node, implementation.hpp:320-364).  Passes flag, on the
 * tree as it stands at the start of the pass, every leaf below max_level
 * whose centre lies within band + h*sqrt(nd)/2 of the geometry's surface,
 * then refine the flagged leaves in increasing node order (children
 * appended at the end); passes repeat until nothing is flagged.  Only valid
 * before any solver on the grid is built.  nrefined (optional) receives the
 * number of refined nodes. */
int hom3g_grid_refine_region(hom3g_ctx* ctx, int grid_id, int max_level, const hom3g_geometry* g,
                             double band, int64_t* nrefined);
/* 2:1 face balance in the same pass structure: a refined node whose
 * same-level face neighbour is missing flags the coarser leaf covering that
 * position. */
int hom3g_grid_balance(hom3g_ctx* ctx, int grid_id, int64_t* nrefined);
int hom3g_grid_size(hom3g_ctx* ctx, int grid_id, int64_t* nnodes);
int hom3g_grid_tree(hom3g_ctx* ctx, int grid_id, int64_t* parent, int64_t* first_child);
/* find_samelvl_neighbor for every node (implementation.hpp:457-543); out is
 * row-major nnodes x 2nd */
int hom3g_grid_neighbors(hom3g_ctx* ctx, int grid_id, int64_t* out);
/* cell_coordinates / cell_length for every node (grid.hpp:344-374) */
int hom3g_grid_geometry(hom3g_ctx* ctx, int grid_id, double* x_rowmajor, double* length);

/* ---- solver ----------------------------------------------------------------
 * Solver(SolverIdx, io::Properties) — solver.hpp:88-95 */
int hom3g_solver_create(hom3g_ctx* ctx, int grid_id, int solver_idx, const hom3g_physics* phys,
                        int* solver_id);
/* "initialDomain" property (solver.hpp:882-886): sign>0 keeps cells with
 * sdf(x)>0, sign<0 keeps sdf(x)<0, sign==0 keeps all. */
int hom3g_solver_set_domain(hom3g_ctx* ctx, int sid, const hom3g_geometry* g, int sign);
int hom3g_solver_set_domain_mask(hom3g_ctx* ctx, int sid, const uint8_t* leaf_mask, int64_t n);
/* append_bc (solver.hpp:159-162).  params are the constant surface values of
 * the Condition; other_sid is the coupled solver for CHT kinds (-1 if none);
 * kratio = k_solid/k_fluid (1 = reference CHT). */
int hom3g_solver_append_bc(hom3g_ctx* ctx, int sid, const hom3g_geometry* g, int kind,
                           const double* params, int nparams, int other_sid, double kratio);
/* append_bc (solver.hpp:159-162) of a Condition evaluated on the host
 * (HOM3G_BC_HOST_CALLBACK): the library fills the other conditions' ghosts
 * on the device, then copies the boundary cells of this one's ghosts to the
 * host, calls f, and copies the ghosts back (one synchronisation per ghost
 * fill; steps with such a condition are not captured as graphs).
 * slope_hook: HOM3G_SLOPE_*. */
int hom3g_solver_append_bc_host(hom3g_ctx* ctx, int sid, const hom3g_geometry* g, hom3g_bc_callback f, void* user,
                                int slope_hook);
/* initialize() without the initial condition (solver.hpp:133-138, 876-977):
 * local cells, neighbours, ghosts, sort_gc, distances — on the device.
 * ghost_order 0 = canonical (stable) order. */
int hom3g_solver_build(hom3g_ctx* ctx, int sid);
/* impose_initial_condition (solver.hpp:980-992) for a built-in IC, then
 * set_dt (solver.hpp:541-551). */
int hom3g_solver_impose_ic(hom3g_ctx* ctx, int sid, int ic_kind, const double* params, int nparams);
/* impose_initial_condition (solver.hpp:980-992) of a host-evaluated initial
 * condition, then set_dt: f is called at every internal cell's 2^nd Gauss
 * points (+-1/sqrt(3) of the half length, quadrature/quadrature.hpp:14-74)
 * in the reference's order, the cell average ((sum f) V / 2^nd) / V is
 * uploaded; a seeded perturbation (below) applies as for built-in kinds. */
int hom3g_solver_impose_ic_host(hom3g_ctx* ctx, int sid, hom3g_ic_callback f, void* user);
/* apply_bcs(lhs) (solver.hpp:561-586): fills the ghosts of the current state
 * (write_domain does this before writing, solver.hpp:329-333). */
int hom3g_solver_apply_bcs(hom3g_ctx* ctx, int sid);
/* Seeded density perturbation applied by hom3g_solver_impose_ic after the
 * initial condition (extension, no reference counterpart): rho *= 1 +
 * amplitude*(2U - 1) at fixed velocity and pressure, U in [0,1) from
 * SplitMix64 keyed by (seed, level, Morton index of the cell at its level). */
int hom3g_solver_set_perturbation(hom3g_ctx* ctx, int sid, uint64_t seed, double amplitude);
/* Faces towards finer cells on a 2:1 refined grid.  The neighbour table
 * holds -(2 + f) for face f; out (may be null) receives 2^(nd-1) cell ids
 * per face in ascending child position, row-major. */
int hom3g_solver_fine_faces(hom3g_ctx* ctx, int sid, int32_t* out, int64_t* nfine);
int hom3g_solver_counts(hom3g_ctx* ctx, int sid, int64_t* ncells, int64_t* first_ghost);
/* nd, nvars and physics of a solver (any pointer may be NULL) */
int hom3g_solver_info(hom3g_ctx* ctx, int sid, int* nd, int* nvars, int* physics);
/* Kernel family chosen at build (HOM3G_KERNEL_GENERIC or _BRICK). */
int hom3g_solver_kernel(hom3g_ctx* ctx, int sid, int* kernel);
/* Download the topology (column-major, ld = ncells). Any pointer may be NULL. */
int hom3g_solver_topology(hom3g_ctx* ctx, int sid, int64_t* neighbors, int64_t* node_idx,
                          int32_t* bc_idx, double* x_center, double* length, double* distances);
/* cells().lhs upload/download (container.hpp:148), column-major with ld */
int hom3g_solver_upload_state(hom3g_ctx* ctx, int sid, const double* lhs, int64_t ld);
int hom3g_solver_download_state(hom3g_ctx* ctx, int sid, double* lhs, int64_t ld);
/* time(), dt(), step() set after an upload (solver.hpp:406-416, used by the
 * next set_dt, :541-551); a dt forced before the reset (force_dt, coupling)
 * is dropped */
int hom3g_solver_reset_time(hom3g_ctx* ctx, int sid, double time, double dt, int64_t step);
/* per-cell volumetric source for the heat solver (extension; the physics
 * returns dt*S as compute_source_term, solver.hpp:462-465) */
int hom3g_solver_set_source(hom3g_ctx* ctx, int sid, const double* src, int64_t n);

/* solve() x nsteps (solver.hpp:140-145): apply_bcs(lhs), set_dt, evolve,
 * advance — entirely on the device, no host synchronisation per step.
 * Returns HOM3G_ERR_DIVERGED ("Solution diverged!", utilities.hpp:73-76,
 * 107-110) when a step's computed or forced dt left [1e-300, 1e300] (the
 * end-time clamp of set_dt is not divergence); the steps after it are still
 * taken on the device, the message names the first such step. */
int hom3g_solver_solve(hom3g_ctx* ctx, int sid, int nsteps);
/* min_dt() (solver.hpp:109, compute_dt :511-517) */
int hom3g_solver_min_dt(hom3g_ctx* ctx, int sid, double* dt);
/* dt(Num) -> force_dt (solver.hpp:115, 522-525) */
int hom3g_solver_force_dt(hom3g_ctx* ctx, int sid, double dt);
/* time(), dt(), step() (solver.hpp:103-112) */
int hom3g_solver_time(hom3g_ctx* ctx, int sid, double* time, double* dt, int64_t* step);
/* FP64 sums of cell volume * conserved variable over internal cells (extension) */
int hom3g_solver_conservation_sums(hom3g_ctx* ctx, int sid, double* out);
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, serving loops, harnesses that forbid
 * hipMalloc): the sums of hom3g_solver_conservation_sums (bit-identical),
 * with the partial sums in a static device scratch of 256 x 8 doubles
 * (16 KiB per GPU, held until the process ends) instead of buffers
 * allocated per call.  sums_dev: nvars doubles of device memory, written on
 * the context's stream without a host synchronisation (hom3g_synchronize
 * before reading them).  Solvers with more than 8 variables are rejected
 * (HOM3G_ERR_CAPACITY).  All calls on one GPU share the scratch: calls that
 * overlap there (other contexts or host threads) race. */
int hom3g_solver_conservation_sums_noalloc(hom3g_ctx* ctx, int sid, double* sums_dev);
/* Non-finite / non-positive density-pressure scan of the internal cells:
 * returns the count of bad cells in *nbad (extension; the reference only
 * checks dt, utilities.hpp:73-76). */
int hom3g_solver_check_state(hom3g_ctx* ctx, int sid, int64_t* nbad);
/* Time nsteps solve() calls on the context stream with HIP events; *ms is the
 * elapsed device time. */
int hom3g_solver_time_solve(hom3g_ctx* ctx, int sid, int nsteps, float* ms);
/* As hom3g_solver_time_solve, with HIP events around every stage-kernel
 * launch: *stage_ms is the mean duration of one stage launch and *nlaunches
 * their number (profiling aid for the roofline report). */
int hom3g_solver_profile_solve(hom3g_ctx* ctx, int sid, int nsteps, float* total_ms, float* stage_ms,
                               int* nlaunches);
int hom3g_solver_destroy(hom3g_ctx* ctx, int sid);
/* A plain device-to-device copy of `bytes` (read `bytes`, write `bytes`),
 * timed with HIP events: the reference point for the HBM byte counters and
 * the achievable bandwidth in the roofline report (no reference counterpart). */
int hom3g_copy_calibration(hom3g_ctx* ctx, int64_t bytes, float* ms);
/* Mean time (*ms) of `reps` device copies of `bytes` with copy variant
 * `variant` (0: grid-stride 32-byte loads and stores, 1: the same with
 * non-temporal loads and stores, 2: grid-stride with four loads in flight
 * and non-temporal stores, 3: one 32-byte element per thread) on `grid` x
 * `block` threads: the sweep behind the achievable-bandwidth figure of the
 * roofline report (no reference counterpart). */
int hom3g_copy_bandwidth(hom3g_ctx* ctx, int64_t bytes, int variant, int grid, int block, int reps, float* ms);

/* ---- coupling (solver/fv/coupling.hpp) ------------------------------------ */
/* nsteps of: coupling::apply(Tr0, cns, heat); cns.solve(); heat.solve();
 * (coupled_cns_heat_test.cpp:235-240) */
int hom3g_couple_cht_step(hom3g_ctx* ctx, int cns_sid, int heat_sid, double Tr0, int nsteps);
/* nsteps of: coupling::apply(s1, s2); s1.solve(); s2.solve(); (coupling.hpp:66-73) */
int hom3g_couple_same_step(hom3g_ctx* ctx, int sid1, int sid2, int nsteps);

/* ---- conjugate heat transfer: objective, tangent, adjoint (no reference
 * counterpart) --------------------------------------------------------------
 * The forward map is the reference's coupled step (coupling::apply, then
 * cns.solve(), then heat.solve(); coupling.hpp:47-63,
 * coupled_cns_heat_test.cpp:235-248), run from the solvers' current states
 * (they end advanced by nsteps).  J = sum_n dt_heat^n Q^{n+1}, Q = wall heat
 * flux out of the solid through its conjugate-interface faces: the sum over
 * the solid's interface ghosts of h^(nd-2) * 2 (T_b - T_s).
 * dts: the fluid dt of every step (heat dt = dt * Tr0 * Pr0 * Re0); NULL in
 * hom3g_cht_objective takes the CFL-limited dts of coupling::apply and, with
 * dts_out, returns them.  The tangent and the adjoint need a prescribed dts
 * (the schedule is frozen so that J is a smooth function of the source). */
int hom3g_cht_objective(hom3g_ctx* ctx, int cns_sid, int heat_sid, double Tr0, int nsteps, const double* dts,
                        double* dts_out, double* J);
/* FP64 forward-mode derivative of J along dsource (one value per internal
 * solid cell). */
int hom3g_cht_tangent(hom3g_ctx* ctx, int cns_sid, int heat_sid, double Tr0, int nsteps, const double* dts,
                      const double* dsource, double* J, double* dJ);
/* Discrete adjoint: dJ/dS for the solid's per-cell source (grad_source, one
 * value per internal solid cell).  ckpt_every > 0: checkpoints every
 * ckpt_every steps and one recomputation per segment; ckpt_every == 0: the
 * forward sweep keeps every step's stage inputs in device memory and the
 * reverse sweep recomputes nothing (HOM3G_ERR_CAPACITY when they do not fit);
 * ckpt_every == -S (1 <= S <= 4096): binomial checkpointing (revolve's
 * schedule) with at most S saved states in device memory besides the
 * initial one, whatever the number of steps.  The gradient is the same,
 * bit for bit, in every mode.
 * The solvers end at step nsteps as after a forward run.  ms_forward /
 * ms_reverse (optional): event times of the forward and reverse sweeps. */
int hom3g_cht_adjoint(hom3g_ctx* ctx, int cns_sid, int heat_sid, double Tr0, int nsteps, const double* dts,
                      int ckpt_every, double* J, double* grad_source, float* ms_forward, float* ms_reverse);

/* Where hom3g_cht_adjoint / hom3g_cht_misfit_adjoint keep their checkpoints
 * (the coupled state every ckpt_every steps, ckpt_every > 0):
 * HOM3G_CKPT_AUTO (default) device memory while they fit, else pinned host
 * memory; HOM3G_CKPT_DEVICE; HOM3G_CKPT_HOST (copied out after each
 * checkpoint step, back before each segment's recomputation).  The gradient
 * is the same either way.  An extension (checkpoint spill beyond HBM): the
 * reference has no adjoint. */
enum { HOM3G_CKPT_AUTO = 0, HOM3G_CKPT_DEVICE = 1, HOM3G_CKPT_HOST = 2 };
int hom3g_set_checkpoint_memory(hom3g_ctx* ctx, int where);

/* Inverse-source form: the same coupled run with the objective
 * J = 1/2 sum_c (T_N,c - T_obs,c)^2 over the internal solid cells at the
 * final step (T_obs: one value per internal solid cell), e.g. a misfit
 * against observed temperatures; the source gradient of J drives a source
 * reconstruction.  Same dts, dsource, ckpt_every conventions as above. */
int hom3g_cht_misfit_objective(hom3g_ctx* ctx, int cns_sid, int heat_sid, double Tr0, int nsteps,
                               const double* dts, const double* T_obs, double* J);
int hom3g_cht_misfit_tangent(hom3g_ctx* ctx, int cns_sid, int heat_sid, double Tr0, int nsteps, const double* dts,
                             const double* T_obs, const double* dsource, double* J, double* dJ);
int hom3g_cht_misfit_adjoint(hom3g_ctx* ctx, int cns_sid, int heat_sid, double Tr0, int nsteps, const double* dts,
                             int ckpt_every, const double* T_obs, double* J, double* grad_source, float* ms_forward,
                             float* ms_reverse);

/* ---- batched instances (no reference counterpart) --------------------------
 * Many independent first-order Euler solvers (AUSM, no reconstruction), each
 * with its own grid, 2:1 refinement, dt, time and step, advanced together by
 * one persistent kernel per RK stage; a step is captured in a hipGraph.
 * Each instance follows solver.hpp:140-145 (set_dt, evolve, advance) exactly,
 * so a batch of one equals hom3g_solver_solve.  Instances must be built and
 * initialised solvers sharing nd, gamma, CFL and time integration, with
 * zero-gradient Neumann or periodic boundaries only; their state is copied
 * into the batch at creation (the solvers themselves are left unchanged).
 * precision: HOM3G_FP64, or HOM3G_FP32_STATE (FP32 RK stage buffers and flux
 * arithmetic, FP64 U^n and RK accumulation). */
int hom3g_batch_create(hom3g_ctx* ctx, const int* solver_ids, int n, int precision, int* batch_id);
/* Launch shape: block threads (a multiple of 64, at most 512), tile = cells
 * per work item (at least block), grid = persistent workgroups.  Changing it
 * drops the captured graph. */
int hom3g_batch_set_config(hom3g_ctx* ctx, int batch_id, int block, int tile, int grid);
int hom3g_batch_get_config(hom3g_ctx* ctx, int batch_id, int* block, int* tile, int* grid, int64_t* nitems);
/* Times stage 1 over block x tile x workgroups-per-CU candidates on the
 * current state (nothing is advanced) and keeps the fastest; outputs may be
 * NULL. */
int hom3g_batch_autotune(hom3g_ctx* ctx, int batch_id, int* block, int* tile, int* grid, float* stage_ms);
int hom3g_batch_solve(hom3g_ctx* ctx, int batch_id, int nsteps, int use_graph);
/* nsteps steps between two events on the library stream; ms per step */
int hom3g_batch_time_solve(hom3g_ctx* ctx, int batch_id, int nsteps, int use_graph, float* ms_per_step);
int hom3g_batch_counts(hom3g_ctx* ctx, int batch_id, int* ninstances, int64_t* total_cells);
/* internal cells of one instance, column-major nvars x n */
int hom3g_batch_download_state(hom3g_ctx* ctx, int batch_id, int instance, double* out, int64_t n);
int hom3g_batch_time(hom3g_ctx* ctx, int batch_id, int instance, double* time, double* dt, int64_t* step);
int hom3g_batch_destroy(hom3g_ctx* ctx, int batch_id);

#ifdef __cplusplus
}
#endif
#endif /* HOM3G_H_ */
