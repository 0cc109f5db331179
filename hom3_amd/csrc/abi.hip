// extern "C" entry points of libhom3g (declared in include/hom3g.h).
// Every entry point converts exceptions into a status code and a message
// retrievable with hom3g_last_error(); nothing here terminates the process.
#include <cmath>
#include <cstring>
#include <string>

#include "internal.hpp"

using namespace hom3g;

namespace {
thread_local std::string g_last_error;

template <class F>
int guarded(F&& f) {
  try {
    f();
    return HOM3G_OK;
  } catch (const Error& e) {
    g_last_error = e.what();
    return e.code;
  } catch (const std::exception& e) {
    g_last_error = e.what();
    return HOM3G_ERR_ARG;
  }
}

Ctx& ctx_ref(hom3g_ctx* c) {
  if (!c) throw Error(HOM3G_ERR_ARG, "null context");
  return *reinterpret_cast<Ctx*>(c);
}
GridDev& grid_ref(Ctx& C, int id) {
  if (id < 0 || id >= (int)C.grids.size() || !C.grids[id]) throw Error(HOM3G_ERR_ARG, "bad grid id");
  return *C.grids[id];
}
SolverDev& solver_ref(Ctx& C, int id) {
  if (id < 0 || id >= (int)C.solvers.size() || !C.solvers[id]) throw Error(HOM3G_ERR_ARG, "bad solver id");
  return *C.solvers[id];
}
SolverDev& built_ref(Ctx& C, int id) {
  SolverDev& S = solver_ref(C, id);
  if (!S.built) throw Error(HOM3G_ERR_STATE, "solver not built (call hom3g_solver_build)");
  return S;
}
void set_device(Ctx& C) { H3_HIP(hipSetDevice(C.device)); }

// euler/quantities.hpp:131-145 and cns/quantities.hpp:23-30, 113-115
PhysConst make_phys_const(const hom3g_physics& p) {
  PhysConst c{};
  c.physics = p.physics;
  c.nd = p.nd;
  c.nvars = (p.physics == HOM3G_HEAT) ? 1 : p.nd + 2;
  c.flux = p.flux;
  c.recon = p.recon;
  c.limiter = p.limiter;
  c.gamma = p.gamma;
  c.gm1 = p.gamma - 1.;
  c.cfl = p.cfl;
  c.cfl_v = p.cfl_viscous;
  c.Pr0 = p.Pr0;
  c.sOverTref = (p.T0 != 0) ? p.Sutherland0 / p.T0 : 0.0;
  c.sOverTrefP1 = c.sOverTref + 1.0;
  c.Re0 = 0.0;
  if (p.physics == HOM3G_CNS) {
    const double gm1 = p.gamma - 1.;
    const double Tinf = 1. / (1 + 0.5 * gm1 * (p.Minfty * p.Minfty));
    const double rinf = std::pow(Tinf, 1. / gm1);
    const double uinf = p.Minfty * std::sqrt(Tinf);
    const double muinf = std::pow(Tinf, 1.5) * c.sOverTrefP1 / (Tinf + c.sOverTref);
    c.Re0 = p.ReInfty * muinf / (rinf * uinf);
  }
  c.q_coef = (c.gm1 != 0 && c.Pr0 != 0) ? c.gamma / (c.gm1 * c.Pr0) : 0.0;
  c.inv_Re0 = (c.Re0 != 0) ? 1.0 / c.Re0 : 0.0;
  c.visc_dt = 0.0;
  if (c.Re0 > 0 && c.Pr0 > 0 && c.gm1 > 0)
    c.visc_dt = 1.0 / (2.0 * c.sOverTrefP1 * std::max(1.0 / c.Re0, 1.0 / (c.Re0 * c.Pr0 * c.gm1)));
  return c;
}

bool coupled_kind(int k) {
  return k == HOM3G_BC_CHT_FLUID || k == HOM3G_BC_CHT_SOLID || k == HOM3G_BC_EULER_COUPLING;
}

void ensure_partners(Ctx& C, SolverDev& S) {
  if (S.partners_built) return;
  bool need = false;
  for (auto& b : S.bcs) need = need || coupled_kind(b.kind);
  if (need) {
    // the coupled solver must be built first; retried on the next call otherwise
    for (auto& b : S.bcs)
      if (coupled_kind(b.kind) &&
          (b.other_sid < 0 || b.other_sid >= (int)C.solvers.size() || !C.solvers[b.other_sid] ||
           !C.solvers[b.other_sid]->built))
        return;
    solver_build_partners(C, S);
  }
  S.partners_built = true;
}
}  // namespace

extern "C" {

const char* hom3g_last_error(void) { return g_last_error.c_str(); }
int hom3g_version(void) { return 10000; }

int hom3g_create(hom3g_ctx** out, int device) {
  return guarded([&] {
    if (!out) throw Error(HOM3G_ERR_ARG, "null output");
    int ndev = 0;
    H3_HIP(hipGetDeviceCount(&ndev));
    if (device < 0 || device >= ndev) throw Error(HOM3G_ERR_ARG, "no such HIP device");
    auto* C = new Ctx();
    C->device = device;
    try {
      H3_HIP(hipSetDevice(device));
      H3_HIP(hipStreamCreateWithFlags(&C->stream, hipStreamNonBlocking));
      H3_HIP(hipEventCreate(&C->ev0));
      H3_HIP(hipEventCreate(&C->ev1));
      noalloc_init(*C);
    } catch (...) {
      delete C;
      throw;
    }
    *out = reinterpret_cast<hom3g_ctx*>(C);
  });
}

int hom3g_destroy(hom3g_ctx* ctx) {
  return guarded([&] {
    if (!ctx) return;
    Ctx* C = reinterpret_cast<Ctx*>(ctx);
    (void)hipSetDevice(C->device);
    (void)hipStreamSynchronize(C->stream);
    for (auto& B : C->batches)
      if (B) batch_release(*B);
    C->batches.clear();
    C->solvers.clear();
    C->grids.clear();
    C->red.release();
    C->red_u64.release();
    for (auto e : C->prof_events) (void)hipEventDestroy(e);
    if (C->ev0) (void)hipEventDestroy(C->ev0);
    if (C->ev1) (void)hipEventDestroy(C->ev1);
    if (C->stream) (void)hipStreamDestroy(C->stream);
    delete C;
  });
}

int hom3g_synchronize(hom3g_ctx* ctx) {
  return guarded([&] {
    Ctx& C = ctx_ref(ctx);
    set_device(C);
    H3_HIP(hipStreamSynchronize(C.stream));
  });
}

// ---- grid -------------------------------------------------------------------
int hom3g_grid_create(hom3g_ctx* ctx, int nd, const double* xmin, const double* xmax, int min_level,
                      int max_solvers, int* grid_id) {
  return guarded([&] {
    Ctx& C = ctx_ref(ctx);
    set_device(C);
    H3_REQUIRE(nd == 2 || nd == 3, "nd must be 2 or 3");
    H3_REQUIRE(xmin && xmax && grid_id, "null argument");
    auto G = std::make_unique<GridDev>();
    G->nd = nd;
    G->max_solvers = max_solvers > 0 ? max_solvers : 1;
    // grid.hpp:108-132 RootCell(min, max)
    double length = 2.220446049250313e-16;
    for (int d = 0; d < nd; ++d) {
      const double ld = xmax[d] - xmin[d];
      length = std::max(length, ld);
      G->root_x[d] = xmin[d] + 0.5 * ld;
    }
    G->root_len = length;
    grid_generate_uniform(C, *G, min_level);
    H3_HIP(hipStreamSynchronize(C.stream));
    C.grids.push_back(std::move(G));
    *grid_id = (int)C.grids.size() - 1;
  });
}

int hom3g_grid_set_tree(hom3g_ctx* ctx, int grid_id, int64_t nnodes, const int64_t* parent,
                        const int64_t* first_child) {
  return guarded([&] {
    Ctx& C = ctx_ref(ctx);
    set_device(C);
    GridDev& G = grid_ref(C, grid_id);
    H3_REQUIRE(nnodes > 0 && parent && first_child, "bad tree");
    G.nnodes = nnodes;
    G.parent.alloc(nnodes);
    G.first_child.alloc(nnodes);
    G.parent.upload(parent, nnodes, C.stream);
    G.first_child.upload(first_child, nnodes, C.stream);
    grid_compute_levels(C, G);
    G.uniform_level = -1;
    H3_HIP(hipStreamSynchronize(C.stream));
  });
}

static void require_no_built_solver(Ctx& C, int grid_id) {
  for (auto& S : C.solvers)
    H3_REQUIRE(!(S && S->grid_id == grid_id && S->built), "grid already has a built solver");
}

int hom3g_grid_refine_region(hom3g_ctx* ctx, int grid_id, int max_level, const hom3g_geometry* geom, double band,
                             int64_t* nrefined) {
  return guarded([&] {
    Ctx& C = ctx_ref(ctx);
    set_device(C);
    GridDev& G = grid_ref(C, grid_id);
    H3_REQUIRE(geom != nullptr, "null geometry");
    require_no_built_solver(C, grid_id);
    const int64_t k = grid_refine_region(C, G, max_level, *geom, band);
    if (nrefined) *nrefined = k;
  });
}

int hom3g_grid_balance(hom3g_ctx* ctx, int grid_id, int64_t* nrefined) {
  return guarded([&] {
    Ctx& C = ctx_ref(ctx);
    set_device(C);
    GridDev& G = grid_ref(C, grid_id);
    require_no_built_solver(C, grid_id);
    const int64_t k = grid_balance(C, G);
    if (nrefined) *nrefined = k;
  });
}

int hom3g_grid_size(hom3g_ctx* ctx, int grid_id, int64_t* n) {
  return guarded([&] {
    Ctx& C = ctx_ref(ctx);
    *n = grid_ref(C, grid_id).nnodes;
  });
}

int hom3g_grid_tree(hom3g_ctx* ctx, int grid_id, int64_t* parent, int64_t* first_child) {
  return guarded([&] {
    Ctx& C = ctx_ref(ctx);
    set_device(C);
    GridDev& G = grid_ref(C, grid_id);
    if (parent) G.parent.download(parent, G.nnodes, C.stream);
    if (first_child) G.first_child.download(first_child, G.nnodes, C.stream);
    H3_HIP(hipStreamSynchronize(C.stream));
  });
}

int hom3g_grid_neighbors(hom3g_ctx* ctx, int grid_id, int64_t* out) {
  return guarded([&] {
    Ctx& C = ctx_ref(ctx);
    set_device(C);
    grid_neighbors(C, grid_ref(C, grid_id), out);
  });
}

int hom3g_grid_geometry(hom3g_ctx* ctx, int grid_id, double* x, double* len) {
  return guarded([&] {
    Ctx& C = ctx_ref(ctx);
    set_device(C);
    grid_geometry(C, grid_ref(C, grid_id), x, len);
  });
}

// ---- solver -------------------------------------------------------------------
int hom3g_solver_create(hom3g_ctx* ctx, int grid_id, int solver_idx, const hom3g_physics* phys,
                        int* solver_id) {
  return guarded([&] {
    Ctx& C = ctx_ref(ctx);
    set_device(C);
    GridDev& G = grid_ref(C, grid_id);
    H3_REQUIRE(phys && solver_id, "null argument");
    H3_REQUIRE(phys->nd == G.nd, "physics nd differs from the grid's");
    H3_REQUIRE(phys->physics >= 0 && phys->physics <= 2, "unknown physics");
    H3_REQUIRE(phys->ti >= 0 && phys->ti <= 2, "unknown time integration");
    H3_REQUIRE(solver_idx >= 0 && solver_idx < G.max_solvers, "solver index exceeds maxNoGridSolvers");
    auto S = std::make_unique<SolverDev>();
    S->grid_id = grid_id;
    S->solver_idx = solver_idx;
    S->P = *phys;
    S->pc = make_phys_const(*phys);
    S->nd = phys->nd;
    S->nvars = S->pc.nvars;
    S->npos = 2 * phys->nd;
    C.solvers.push_back(std::move(S));
    *solver_id = (int)C.solvers.size() - 1;
    C.solvers.back()->id = *solver_id;
  });
}

int hom3g_solver_set_domain(hom3g_ctx* ctx, int sid, const hom3g_geometry* g, int sign) {
  return guarded([&] {
    Ctx& C = ctx_ref(ctx);
    SolverDev& S = solver_ref(C, sid);
    H3_REQUIRE(!S.built, "domain must be set before build");
    if (g) S.domain_geom = *g;
    S.domain_sign = sign;
  });
}

int hom3g_solver_set_domain_mask(hom3g_ctx* ctx, int sid, const uint8_t* mask, int64_t n) {
  return guarded([&] {
    Ctx& C = ctx_ref(ctx);
    SolverDev& S = solver_ref(C, sid);
    H3_REQUIRE(!S.built, "domain must be set before build");
    S.domain_mask.assign(mask, mask + n);
  });
}

int hom3g_solver_append_bc(hom3g_ctx* ctx, int sid, const hom3g_geometry* g, int kind, const double* params,
                           int nparams, int other_sid, double kratio) {
  return guarded([&] {
    Ctx& C = ctx_ref(ctx);
    SolverDev& S = solver_ref(C, sid);
    H3_REQUIRE(!S.built, "boundary conditions must be appended before build");
    H3_REQUIRE(g, "null geometry");
    H3_REQUIRE(kind >= 0 && kind <= HOM3G_BC_EULER_COUPLING, "unknown boundary condition kind");
    H3_REQUIRE(nparams >= 0 && nparams <= 8, "at most 8 boundary parameters");
    H3_REQUIRE((int)S.bcs.size() < kMaxBC, "too many boundary conditions");
    BCDesc b;
    b.kind = kind;
    b.geom = *g;
    b.nparams = nparams;
    for (int k = 0; k < nparams; ++k) b.params[k] = params[k];
    b.other_sid = other_sid;
    b.kratio = kratio;
    S.bcs.push_back(b);
  });
}

int hom3g_solver_append_bc_host(hom3g_ctx* ctx, int sid, const hom3g_geometry* g, hom3g_bc_callback f, void* user,
                                int slope_hook) {
  return guarded([&] {
    Ctx& C = ctx_ref(ctx);
    SolverDev& S = solver_ref(C, sid);
    H3_REQUIRE(!S.built, "boundary conditions must be appended before build");
    H3_REQUIRE(g && f, "null geometry or callback");
    H3_REQUIRE(slope_hook == HOM3G_SLOPE_DEFAULT || slope_hook == HOM3G_SLOPE_WALL, "unknown slope hook");
    H3_REQUIRE((int)S.bcs.size() < kMaxBC, "too many boundary conditions");
    BCDesc b;
    b.kind = HOM3G_BC_HOST_CALLBACK;
    b.geom = *g;
    b.cb = f;
    b.user = user;
    b.hook = slope_hook;
    S.bcs.push_back(b);
    S.has_host_bc = true;
  });
}

int hom3g_solver_build(hom3g_ctx* ctx, int sid) {
  return guarded([&] {
    Ctx& C = ctx_ref(ctx);
    set_device(C);
    SolverDev& S = solver_ref(C, sid);
    H3_REQUIRE(!S.built, "solver already built");
    GridDev& G = grid_ref(C, S.grid_id);
    solver_build(C, G, S);
    // kernel family
    if (S.P.kernel == HOM3G_KERNEL_BRICK) {
      H3_REQUIRE(S.brick_ok, "brick kernel requested but the topology/physics does not allow it");
      S.kernel = HOM3G_KERNEL_BRICK;
    } else if (S.P.kernel == HOM3G_KERNEL_AUTO && S.brick_ok) {
      S.kernel = HOM3G_KERNEL_BRICK;
    } else {
      S.kernel = HOM3G_KERNEL_GENERIC;
    }
    if (S.kernel != HOM3G_KERNEL_BRICK) S.brick_table = false;
    solver_alloc_state(C, S);
  });
}

int hom3g_solver_impose_ic(hom3g_ctx* ctx, int sid, int kind, const double* params, int np) {
  return guarded([&] {
    Ctx& C = ctx_ref(ctx);
    set_device(C);
    SolverDev& S = built_ref(C, sid);
    solver_impose_ic(C, S, kind, params, np);
    solver_apply_perturbation(C, S);
    ensure_partners(C, S);
    solver_init_time(C, S);
    H3_HIP(hipStreamSynchronize(C.stream));
  });
}

int hom3g_solver_impose_ic_host(hom3g_ctx* ctx, int sid, hom3g_ic_callback f, void* user) {
  return guarded([&] {
    Ctx& C = ctx_ref(ctx);
    set_device(C);
    SolverDev& S = built_ref(C, sid);
    solver_impose_ic_host(C, S, f, user);
    solver_apply_perturbation(C, S);
    ensure_partners(C, S);
    solver_init_time(C, S);
    H3_HIP(hipStreamSynchronize(C.stream));
  });
}

int hom3g_solver_apply_bcs(hom3g_ctx* ctx, int sid) {
  return guarded([&] {
    Ctx& C = ctx_ref(ctx);
    set_device(C);
    SolverDev& S = built_ref(C, sid);
    ensure_partners(C, S);
    solver_fill_ghosts(C, S, S.lhs.p);
    S.dtmin_current = false;
    H3_HIP(hipStreamSynchronize(C.stream));
  });
}

int hom3g_solver_set_perturbation(hom3g_ctx* ctx, int sid, uint64_t seed, double amplitude) {
  return guarded([&] {
    Ctx& C = ctx_ref(ctx);
    SolverDev& S = solver_ref(C, sid);
    S.perturb_seed = seed;
    S.perturb_amp = amplitude;
  });
}

int hom3g_solver_fine_faces(hom3g_ctx* ctx, int sid, int32_t* out, int64_t* nfine) {
  return guarded([&] {
    Ctx& C = ctx_ref(ctx);
    set_device(C);
    SolverDev& S = built_ref(C, sid);
    if (nfine) *nfine = S.nfine;
    if (out && S.nfine > 0) {
      S.fine.download(out, (size_t)S.nfine << (S.nd - 1), C.stream);
      H3_HIP(hipStreamSynchronize(C.stream));
    }
  });
}

int hom3g_solver_info(hom3g_ctx* ctx, int sid, int* nd, int* nvars, int* physics) {
  return guarded([&] {
    Ctx& C = ctx_ref(ctx);
    SolverDev& S = solver_ref(C, sid);
    if (nd) *nd = S.nd;
    if (nvars) *nvars = S.nvars;
    if (physics) *physics = S.P.physics;
  });
}

int hom3g_solver_counts(hom3g_ctx* ctx, int sid, int64_t* ncells, int64_t* first_ghost) {
  return guarded([&] {
    Ctx& C = ctx_ref(ctx);
    SolverDev& S = built_ref(C, sid);
    if (ncells) *ncells = S.ncells;
    if (first_ghost) *first_ghost = S.n_int;
  });
}

int hom3g_solver_kernel(hom3g_ctx* ctx, int sid, int* kernel) {
  return guarded([&] {
    Ctx& C = ctx_ref(ctx);
    *kernel = built_ref(C, sid).kernel;
  });
}

int hom3g_solver_topology(hom3g_ctx* ctx, int sid, int64_t* neighbors, int64_t* node_idx, int32_t* bc_idx,
                          double* x_center, double* length, double* distances) {
  return guarded([&] {
    Ctx& C = ctx_ref(ctx);
    set_device(C);
    SolverDev& S = built_ref(C, sid);
    const int64_t n = S.ncells, cap = S.cap;
    if (neighbors) {
      std::vector<int32_t> t((size_t)S.npos * cap);
      S.nghbr.download(t.data(), t.size(), C.stream);
      H3_HIP(hipStreamSynchronize(C.stream));
      for (int p = 0; p < S.npos; ++p)
        for (int64_t c = 0; c < n; ++c) {
          const int32_t v = t[(size_t)p * cap + c];
          // -1: none; <= -2: face towards finer cells, kept as -(2 + f)
          neighbors[(size_t)p * n + c] = v == -1 ? HOM3G_INVALID_ID : (int64_t)v;
        }
    }
    auto col = [&](auto& buf, auto* out, int ncols) {
      for (int k = 0; k < ncols; ++k) buf.download(out + (size_t)k * n, n, C.stream, (size_t)k * cap);
    };
    if (node_idx) col(S.node_idx, node_idx, 1);
    if (bc_idx) col(S.bc_idx, bc_idx, 1);
    if (x_center) col(S.x_center, x_center, S.nd);
    if (length) col(S.length, length, 1);
    if (distances) col(S.dist, distances, S.npos);
    H3_HIP(hipStreamSynchronize(C.stream));
  });
}

int hom3g_solver_upload_state(hom3g_ctx* ctx, int sid, const double* lhs, int64_t ld) {
  return guarded([&] {
    Ctx& C = ctx_ref(ctx);
    set_device(C);
    SolverDev& S = built_ref(C, sid);
    H3_REQUIRE(lhs && ld >= S.ncells, "bad state buffer");
    for (int v = 0; v < S.nvars; ++v) S.lhs.upload(lhs + (size_t)v * ld, S.ncells, C.stream, (size_t)v * S.cap);
    S.dtmin_current = false;
    ensure_partners(C, S);
    H3_HIP(hipStreamSynchronize(C.stream));
  });
}

int hom3g_solver_download_state(hom3g_ctx* ctx, int sid, double* lhs, int64_t ld) {
  return guarded([&] {
    Ctx& C = ctx_ref(ctx);
    set_device(C);
    SolverDev& S = built_ref(C, sid);
    H3_REQUIRE(lhs && ld >= S.ncells, "bad state buffer");
    if (S.kernel == HOM3G_KERNEL_BRICK) solver_fill_ghosts(C, S, S.lhs.p);
    for (int v = 0; v < S.nvars; ++v) S.lhs.download(lhs + (size_t)v * ld, S.ncells, C.stream, (size_t)v * S.cap);
    H3_HIP(hipStreamSynchronize(C.stream));
  });
}

int hom3g_solver_reset_time(hom3g_ctx* ctx, int sid, double time, double dt, int64_t step) {
  return guarded([&] {
    Ctx& C = ctx_ref(ctx);
    set_device(C);
    SolverDev& S = built_ref(C, sid);
    solver_pull_scalars(C, S);
    S.host_scal.time = time;
    S.host_scal.dt = dt;
    S.host_scal.step = step;
    S.host_scal.forced_step = -1;  // a dt forced before the reset does not carry over
    S.host_scal.diverged_step = -1;
    solver_push_scalars(C, S);
    S.dtmin_current = false;
    H3_HIP(hipStreamSynchronize(C.stream));
  });
}

int hom3g_solver_set_source(hom3g_ctx* ctx, int sid, const double* src, int64_t n) {
  return guarded([&] {
    Ctx& C = ctx_ref(ctx);
    set_device(C);
    SolverDev& S = built_ref(C, sid);
    H3_REQUIRE(S.P.physics == HOM3G_HEAT, "volumetric sources are supported for the heat solver");
    H3_REQUIRE(n == S.n_int, "source length must equal the number of internal cells");
    S.source.alloc(n);
    S.source.upload(src, n, C.stream);
    S.dtmin_current = false;
    H3_HIP(hipStreamSynchronize(C.stream));
  });
}

int hom3g_solver_solve(hom3g_ctx* ctx, int sid, int nsteps) {
  return guarded([&] {
    Ctx& C = ctx_ref(ctx);
    set_device(C);
    SolverDev& S = built_ref(C, sid);
    H3_REQUIRE(nsteps >= 0, "negative step count");
    ensure_partners(C, S);
    solver_solve(C, S, nsteps);
    // solution_diverged (utilities.hpp:73-76, 107-110) on the steps just taken
    solver_pull_scalars(C, S);
    if (S.host_scal.diverged_step >= 0)
      throw Error(HOM3G_ERR_DIVERGED,
                  "Solution diverged! (dt outside [1e-300, 1e300] at step " +
                      std::to_string(S.host_scal.diverged_step) + ")");
  });
}

int hom3g_solver_min_dt(hom3g_ctx* ctx, int sid, double* dt) {
  return guarded([&] {
    Ctx& C = ctx_ref(ctx);
    set_device(C);
    SolverDev& S = built_ref(C, sid);
    *dt = solver_min_dt(C, S);
  });
}

int hom3g_solver_force_dt(hom3g_ctx* ctx, int sid, double dt) {
  return guarded([&] {
    Ctx& C = ctx_ref(ctx);
    set_device(C);
    SolverDev& S = built_ref(C, sid);
    solver_pull_scalars(C, S);
    S.host_scal.forced_step = S.host_scal.step;
    S.host_scal.forced_dt = dt;
    solver_push_scalars(C, S);
    H3_HIP(hipStreamSynchronize(C.stream));
  });
}

int hom3g_solver_time(hom3g_ctx* ctx, int sid, double* time, double* dt, int64_t* step) {
  return guarded([&] {
    Ctx& C = ctx_ref(ctx);
    set_device(C);
    SolverDev& S = built_ref(C, sid);
    solver_pull_scalars(C, S);
    if (time) *time = S.host_scal.time;
    if (dt) *dt = S.host_scal.dt;
    if (step) *step = S.host_scal.step;
  });
}

int hom3g_solver_conservation_sums(hom3g_ctx* ctx, int sid, double* out) {
  return guarded([&] {
    Ctx& C = ctx_ref(ctx);
    set_device(C);
    solver_conservation_sums(C, built_ref(C, sid), out);
  });
}

int hom3g_solver_conservation_sums_noalloc(hom3g_ctx* ctx, int sid, double* sums_dev) {
  return guarded([&] {
    Ctx& C = ctx_ref(ctx);
    set_device(C);
    solver_conservation_sums_noalloc(C, built_ref(C, sid), sums_dev);
  });
}

int hom3g_solver_check_state(hom3g_ctx* ctx, int sid, int64_t* nbad) {
  return guarded([&] {
    Ctx& C = ctx_ref(ctx);
    set_device(C);
    *nbad = solver_check_state(C, built_ref(C, sid));
  });
}

int hom3g_solver_time_solve(hom3g_ctx* ctx, int sid, int nsteps, float* ms) {
  return guarded([&] {
    Ctx& C = ctx_ref(ctx);
    set_device(C);
    SolverDev& S = built_ref(C, sid);
    ensure_partners(C, S);
    H3_HIP(hipStreamSynchronize(C.stream));
    H3_HIP(hipEventRecord(C.ev0, C.stream));
    solver_solve(C, S, nsteps);
    H3_HIP(hipEventRecord(C.ev1, C.stream));
    H3_HIP(hipEventSynchronize(C.ev1));
    H3_HIP(hipEventElapsedTime(ms, C.ev0, C.ev1));
  });
}

int hom3g_solver_profile_solve(hom3g_ctx* ctx, int sid, int nsteps, float* total_ms, float* stage_ms,
                               int* nlaunches) {
  return guarded([&] {
    Ctx& C = ctx_ref(ctx);
    set_device(C);
    SolverDev& S = built_ref(C, sid);
    ensure_partners(C, S);
    H3_HIP(hipStreamSynchronize(C.stream));
    C.prof_on = true;
    C.prof_used = 0;
    try {
      H3_HIP(hipEventRecord(C.ev0, C.stream));
      solver_solve(C, S, nsteps);
      H3_HIP(hipEventRecord(C.ev1, C.stream));
      H3_HIP(hipEventSynchronize(C.ev1));
    } catch (...) {
      C.prof_on = false;
      throw;
    }
    C.prof_on = false;
    H3_HIP(hipEventElapsedTime(total_ms, C.ev0, C.ev1));
    float sum = 0;
    for (size_t k = 0; k + 1 < C.prof_used; k += 2) {
      float ms = 0;
      H3_HIP(hipEventElapsedTime(&ms, C.prof_events[k], C.prof_events[k + 1]));
      sum += ms;
    }
    const int n = (int)(C.prof_used / 2);
    if (nlaunches) *nlaunches = n;
    if (stage_ms) *stage_ms = n ? sum / n : 0.f;
  });
}

int hom3g_copy_calibration(hom3g_ctx* ctx, int64_t bytes, float* ms) {
  return guarded([&] {
    Ctx& C = ctx_ref(ctx);
    set_device(C);
    H3_REQUIRE(bytes >= 1024 && bytes % 32 == 0, "copy size must be a positive multiple of 32 bytes");
    *ms = copy_calibration(C, bytes);
  });
}

int hom3g_copy_bandwidth(hom3g_ctx* ctx, int64_t bytes, int variant, int grid, int block, int reps, float* ms) {
  return guarded([&] {
    Ctx& C = ctx_ref(ctx);
    set_device(C);
    H3_REQUIRE(bytes >= 1024 && bytes % 32 == 0, "copy size must be a positive multiple of 32 bytes");
    H3_REQUIRE(variant >= 0 && variant <= 3, "copy variant must be 0..3");
    H3_REQUIRE(grid >= 1 && block >= 64 && block <= 1024 && block % 64 == 0 && reps >= 1, "bad launch shape");
    *ms = copy_bandwidth(C, bytes, variant, grid, block, reps);
  });
}

int hom3g_solver_destroy(hom3g_ctx* ctx, int sid) {
  return guarded([&] {
    Ctx& C = ctx_ref(ctx);
    set_device(C);
    solver_ref(C, sid);
    H3_HIP(hipStreamSynchronize(C.stream));
    C.solvers[sid].reset();
  });
}

int hom3g_couple_cht_step(hom3g_ctx* ctx, int cns_sid, int heat_sid, double Tr0, int nsteps) {
  return guarded([&] {
    Ctx& C = ctx_ref(ctx);
    set_device(C);
    SolverDev& A = built_ref(C, cns_sid);
    SolverDev& B = built_ref(C, heat_sid);
    H3_REQUIRE(A.P.physics == HOM3G_CNS && B.P.physics == HOM3G_HEAT, "CHT coupling needs a CNS and a heat solver");
    ensure_partners(C, A);
    ensure_partners(C, B);
    for (int i = 0; i < nsteps; ++i) couple_cht_step(C, A, B, Tr0);
    H3_HIP(hipStreamSynchronize(C.stream));
  });
}

int hom3g_couple_same_step(hom3g_ctx* ctx, int s1, int s2, int nsteps) {
  return guarded([&] {
    Ctx& C = ctx_ref(ctx);
    set_device(C);
    SolverDev& A = built_ref(C, s1);
    SolverDev& B = built_ref(C, s2);
    H3_REQUIRE(A.P.physics == B.P.physics, "same-physics coupling needs two solvers of one physics");
    ensure_partners(C, A);
    ensure_partners(C, B);
    for (int i = 0; i < nsteps; ++i) couple_same_step(C, A, B);
    H3_HIP(hipStreamSynchronize(C.stream));
  });
}

// ---- conjugate heat transfer objective, tangent and adjoint --------------------
int hom3g_cht_objective(hom3g_ctx* ctx, int cns_sid, int heat_sid, double Tr0, int nsteps, const double* dts,
                        double* dts_out, double* J) {
  return guarded([&] {
    Ctx& C = ctx_ref(ctx);
    set_device(C);
    SolverDev& A = built_ref(C, cns_sid);
    SolverDev& B = built_ref(C, heat_sid);
    H3_REQUIRE(nsteps >= 0 && J, "bad argument");
    ensure_partners(C, A);
    ensure_partners(C, B);
    ChtRunStats st;
    cht_objective_run(C, A, B, Tr0, nsteps, dts, dts_out, st);
    *J = st.J;
  });
}

int hom3g_cht_tangent(hom3g_ctx* ctx, int cns_sid, int heat_sid, double Tr0, int nsteps, const double* dts,
                      const double* dsource, double* J, double* dJ) {
  return guarded([&] {
    Ctx& C = ctx_ref(ctx);
    set_device(C);
    SolverDev& A = built_ref(C, cns_sid);
    SolverDev& B = built_ref(C, heat_sid);
    H3_REQUIRE(nsteps >= 0 && J && dJ, "bad argument");
    ensure_partners(C, A);
    ensure_partners(C, B);
    cht_tangent_run(C, A, B, Tr0, nsteps, dts, dsource, J, dJ);
  });
}

int hom3g_cht_adjoint(hom3g_ctx* ctx, int cns_sid, int heat_sid, double Tr0, int nsteps, const double* dts,
                      int ckpt_every, double* J, double* grad_source, float* ms_forward, float* ms_reverse) {
  return guarded([&] {
    Ctx& C = ctx_ref(ctx);
    set_device(C);
    SolverDev& A = built_ref(C, cns_sid);
    SolverDev& B = built_ref(C, heat_sid);
    H3_REQUIRE(nsteps >= 1 && J, "bad argument");
    ensure_partners(C, A);
    ensure_partners(C, B);
    ChtRunStats st;
    cht_adjoint_run(C, A, B, Tr0, nsteps, dts, ckpt_every, grad_source, st);
    *J = st.J;
    if (ms_forward) *ms_forward = st.ms_forward;
    if (ms_reverse) *ms_reverse = st.ms_reverse;
  });
}

int hom3g_set_checkpoint_memory(hom3g_ctx* ctx, int where) {
  return guarded([&] {
    Ctx& C = ctx_ref(ctx);
    H3_REQUIRE(where == HOM3G_CKPT_AUTO || where == HOM3G_CKPT_DEVICE || where == HOM3G_CKPT_HOST,
               "checkpoint memory must be HOM3G_CKPT_AUTO, _DEVICE or _HOST");
    C.ckpt_memory = where;
  });
}

int hom3g_cht_misfit_objective(hom3g_ctx* ctx, int cns_sid, int heat_sid, double Tr0, int nsteps,
                               const double* dts, const double* T_obs, double* J) {
  return guarded([&] {
    Ctx& C = ctx_ref(ctx);
    set_device(C);
    SolverDev& A = built_ref(C, cns_sid);
    SolverDev& B = built_ref(C, heat_sid);
    H3_REQUIRE(nsteps >= 0 && J && T_obs, "bad argument");
    ensure_partners(C, A);
    ensure_partners(C, B);
    ChtRunStats st;
    cht_objective_run(C, A, B, Tr0, nsteps, dts, nullptr, st, T_obs);
    *J = st.J;
  });
}

int hom3g_cht_misfit_tangent(hom3g_ctx* ctx, int cns_sid, int heat_sid, double Tr0, int nsteps, const double* dts,
                             const double* T_obs, const double* dsource, double* J, double* dJ) {
  return guarded([&] {
    Ctx& C = ctx_ref(ctx);
    set_device(C);
    SolverDev& A = built_ref(C, cns_sid);
    SolverDev& B = built_ref(C, heat_sid);
    H3_REQUIRE(nsteps >= 0 && J && dJ && T_obs, "bad argument");
    ensure_partners(C, A);
    ensure_partners(C, B);
    cht_tangent_run(C, A, B, Tr0, nsteps, dts, dsource, J, dJ, T_obs);
  });
}

int hom3g_cht_misfit_adjoint(hom3g_ctx* ctx, int cns_sid, int heat_sid, double Tr0, int nsteps, const double* dts,
                             int ckpt_every, const double* T_obs, double* J, double* grad_source, float* ms_forward,
                             float* ms_reverse) {
  return guarded([&] {
    Ctx& C = ctx_ref(ctx);
    set_device(C);
    SolverDev& A = built_ref(C, cns_sid);
    SolverDev& B = built_ref(C, heat_sid);
    H3_REQUIRE(nsteps >= 1 && J && T_obs, "bad argument");
    ensure_partners(C, A);
    ensure_partners(C, B);
    ChtRunStats st;
    cht_adjoint_run(C, A, B, Tr0, nsteps, dts, ckpt_every, grad_source, st, T_obs);
    *J = st.J;
    if (ms_forward) *ms_forward = st.ms_forward;
    if (ms_reverse) *ms_reverse = st.ms_reverse;
  });
}

// ---- batched instances -------------------------------------------------------
static BatchDev& batch_ref(Ctx& C, int id) {
  H3_REQUIRE(id >= 0 && id < (int)C.batches.size() && C.batches[id], "invalid batch id");
  return *C.batches[id];
}

int hom3g_batch_create(hom3g_ctx* ctx, const int* sids, int n, int precision, int* batch_id) {
  return guarded([&] {
    Ctx& C = ctx_ref(ctx);
    set_device(C);
    H3_REQUIRE(sids && batch_id && n > 0, "null argument");
    H3_REQUIRE(precision == HOM3G_FP64 || precision == HOM3G_FP32_STATE, "unknown precision");
    std::vector<SolverDev*> inst;
    for (int i = 0; i < n; ++i) inst.push_back(&built_ref(C, sids[i]));
    auto B = std::make_unique<BatchDev>();
    batch_create(C, inst, precision, *B);
    C.batches.push_back(std::move(B));
    *batch_id = (int)C.batches.size() - 1;
  });
}

int hom3g_batch_set_config(hom3g_ctx* ctx, int bid, int block, int tile, int grid) {
  return guarded([&] {
    Ctx& C = ctx_ref(ctx);
    set_device(C);
    batch_set_config(C, batch_ref(C, bid), block, tile, grid);
  });
}

int hom3g_batch_get_config(hom3g_ctx* ctx, int bid, int* block, int* tile, int* grid, int64_t* nitems) {
  return guarded([&] {
    Ctx& C = ctx_ref(ctx);
    BatchDev& B = batch_ref(C, bid);
    if (block) *block = B.block;
    if (tile) *tile = B.tile;
    if (grid) *grid = B.grid;
    if (nitems) *nitems = (int64_t)B.h_items.size();
  });
}

int hom3g_batch_autotune(hom3g_ctx* ctx, int bid, int* block, int* tile, int* grid, float* stage_ms) {
  return guarded([&] {
    Ctx& C = ctx_ref(ctx);
    set_device(C);
    batch_autotune(C, batch_ref(C, bid), block, tile, grid, stage_ms);
  });
}

int hom3g_batch_solve(hom3g_ctx* ctx, int bid, int nsteps, int use_graph) {
  return guarded([&] {
    Ctx& C = ctx_ref(ctx);
    set_device(C);
    batch_solve(C, batch_ref(C, bid), nsteps, use_graph != 0);
    H3_HIP(hipStreamSynchronize(C.stream));
  });
}

int hom3g_batch_time_solve(hom3g_ctx* ctx, int bid, int nsteps, int use_graph, float* ms_per_step) {
  return guarded([&] {
    Ctx& C = ctx_ref(ctx);
    set_device(C);
    H3_REQUIRE(ms_per_step != nullptr, "null argument");
    *ms_per_step = batch_time_solve(C, batch_ref(C, bid), nsteps, use_graph != 0);
  });
}

int hom3g_batch_counts(hom3g_ctx* ctx, int bid, int* ninstances, int64_t* total_cells) {
  return guarded([&] {
    Ctx& C = ctx_ref(ctx);
    BatchDev& B = batch_ref(C, bid);
    if (ninstances) *ninstances = B.ninst;
    if (total_cells) *total_cells = B.ncell;
  });
}

int hom3g_batch_download_state(hom3g_ctx* ctx, int bid, int inst, double* out, int64_t n) {
  return guarded([&] {
    Ctx& C = ctx_ref(ctx);
    set_device(C);
    BatchDev& B = batch_ref(C, bid);
    H3_REQUIRE(inst >= 0 && inst < B.ninst && out, "bad argument");
    H3_REQUIRE(n == B.nint[inst], "n must equal the instance's internal cell count");
    batch_download(C, B, inst, out);
  });
}

int hom3g_batch_time(hom3g_ctx* ctx, int bid, int inst, double* time, double* dt, int64_t* step) {
  return guarded([&] {
    Ctx& C = ctx_ref(ctx);
    set_device(C);
    DevScalars s;
    batch_scalars(C, batch_ref(C, bid), inst, &s);
    if (time) *time = s.time;
    if (dt) *dt = s.dt;
    if (step) *step = s.step;
  });
}

int hom3g_batch_destroy(hom3g_ctx* ctx, int bid) {
  return guarded([&] {
    Ctx& C = ctx_ref(ctx);
    set_device(C);
    BatchDev& B = batch_ref(C, bid);
    H3_HIP(hipStreamSynchronize(C.stream));
    batch_release(B);
    C.batches[bid].reset();
  });
}

}  // extern "C"
