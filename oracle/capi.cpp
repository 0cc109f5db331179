// CPU ORACLE — test infrastructure only (see oracle.hpp).
// Plain C entry points so the pytest harness can drive the oracle through
// ctypes.  Errors are reported as negative return codes plus
// orc_last_error(); nothing here calls exit().
#include <type_traits>
#include "oracle.hpp"
#include <chrono>
#include <cstring>
#include <stdexcept>
#include <string>
#ifdef _OPENMP
#include <omp.h>
#endif

using namespace orc;

namespace {
thread_local std::string g_err;
template <class F> int guard(F&& f) {
  try { f(); return 0; } catch (const std::exception& e) { g_err = e.what(); return -1; }
}
}  // namespace

extern "C" {

const char* orc_last_error() { return g_err.c_str(); }

// Threads of the OpenMP timing build (1 in the checker builds, which are
// compiled without OpenMP); returns the number in effect.
int orc_set_threads(int n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
  return omp_get_max_threads();
#else
  (void)n;
  return 1;
#endif
}

// ---- grid ----------------------------------------------------------------
void* orc_grid_create(int nd, const double* xmin, const double* xmax, int max_solvers) {
  Num lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
  for (int d = 0; d < nd && d < 3; ++d) { lo[d] = xmin[d]; hi[d] = xmax[d]; }
  return new Grid(nd, lo, hi, max_solvers);
}
void orc_grid_destroy(void* g) { delete static_cast<Grid*>(g); }
int orc_grid_min_level(void* g, int level) { return guard([&] { static_cast<Grid*>(g)->min_level(level); }); }
int64_t orc_grid_size(void* g) { return static_cast<Grid*>(g)->size(); }
int64_t orc_grid_refine(void* g, int64_t n) {
  int64_t r = -1;
  guard([&] { r = static_cast<Grid*>(g)->refine(n); });
  return r;
}
int orc_grid_refine_region(void* g, int max_level, const GeomDesc* geom, double band) {
  return guard([&] { static_cast<Grid*>(g)->refine_region(max_level, *geom, band); });
}
int orc_grid_balance(void* g) { return guard([&] { static_cast<Grid*>(g)->balance_2to1(); }); }
void orc_grid_tree(void* g, int64_t* parent, int64_t* first_child) {
  auto* G = static_cast<Grid*>(g);
  for (Ind n = 0; n < G->size(); ++n) { parent[n] = G->parent[n]; first_child[n] = G->first_child[n]; }
}
int64_t orc_grid_find_samelvl_neighbor(void* g, int64_t n, int pos) {
  return static_cast<Grid*>(g)->find_samelvl_neighbor(n, pos);
}
int orc_grid_level(void* g, int64_t n) { return static_cast<Grid*>(g)->level(n); }
int orc_grid_position_in_parent(void* g, int64_t n) { return static_cast<Grid*>(g)->position_in_parent(n); }
int orc_grid_is_leaf(void* g, int64_t n) { return static_cast<Grid*>(g)->is_leaf(n) ? 1 : 0; }
void orc_grid_cell_coordinates(void* g, int64_t n, double* x) {
  auto* G = static_cast<Grid*>(g);
  Num xn[3] = {0, 0, 0};
  G->cell_coordinates(n, xn);
  for (int d = 0; d < G->nd; ++d) x[d] = (double)xn[d];
}
double orc_grid_cell_length(void* g, int64_t n) { return static_cast<Grid*>(g)->cell_length(n); }
int orc_grid_is_cut_by(void* g, int64_t n, const GeomDesc* geom) { return static_cast<Grid*>(g)->is_cut_by(n, *geom) ? 1 : 0; }
double orc_signed_distance(const GeomDesc* geom, const double* x, int nd) {
  Num xn[3] = {0, 0, 0};
  for (int d = 0; d < nd && d < 3; ++d) xn[d] = x[d];
  return (double)signed_distance(*geom, xn, nd);
}
void orc_grid_all_neighbors(void* g, int64_t* out) {
  auto* G = static_cast<Grid*>(g);
  const Ind N = G->size();
  for (Ind n = 0; n < N; ++n)
    for (int p = 0; p < G->npos(); ++p) out[n * G->npos() + p] = G->find_samelvl_neighbor(n, p);
}

// ---- solver --------------------------------------------------------------
void* orc_solver_create(void* g, int sidx, const PhysDesc* p) {
  return new Solver(*static_cast<Grid*>(g), sidx, *p);
}
void orc_solver_destroy(void* s) { delete static_cast<Solver*>(s); }
int orc_solver_append_bc(void* s, const GeomDesc* geom, int kind, const double* params, int np,
                         void* other, double kratio) {
  return guard([&] {
    BC bc;
    bc.geom = *geom;
    bc.kind = kind;
    bc.params.assign(params, params + np);
    bc.other = static_cast<Solver*>(other);
    bc.kratio = kratio;
    static_cast<Solver*>(s)->bcs.push_back(bc);
  });
}
int orc_solver_set_bc_other(void* s, int k, void* other) {
  return guard([&] { static_cast<Solver*>(s)->bcs.at(k).other = static_cast<Solver*>(other); });
}
void orc_solver_set_domain(void* s, const GeomDesc* geom, int sign) {
  auto* S = static_cast<Solver*>(s);
  if (geom) S->domain_geom = *geom;
  S->domain_sign = sign;
}
void orc_solver_set_domain_mask(void* s, const uint8_t* mask, int64_t n) {
  static_cast<Solver*>(s)->domain_mask.assign(mask, mask + n);
}
void orc_solver_set_ic(void* s, int kind, const double* params, int np, ic_callback_t cb, void* ctx) {
  auto* S = static_cast<Solver*>(s);
  S->ic_kind = kind;
  S->ic_params.assign(params, params + np);
  S->ic_cb = cb;
  S->ic_ctx = ctx;
}
void orc_solver_set_ghost_order(void* s, int order) { static_cast<Solver*>(s)->ghost_order = order; }
int orc_solver_initialize(void* s) { return guard([&] { static_cast<Solver*>(s)->initialize(); }); }
int orc_solver_solve(void* s, int nsteps) {
  return guard([&] { for (int i = 0; i < nsteps; ++i) static_cast<Solver*>(s)->solve(); });
}
int orc_solver_apply_bcs_lhs(void* s) { return guard([&] { auto* S = static_cast<Solver*>(s); S->apply_bcs(S->lhs.data()); }); }
int orc_solver_evolve(void* s) { return guard([&] { static_cast<Solver*>(s)->evolve(); }); }
int orc_solver_set_dt(void* s) { return guard([&] { static_cast<Solver*>(s)->set_dt(); }); }
void orc_solver_advance(void* s) { auto* S = static_cast<Solver*>(s); S->time_ += S->dt_; ++S->step_; }
double orc_solver_compute_dt(void* s) { return static_cast<Solver*>(s)->compute_dt(); }
void orc_solver_force_dt(void* s, double dt) { static_cast<Solver*>(s)->force_dt(dt); }
double orc_solver_dt(void* s) { return static_cast<Solver*>(s)->dt_; }
double orc_solver_time(void* s) { return static_cast<Solver*>(s)->time_; }
int64_t orc_solver_step(void* s) { return static_cast<Solver*>(s)->step_; }
void orc_solver_set_time(void* s, double t, double dt, int64_t step) {
  auto* S = static_cast<Solver*>(s); S->time_ = t; S->dt_ = dt; S->step_ = step;
}
int64_t orc_solver_ncells(void* s) { return static_cast<Solver*>(s)->ncells; }
int64_t orc_solver_first_gc(void* s) { return static_cast<Solver*>(s)->firstGC; }
int orc_solver_nvars(void* s) { return static_cast<Solver*>(s)->nvars; }
int64_t orc_solver_cap(void* s) { return static_cast<Solver*>(s)->cap; }
void orc_solver_quantities(void* s, double* out) {
  auto* S = static_cast<Solver*>(s);
  out[0] = S->qty.T_infinity(); out[1] = S->qty.rho_infinity(); out[2] = S->qty.u_infinity();
  out[3] = S->qty.p_infinity(); out[4] = S->qty.Re0; out[5] = S->qty.sOverTref; out[6] = S->qty.sOverTrefP1;
}

// Copy a SoA field out with row count = ncells (column-major, ld = ncells).
void orc_solver_get(void* s, const char* field, void* out) {
  auto* S = static_cast<Solver*>(s);
  const Ind n = S->ncells, cap = S->cap;
  auto cp = [&](const auto& arr, int ncols) {
    // floating-point fields leave as double whatever the oracle computes in
    using T0 = typename std::decay<decltype(arr[0])>::type;
    using T = typename std::conditional<std::is_floating_point<T0>::value, double, T0>::type;
    T* o = static_cast<T*>(out);
    for (int k = 0; k < ncols; ++k)
      for (Ind c = 0; c < n; ++c) o[(size_t)k * n + c] = (T)arr[(size_t)k * cap + c];
  };
  const std::string f(field);
  if (f == "lhs") cp(S->lhs, S->nvars);
  else if (f == "rhs") cp(S->rhs, S->nvars);
  else if (f == "neighbors") cp(S->neighbors, S->npos);
  else if (f == "x_center") cp(S->x_center, S->nd);
  else if (f == "distances") cp(S->distances, S->npos);
  else if (f == "length") cp(S->length, 1);
  else if (f == "node_idx") cp(S->node_idx, 1);
  else if (f == "bc_idx") {
    int32_t* o = static_cast<int32_t*>(out);
    for (Ind c = 0; c < n; ++c) o[c] = S->bc_idx[c];
  }
}
// fine-face table of the 2:1 refinement extension (row-major, 2^(nd-1) per face)
int64_t orc_solver_nfine(void* s) {
  auto* S = static_cast<Solver*>(s);
  return (int64_t)(S->fine.size() / S->nfine_per());
}
void orc_solver_fine(void* s, int64_t* out) {
  auto* S = static_cast<Solver*>(s);
  for (size_t k = 0; k < S->fine.size(); ++k) out[k] = S->fine[k];
}
void orc_solver_set_perturbation(void* s, uint64_t seed, double amp) {
  auto* S = static_cast<Solver*>(s);
  S->perturb_seed = seed;
  S->perturb_amp = amp;
}
int orc_couple_cns_heat_fixed(double Tr0, double dt, void* cns, void* heat) {
  return guard([&] { couple_cns_heat_fixed(Tr0, dt, *static_cast<Solver*>(cns), *static_cast<Solver*>(heat)); });
}
double orc_cht_wall_heat_flux(void* heat) { return cht_wall_heat_flux(*static_cast<Solver*>(heat)); }
uint64_t orc_node_morton(void* g, int64_t n) { return node_morton(*static_cast<Grid*>(g), n); }
void orc_solver_set_lhs(void* s, const double* in) {
  auto* S = static_cast<Solver*>(s);
  const Ind n = S->ncells, cap = S->cap;
  for (int v = 0; v < S->nvars; ++v)
    for (Ind c = 0; c < n; ++c) S->lhs[(size_t)v * cap + c] = in[(size_t)v * n + c];
}
void orc_solver_set_source(void* s, const double* src, int64_t n) {
  static_cast<Solver*>(s)->source.assign(src, src + n);
}
void orc_solver_conservation_sums(void* s, double* out) {
  auto* S = static_cast<Solver*>(s);
  Num o[5] = {0, 0, 0, 0, 0};
  S->conservation_sums(o);
  for (int v = 0; v < S->nvars; ++v) out[v] = (double)o[v];
}
int64_t orc_solver_periodic_image(void* s, int64_t g) { return static_cast<Solver*>(s)->periodic_image(g); }
int64_t orc_solver_cht_partner(void* s, int64_t g) {
  auto* S = static_cast<Solver*>(s);
  int64_t r = -1;
  guard([&] { r = S->cht_partner(g, S->bcs.at(S->bc_idx[g])); });
  return r;
}
// One flux evaluation (for unit tests of the numerical fluxes)
void orc_solver_face_flux(void* s, int64_t l, int64_t r, int d, double* out) {
  auto* S = static_cast<Solver*>(s);
  Num o[5] = {0, 0, 0, 0, 0};
  S->compute_num_flux(S->lhs.data(), l, r, d, S->length[r], S->dt_, o);
  for (int v = 0; v < S->nvars; ++v) out[v] = (double)o[v];
}
void orc_solver_num_flux(void* s, int64_t c, double* out) {
  auto* S = static_cast<Solver*>(s);
  Num o[5] = {0, 0, 0, 0, 0};
  S->num_flux(S->lhs.data(), c, o);
  for (int v = 0; v < S->nvars; ++v) out[v] = (double)o[v];
}
// AUSM pieces on explicit states (pins of the flux formulas)
void orc_ausm_split(int sgn, double M, double* m, double* p) {
  *m = (double)ausm_m_split(sgn, M);
  *p = (double)ausm_p_split(sgn, M);
}
void orc_solver_ausm_states(void* s, const double* ql, const double* qr, int d, double* out) {
  auto* S = static_cast<Solver*>(s);
  Num a[5] = {0, 0, 0, 0, 0}, b[5] = {0, 0, 0, 0, 0}, o[5] = {0, 0, 0, 0, 0};
  for (int v = 0; v < S->nvars; ++v) { a[v] = ql[v]; b[v] = qr[v]; }
  S->ausm_states(a, b, d, o);
  for (int v = 0; v < S->nvars; ++v) out[v] = (double)o[v];
}
// Point value of the solver's initial condition (conserved variables)
int orc_solver_ic_point(void* s, const double* x, double* out) {
  return guard([&] {
    auto* S = static_cast<Solver*>(s);
    Num xn[3] = {0, 0, 0}, o[5] = {0, 0, 0, 0, 0};
    for (int d = 0; d < S->nd; ++d) xn[d] = x[d];
    S->ic_point(xn, o);
    for (int v = 0; v < S->nvars; ++v) out[v] = (double)o[v];
  });
}
double orc_solver_slope(void* s, int64_t c, int v, int dir) {
  auto* S = static_cast<Solver*>(s);
  return S->slope(S->lhs.data(), c, v, dir);
}
double orc_solver_cell_dt(void* s, int64_t c) {
  auto* S = static_cast<Solver*>(s);
  return S->cell_dt(S->lhs.data(), c);
}

// coupling
void orc_couple_cns_heat(double Tr0, void* cns, void* heat) {
  couple_cns_heat(Tr0, *static_cast<Solver*>(cns), *static_cast<Solver*>(heat));
}
void orc_couple_same(void* a, void* b) { couple_same(*static_cast<Solver*>(a), *static_cast<Solver*>(b)); }

// Time `nsteps` calls of solve() (step loop only), returns seconds.
double orc_time_solve(void* s, int nsteps) {
  auto* S = static_cast<Solver*>(s);
  const auto t0 = std::chrono::steady_clock::now();
  for (int i = 0; i < nsteps; ++i) S->solve();
  const auto t1 = std::chrono::steady_clock::now();
  return std::chrono::duration<double>(t1 - t0).count();
}

// PRNG helpers
double orc_prng_uniform(uint64_t seed, uint64_t stream, uint64_t counter) { return prng_uniform(seed, stream, counter); }
uint64_t orc_morton_encode(int nd, const uint32_t* ijk) { return morton_encode(nd, ijk); }

}  // extern "C"
