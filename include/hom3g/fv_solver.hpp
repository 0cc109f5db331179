// C++ drop-in facade for hom3's finite-volume solver over the hom3g C ABI.
//
// Mirrors the member surface a hom3 driver uses (src/solver/fv/solver.hpp,
// src/solver/fv/utilities.hpp, src/solver/fv/coupling.hpp) without Eigen or
// Boost: a hom3 test such as euler_fv_solver_test.cpp reads the same with
// hom3g::fv in place of hom3::solver::fv.  Header-only; link libhom3g.so.
//
//   hom3g::Context ctx(0);
//   hom3g::Grid grid(ctx, 3, {0,0,0}, {1,1,1}, /*minRefLevel=*/5);
//   hom3g::fv::Solver s(grid, hom3g::fv::physics::euler, props);   // Solver(SolverIdx, io::Properties)
//   s.set_initial_condition(HOM3G_IC_SHOCK_TUBE, {...});
//   hom3g::fv::append_bcs(s, hom3g::fv::Condition::neumann());        // utilities.hpp:27-33
//   hom3g::fv::run_solver(grid, s, maxNoTimeSteps);                   // utilities.hpp:98-113
#pragma once
#include <cmath>
#include <limits>
#include <stdexcept>
#include <string>
#include <vector>

#include "../hom3g.h"

namespace hom3g {

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

inline void check(int rc) {
  if (rc != HOM3G_OK) throw Error(rc, hom3g_last_error());
}

class Context {
 public:
  explicit Context(int device = 0) { check(hom3g_create(&h_, device)); }
  ~Context() { if (h_) hom3g_destroy(h_); }
  Context(const Context&) = delete;
  Context& operator=(const Context&) = delete;
  hom3g_ctx* get() const { return h_; }
  void synchronize() { check(hom3g_synchronize(h_)); }

 private:
  hom3g_ctx* h_ = nullptr;
};

// grid::CartesianHSP<nd> + generation::MinLevel (grid/grid.hpp:147-221)
class Grid {
 public:
  Grid(Context& ctx, int nd, std::vector<double> xmin, std::vector<double> xmax, int min_level,
       int max_solvers = 1)
      : ctx_(ctx), nd_(nd), xmin_(xmin), xmax_(xmax) {
    check(hom3g_grid_create(ctx.get(), nd, xmin.data(), xmax.data(), min_level, max_solvers, &id_));
    root_length_ = std::numeric_limits<double>::epsilon();
    for (int d = 0; d < nd; ++d) {
      root_length_ = std::max(root_length_, xmax[d] - xmin[d]);
      root_center_[d] = xmin[d] + 0.5 * (xmax[d] - xmin[d]);
    }
  }
  int id() const { return id_; }
  int nd() const { return nd_; }
  Context& context() const { return ctx_; }
  double cell_length_at_level(int level) const { return root_length_ / std::pow(2, level); }
  double root_length() const { return root_length_; }
  const double* root_center() const { return root_center_; }
  int64_t size() const {
    int64_t n = 0;
    check(hom3g_grid_size(ctx_.get(), id_, &n));
    return n;
  }
  // find_samelvl_neighbor for every node, row-major (nnodes x 2nd)
  std::vector<int64_t> neighbors() const {
    std::vector<int64_t> out(size() * 2 * nd_);
    check(hom3g_grid_neighbors(ctx_.get(), id_, out.data()));
    return out;
  }

 private:
  Context& ctx_;
  int nd_, id_ = -1;
  std::vector<double> xmin_, xmax_;
  double root_length_ = 0, root_center_[3] = {0, 0, 0};
};

namespace fv {

enum class physics { euler = HOM3G_EULER, cns = HOM3G_CNS, heat = HOM3G_HEAT };

// io::Properties keys the hot path reads
struct Properties {
  double timeEnd = 1e30, CFL = 0.2, CFL_viscous = 0.4, gamma = 1.4, Rspecific = 287.058, T0 = 273.15,
         rho0 = 1.225, Minfty = 0.1, ReInfty = 100.0, Pr0 = 0.72, Sutherland0 = 110.4, diffusivity = 1.0,
         T_ref = 273.15;
  int64_t maxNoCells = 0;
  int flux = HOM3G_FLUX_AUSM, time_integration = HOM3G_TI_EULER_FWD, recon = HOM3G_RECON_NONE,
      limiter = HOM3G_LIM_MINMOD, precision = HOM3G_FP64, kernel = HOM3G_KERNEL_AUTO;
};

class Solver;

// A boundary Condition: one of the reference's condition classes with
// constant surface distributions, or a coupling condition.
struct Condition {
  int kind;
  std::vector<double> params;
  const Solver* other = nullptr;
  double kratio = 1.0;
  static Condition neumann(std::vector<double> flux = {}) { return {HOM3G_BC_NEUMANN, flux}; }
  static Condition dirichlet(std::vector<double> v) { return {HOM3G_BC_DIRICHLET, v}; }
  static Condition isentropic_vortex() { return {HOM3G_BC_ISENTROPIC_VORTEX, {}}; }
  static Condition inflow(std::vector<double> pv) { return {HOM3G_BC_INFLOW, pv}; }
  static Condition outflow(std::vector<double> pv) { return {HOM3G_BC_OUTFLOW, pv}; }
  static Condition adiabatic_noslip() { return {HOM3G_BC_ADIABATIC_NOSLIP, {}}; }
  static Condition adiabatic_slip() { return {HOM3G_BC_ADIABATIC_SLIP, {}}; }
  static Condition isothermal_noslip(double T) { return {HOM3G_BC_ISOTHERMAL_NOSLIP, {T}}; }
  static Condition heat_dirichlet(double T) { return {HOM3G_BC_HEAT_DIRICHLET, {T}}; }
  static Condition heat_neumann(double q) { return {HOM3G_BC_HEAT_NEUMANN, {q}}; }
  static Condition periodic() { return {HOM3G_BC_PERIODIC, {}}; }
};

inline hom3g_geometry edge(std::vector<double> point, std::vector<double> normal) {
  hom3g_geometry g{};
  g.kind = HOM3G_GEOM_EDGE;
  for (size_t d = 0; d < point.size(); ++d) { g.a[d] = point[d]; g.b[d] = normal[d]; }
  return g;
}
inline hom3g_geometry square(std::vector<double> centre, std::vector<double> lengths, bool inverted = false) {
  hom3g_geometry g{};
  g.kind = HOM3G_GEOM_SQUARE;
  g.invert = inverted ? 1 : 0;
  for (size_t d = 0; d < centre.size(); ++d) { g.a[d] = centre[d]; g.b[d] = lengths[d]; }
  return g;
}

// solver::fv::Solver<PhysicsTT, TimeIntegration> (solver.hpp:60-145)
class Solver {
 public:
  Solver(Grid& grid, physics ph, const Properties& p, int solver_idx = 0) : grid_(grid), props_(p) {
    hom3g_physics d{};
    d.physics = static_cast<int>(ph);
    d.nd = grid.nd();
    d.flux = (ph == physics::heat) ? HOM3G_FLUX_THREE_POINT : p.flux;
    d.ti = p.time_integration;
    d.recon = p.recon;
    d.limiter = p.limiter;
    d.gamma = p.gamma;
    d.cfl = p.CFL;
    d.cfl_viscous = p.CFL_viscous;
    d.Rspecific = p.Rspecific;
    d.T0 = p.T0;
    d.rho0 = p.rho0;
    d.Minfty = p.Minfty;
    d.ReInfty = p.ReInfty;
    d.Pr0 = p.Pr0;
    d.Sutherland0 = p.Sutherland0;
    d.diffusivity = p.diffusivity;
    d.T_ref = p.T_ref;
    d.time_end = p.timeEnd;
    d.max_cells = p.maxNoCells;
    d.precision = p.precision;
    d.kernel = p.kernel;
    nvars_ = (ph == physics::heat) ? 1 : grid.nd() + 2;
    check(hom3g_solver_create(ctx(), grid.id(), solver_idx, &d, &id_));
  }

  // -- input (solver.hpp:150-169)
  void set_initial_condition(int ic_kind, std::vector<double> params) {
    ic_kind_ = ic_kind;
    ic_params_ = std::move(params);
  }
  void set_initial_domain(const hom3g_geometry& g, int sign) { check(hom3g_solver_set_domain(ctx(), id_, &g, sign)); }
  void append_bc(const hom3g_geometry& g, const Condition& c) {
    check(hom3g_solver_append_bc(ctx(), id_, &g, c.kind, c.params.data(), (int)c.params.size(),
                                 c.other ? c.other->id_ : -1, c.kratio));
  }

  // -- solver interface (solver.hpp:100-145)
  void initialize() {
    check(hom3g_solver_build(ctx(), id_));
    check(hom3g_solver_impose_ic(ctx(), id_, ic_kind_, ic_params_.data(), (int)ic_params_.size()));
  }
  void solve() { check(hom3g_solver_solve(ctx(), id_, 1)); }
  void solve(int nsteps) { check(hom3g_solver_solve(ctx(), id_, nsteps)); }
  int64_t step() const { double t, dt; int64_t s; check(hom3g_solver_time(ctx(), id_, &t, &dt, &s)); return s; }
  double time() const { double t, dt; int64_t s; check(hom3g_solver_time(ctx(), id_, &t, &dt, &s)); return t; }
  double dt() const { double t, dt; int64_t s; check(hom3g_solver_time(ctx(), id_, &t, &dt, &s)); return dt; }
  void dt(double forced) { check(hom3g_solver_force_dt(ctx(), id_, forced)); }
  double min_dt() const { double d; check(hom3g_solver_min_dt(ctx(), id_, &d)); return d; }
  double final_time() const { return props_.timeEnd; }
  int id() const { return id_; }
  int nvars() const { return nvars_; }
  hom3g_ctx* context_handle() const { return ctx(); }

  // -- data: cells().lhs in the reference's column-major layout
  int64_t ncells() const { int64_t n, f; check(hom3g_solver_counts(ctx(), id_, &n, &f)); return n; }
  int64_t first_ghost() const { int64_t n, f; check(hom3g_solver_counts(ctx(), id_, &n, &f)); return f; }
  std::vector<double> lhs() const {
    const int64_t n = ncells();
    std::vector<double> q((size_t)nvars_ * n);
    check(hom3g_solver_download_state(ctx(), id_, q.data(), n));
    return q;
  }
  void set_lhs(const std::vector<double>& q, int64_t ld) { check(hom3g_solver_upload_state(ctx(), id_, q.data(), ld)); }

 private:
  hom3g_ctx* ctx() const { return grid_.context().get(); }
  Grid& grid_;
  Properties props_;
  int id_ = -1, nvars_ = 0, ic_kind_ = HOM3G_IC_CONSTANT_PV;
  std::vector<double> ic_params_;
};

// grid/helpers.hpp:112-160 make_boundaries + utilities.hpp:27-33 append_bcs
inline std::vector<hom3g_geometry> cube_boundaries(const Grid& g) {
  const double L = g.root_length(), eps = L * 1e-15;
  const double x_min = g.root_center()[0] - 0.5 * L + eps;
  const double x_max = g.root_center()[0] + 0.5 * L - eps;
  const double off = x_min + 0.5 * L;
  std::vector<hom3g_geometry> out;
  for (int d = 0; d < g.nd(); ++d)
    for (int side = 0; side < 2; ++side) {
      std::vector<double> a(g.nd(), off), n(g.nd(), 0.0);
      a[d] = side ? x_max : x_min;
      n[d] = side ? -1.0 : 1.0;
      out.push_back(edge(a, n));
    }
  return out;
}
inline void append_bcs(Solver& s, const std::vector<Condition>& conds, const Grid& g) {
  const auto geos = cube_boundaries(g);
  for (size_t k = 0; k < geos.size(); ++k) s.append_bc(geos[k], conds[conds.size() == 1 ? 0 : k]);
}

// utilities.hpp:73-76, 98-113 (no VTK output)
inline bool solution_diverged(const Solver& s) { return s.dt() < 1e-300 || s.dt() > 1e300; }
inline void run_solver(Grid&, Solver& s, int64_t max_steps) {
  s.initialize();
  while (!(s.time() >= s.final_time() || s.step() >= max_steps)) {
    s.solve();
    if (solution_diverged(s)) throw Error(HOM3G_ERR_DIVERGED, "Solution diverged!");
  }
}

// coupling.hpp:47-63 applied each step before cns.solve(); heat.solve()
inline void couple_and_step(double Tr0, Solver& cns, Solver& heat, int nsteps = 1) {
  check(hom3g_couple_cht_step(cns.context_handle(), cns.id(), heat.id(), Tr0, nsteps));
}

}  // namespace fv
}  // namespace hom3g
