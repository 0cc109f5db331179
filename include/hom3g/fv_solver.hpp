// C++ drop-in facade for hom3's finite-volume solver over the hom3g C ABI.
//
// Mirrors the member surface a hom3 driver uses (src/solver/fv/solver.hpp,
// src/solver/fv/utilities.hpp, src/solver/fv/coupling.hpp,
// src/grid/helpers.hpp) without Eigen or Boost: a hom3 test such as
// euler_fv_solver_test.cpp reads the same with hom3g::fv in place of
// hom3::solver::fv.  Eigen's NumA<n> becomes std::vector<double>; the
// reference's TERMINATE becomes a thrown hom3g::Error.  Header-only; link
// libhom3g.so.
//
//   hom3g::Context ctx(0);
//   hom3g::Grid grid(ctx, hom3g::RootCell{{-8, -8}, {8, 8}}, /*minRefLevel=*/5);
//   hom3g::fv::Properties p;  p.timeEnd = 2.0;
//   hom3g::fv::Solver s(grid, hom3g::fv::physics::euler, p);        // Solver(SolverIdx, io::Properties)
//   s.set_initial_condition([](const hom3g::fv::NumA& x) {          // solver.hpp:153-156
//     return hom3g::fv::euler::ic::isentropic_vortex(x, 0); });
//   hom3g::fv::append_bcs(s, grid.root_cell(),                      // utilities.hpp:15-21
//                         hom3g::fv::make_conditions(2, hom3g::fv::Condition::isentropic_vortex()));
//   hom3g::fv::run_solver(grid, s, maxNoTimeSteps, outputInterval); // utilities.hpp:98-113
#pragma once
#include <cmath>
#include <cstdio>
#include <functional>
#include <limits>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "../hom3g.h"
#include "vtk.hpp"

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

// grid::RootCell<nd> (grid/grid.hpp:106-139): the cube enclosing [x_min, x_max]
struct RootCell {
  std::vector<double> x_min, x_max;
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
  // grid::Grid<nd>{grid::helpers::cube::properties<nd>(rootCell, minRefLevel)}
  Grid(Context& ctx, const RootCell& rc, int min_level, int max_solvers = 1)
      : Grid(ctx, (int)rc.x_min.size(), rc.x_min, rc.x_max, min_level, max_solvers) {}
  int id() const { return id_; }
  int nd() const { return nd_; }
  Context& context() const { return ctx_; }
  double cell_length_at_level(int level) const { return root_length_ / std::pow(2, level); }
  double root_length() const { return root_length_; }
  const double* root_center() const { return root_center_; }
  RootCell root_cell() const { return RootCell{xmin_, xmax_}; }
  void initialize() {}  // the tree is generated at construction
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

using NumA = std::vector<double>;
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

// Nondimensional quantities (euler/quantities.hpp:131-189, cns/quantities.hpp:23-115)
struct Quantities {
  double gamma_ = 1.4, Minfty = 0.1, ReInfty = 100.0, Pr0_ = 0.72, sOverTref = 0, sOverTrefP1 = 1;
  explicit Quantities(const Properties& p)
      : gamma_(p.gamma), Minfty(p.Minfty), ReInfty(p.ReInfty), Pr0_(p.Pr0),
        sOverTref(p.T0 != 0 ? p.Sutherland0 / p.T0 : 0.0), sOverTrefP1(sOverTref + 1.0) {}
  double gamma() const { return gamma_; }
  double gammaM1() const { return gamma_ - 1.0; }
  double T_infinity() const { return 1. / (1 + 0.5 * gammaM1() * (Minfty * Minfty)); }
  double rho_infinity() const { return std::pow(T_infinity(), 1. / gammaM1()); }
  double u_infinity() const { return Minfty * std::sqrt(T_infinity()); }
  double p_infinity() const { return std::pow(T_infinity(), gamma_ / gammaM1()) / gamma_; }
  double sutherlands_law(double T) const { return std::pow(T, 1.5) * sOverTrefP1 / (T + sOverTref); }
  double mu_infinity() const { return sutherlands_law(T_infinity()); }
  double Re0() const { return ReInfty * mu_infinity() / (rho_infinity() * u_infinity()); }
  double Pr0() const { return Pr0_; }
};

namespace euler {
// conserved from primitive variables (u.., rho, p) (euler/quantities.hpp:153-166)
inline NumA cv(const NumA& pv, double gamma = 1.4) {
  const int nd = (int)pv.size() - 2;
  NumA c(pv.size());
  double vm2 = 0;
  for (int d = 0; d < nd; ++d) {
    c[d] = pv[nd] * pv[d];
    vm2 += std::pow(pv[d], 2);
  }
  c[nd] = pv[nd];
  c[nd + 1] = pv[nd + 1] / (gamma - 1.0) + 0.5 * pv[nd] * vm2;
  return c;
}
// primitive from conserved variables (euler/quantities.hpp:173-186)
inline NumA pv(const NumA& cvs, double gamma = 1.4) {
  const int nd = (int)cvs.size() - 2;
  NumA p(cvs.size());
  double vm2 = 0;
  for (int d = 0; d < nd; ++d) {
    p[d] = cvs[d] / cvs[nd];
    vm2 += std::pow(p[d], 2);
  }
  p[nd] = cvs[nd];
  p[nd + 1] = (gamma - 1.0) * (cvs[nd + 1] - 0.5 * cvs[nd] * vm2);
  return p;
}
namespace ic {
// The reference's isentropic vortex, 2D (euler/initial_conditions.hpp:14-46):
// beta 5, gamma 1.4, centre (-5, -5), free stream rho = p = u = v = 1,
// advected with (1, 1); conserved variables at x and time t.
inline NumA isentropic_vortex(const NumA& x, double t) {
  const double pi = 3.141592653589793238462643383279502884;
  const double beta = 5, gamma = 1.4, x0[2] = {-5.0, -5.0}, u = 1, v = 1;
  const double x_t = x[0] - u * t, y_t = x[1] - v * t;
  const double r = std::sqrt(std::pow(x_t - x0[0], 2) + std::pow(y_t - x0[1], 2));
  const double eT1mr2 = std::exp(1 - std::pow(r, 2));
  const double eT21mr2 = std::exp(2 * (1 - std::pow(r, 2)));
  NumA p(4);
  p[0] = u - beta * eT1mr2 * (y_t - x0[1]) / (2 * pi);
  p[1] = v + beta * eT1mr2 * (x_t - x0[0]) / (2 * pi);
  p[2] = std::pow((1 - ((gamma - 1) * std::pow(beta, 2) * eT21mr2) / (16 * gamma * std::pow(pi, 2))),
                  1 / (gamma - 1));
  p[3] = std::pow(p[2], gamma);
  return cv(p, 1.4);
}
}  // namespace ic
}  // namespace euler

class Solver;

// What a host-evaluated Condition sees: see hom3g_bc_callback.
struct GhostSet {
  int buffer;        // 0 lhs (lhs_tag), 1 rhs (rhs_tag), 2 SSP-RK3 stage buffer
  double time;
  int64_t n;
  const int64_t* ghosts;
  const int64_t* bndry;
  const double* x_ghost;
  const double* x_bndry;
  double* q_ghost;
  double* q_bndry;
  int nd, nvars;
  double xg(int64_t i, int d) const { return x_ghost[d * n + i]; }
  double xb(int64_t i, int d) const { return x_bndry[d * n + i]; }
  double& Qg(int64_t i, int v) { return q_ghost[v * n + i]; }
  double& Qb(int64_t i, int v) { return q_bndry[v * n + i]; }
  // cell_dx(bndry, ghost) (solver.hpp:590-592)
  double dist(int64_t i) const {
    double s = 0;
    for (int d = 0; d < nd; ++d) { const double t = xb(i, d) - xg(i, d); s += t * t; }
    return std::sqrt(s);
  }
};

// A boundary Condition: one of the reference's condition classes with
// constant surface distributions, a coupling condition, or any other
// Condition evaluated on the host with one of the two slope hooks.
struct Condition {
  int kind;
  std::vector<double> params;
  const Solver* other = nullptr;
  double kratio = 1.0;
  std::function<void(GhostSet&)> host;
  int hook = HOM3G_SLOPE_DEFAULT;
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
  static Condition on_host(std::function<void(GhostSet&)> f, int slope_hook = HOM3G_SLOPE_DEFAULT) {
    Condition c{HOM3G_BC_HOST_CALLBACK, {}};
    c.host = std::move(f);
    c.hook = slope_hook;
    return c;
  }
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
  using InitialCondition = std::function<NumA(const NumA&)>;
  // solver.hpp:298-313
  struct BoundaryInfo {
    int64_t bndryIdx;
    int bndryPos;
    int64_t ghostIdx;
    int ghostPos;
  };

  Solver(Grid& grid, physics ph, const Properties& p, int solver_idx = 0)
      : quantities(p), grid_(grid), props_(p), physics_(ph), solver_idx_(solver_idx) {
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
    nd_ = grid.nd();
    nvars_ = (ph == physics::heat) ? 1 : grid.nd() + 2;
    check(hom3g_solver_create(ctx(), grid.id(), solver_idx, &d, &id_));
  }
  Solver(const Solver&) = delete;
  Solver& operator=(const Solver&) = delete;

  // -- input (solver.hpp:150-169)
  void set_initial_condition(int ic_kind, std::vector<double> params) {
    ic_kind_ = ic_kind;
    ic_params_ = std::move(params);
    ic_fn_ = nullptr;
  }
  // the "initialCondition" property: cell averages at the Gauss points,
  // evaluated on the host by the library (hom3g_solver_impose_ic_host)
  void set_initial_condition(InitialCondition f) { ic_fn_ = std::move(f); }
  void set_initial_domain(const hom3g_geometry& g, int sign) { check(hom3g_solver_set_domain(ctx(), id_, &g, sign)); }
  void append_bc(const hom3g_geometry& g, const Condition& c) {
    if (c.kind == HOM3G_BC_HOST_CALLBACK) {
      auto hb = std::make_unique<HostBC>(HostBC{c.host, nd_, nvars_});
      check(hom3g_solver_append_bc_host(ctx(), id_, &g, &Solver::host_bc_trampoline, hb.get(), c.hook));
      host_bcs_.push_back(std::move(hb));
      return;
    }
    check(hom3g_solver_append_bc(ctx(), id_, &g, c.kind, c.params.data(), (int)c.params.size(),
                                 c.other ? c.other->id_ : -1, c.kratio));
  }
  // solver.hpp:165-169
  void append_bcs(const std::vector<std::pair<hom3g_geometry, Condition>>& bcs) {
    for (const auto& b : bcs) append_bc(b.first, b.second);
  }

  // -- solver interface (solver.hpp:100-145)
  void initialize() {
    check(hom3g_solver_build(ctx(), id_));
    if (ic_fn_)
      check(hom3g_solver_impose_ic_host(ctx(), id_, &Solver::ic_trampoline, this));
    else
      check(hom3g_solver_impose_ic(ctx(), id_, ic_kind_, ic_params_.data(), (int)ic_params_.size()));
    load_topology();
    dirty();
  }
  void solve() { solve(1); }
  void solve(int nsteps) {
    dirty();
    check(hom3g_solver_solve(ctx(), id_, nsteps));
  }
  void apply_bcs() {
    dirty();
    check(hom3g_solver_apply_bcs(ctx(), id_));
  }
  int64_t step() const { double t, dt; int64_t s; check(hom3g_solver_time(ctx(), id_, &t, &dt, &s)); return s; }
  double time() const { double t, dt; int64_t s; check(hom3g_solver_time(ctx(), id_, &t, &dt, &s)); return t; }
  double dt() const { double t, dt; int64_t s; check(hom3g_solver_time(ctx(), id_, &t, &dt, &s)); return dt; }
  void dt(double forced) { check(hom3g_solver_force_dt(ctx(), id_, forced)); }
  double min_dt() const { double d; check(hom3g_solver_min_dt(ctx(), id_, &d)); return d; }
  double final_time() const { return props_.timeEnd; }
  int id() const { return id_; }
  int nd() const { return nd_; }
  int nvars() const { return nvars_; }
  int solver_idx() const { return solver_idx_; }
  hom3g_ctx* context_handle() const { return ctx(); }
  const Grid& grid() const { return grid_; }

  // -- output (solver.hpp:316-376)
  std::string domain_name() const {
    const char* ph = physics_ == physics::euler ? "Euler" : (physics_ == physics::cns ? "CNS" : "Heat");
    return std::string("Fv_") + ph + std::to_string(nd_) + "D_Id" + std::to_string(solver_idx_);
  }

  // -- data: cells() in the reference's column-major layout
  int64_t ncells() const { int64_t n, f; check(hom3g_solver_counts(ctx(), id_, &n, &f)); return n; }
  int64_t first_ghost() const { int64_t n, f; check(hom3g_solver_counts(ctx(), id_, &n, &f)); return f; }
  std::vector<double> lhs() const {
    const int64_t n = ncells();
    std::vector<double> q((size_t)nvars_ * n);
    check(hom3g_solver_download_state(ctx(), id_, q.data(), n));
    return q;
  }
  void set_lhs(const std::vector<double>& q, int64_t ld) {
    dirty();
    check(hom3g_solver_upload_state(ctx(), id_, q.data(), ld));
  }
  // Q<lhs>(c, v) (solver.hpp:188-220), read from a host copy of the state
  // taken at the first access after the state last changed
  double Q(int64_t c, int v) const {
    refresh();
    return cache_[(size_t)v * cache_ld_ + c];
  }
  NumA Q(int64_t c) const {
    NumA q(nvars_);
    for (int v = 0; v < nvars_; ++v) q[v] = Q(c, v);
    return q;
  }
  // boundary_info (solver.hpp:298-313): the ghost's single neighbour
  BoundaryInfo boundary_info(int64_t g) const {
    const int npos = 2 * nd_;
    for (int p = 0; p < npos; ++p) {
      const int64_t nb = nghbr_[(size_t)p * topo_ld_ + g];
      if (nb != HOM3G_INVALID_ID) return {nb, p, g, p ^ 1};
    }
    return {HOM3G_INVALID_ID, -1, g, -1};
  }
  int64_t neighbor(int64_t c, int pos) const { return nghbr_[(size_t)pos * topo_ld_ + c]; }
  double x_center(int64_t c, int d) const { return xc_[(size_t)d * topo_ld_ + c]; }
  double length(int64_t c) const { return len_[c]; }
  // solver.hpp:590-592: Euclidean distance of the cell centres
  double cell_dx(int64_t a, int64_t b) const {
    double s = 0;
    for (int d = 0; d < nd_; ++d) { const double t = x_center(a, d) - x_center(b, d); s += t * t; }
    return std::sqrt(s);
  }
  // physics accessors (euler/physics.hpp:53-165) on the host copy
  double rho(int64_t c) const { return Q(c, nd_); }
  double rho_u(int64_t c, int d) const { return Q(c, d); }
  double rho_E(int64_t c) const { return Q(c, nd_ + 1); }
  double u(int64_t c, int d) const { return rho_u(c, d) / rho(c); }
  double p(int64_t c) const {
    double s = 0;
    for (int d = 0; d < nd_; ++d) s += u(c, d) * u(c, d);
    return quantities.gammaM1() * (rho_E(c) - 0.5 * rho(c) * s);
  }
  double a(int64_t c) const { return std::sqrt(quantities.gamma() * p(c) / rho(c)); }
  double T(int64_t c) const { return nvars_ == 1 ? Q(c, 0) : quantities.gamma() * p(c) / rho(c); }
  NumA pv(int64_t c) const { return euler::pv(Q(c), quantities.gamma()); }
  NumA cv(const NumA& pvs) const { return euler::cv(pvs, quantities.gamma()); }

  Quantities quantities;

 private:
  static void ic_trampoline(void* user, const double* x, double* q) {
    Solver* s = static_cast<Solver*>(user);
    const NumA r = s->ic_fn_(NumA(x, x + s->nd_));
    for (int v = 0; v < s->nvars_; ++v) q[v] = r[v];
  }
  // one per host-evaluated condition: the library hands it back as `user`
  struct HostBC {
    std::function<void(GhostSet&)> fn;
    int nd, nvars;
  };
  static void host_bc_trampoline(void* user, int buffer, double time, int64_t n, const int64_t* ghosts,
                                 const int64_t* bndry, const double* xg, const double* xb, double* qg, double* qb) {
    HostBC* h = static_cast<HostBC*>(user);
    GhostSet gs{buffer, time, n, ghosts, bndry, xg, xb, qg, qb, h->nd, h->nvars};
    h->fn(gs);
  }
  void load_topology() {
    const int64_t n = ncells();
    topo_ld_ = n;
    nghbr_.assign((size_t)2 * nd_ * n, 0);
    xc_.assign((size_t)nd_ * n, 0);
    len_.assign(n, 0);
    check(hom3g_solver_topology(ctx(), id_, nghbr_.data(), nullptr, nullptr, xc_.data(), len_.data(), nullptr));
  }
  void dirty() { cache_valid_ = false; }
  void refresh() const {
    if (cache_valid_) return;
    cache_ = lhs();
    cache_ld_ = ncells();
    cache_valid_ = true;
  }
  hom3g_ctx* ctx() const { return grid_.context().get(); }
  Grid& grid_;
  Properties props_;
  physics physics_;
  int solver_idx_ = 0, nd_ = 0;
  int id_ = -1, nvars_ = 0, ic_kind_ = HOM3G_IC_CONSTANT_PV;
  std::vector<double> ic_params_;
  InitialCondition ic_fn_;
  std::vector<std::unique_ptr<HostBC>> host_bcs_;
  int64_t topo_ld_ = 0;
  std::vector<int64_t> nghbr_;
  std::vector<double> xc_, len_;
  mutable std::vector<double> cache_;
  mutable int64_t cache_ld_ = 0;
  mutable bool cache_valid_ = false;
};

// grid/helpers.hpp:102-109 make_conditions: one condition per cube face
inline std::vector<Condition> make_conditions(int nd, const Condition& c) {
  return std::vector<Condition>((size_t)2 * nd, c);
}

// grid/helpers.hpp:112-160 make_boundaries (left, right, bottom, top,
// front, back Edges inset by 1e-15 * length)
inline std::vector<hom3g_geometry> cube_boundaries(const RootCell& rc) {
  const int nd = (int)rc.x_min.size();
  double L = std::numeric_limits<double>::epsilon(), centre0 = 0;
  for (int d = 0; d < nd; ++d) L = std::max(L, rc.x_max[d] - rc.x_min[d]);
  centre0 = rc.x_min[0] + 0.5 * (rc.x_max[0] - rc.x_min[0]);
  const double eps = L * 1e-15;
  const double x_min = centre0 - 0.5 * L + eps, x_max = centre0 + 0.5 * L - eps;
  const double off = x_min + 0.5 * L;
  std::vector<hom3g_geometry> out;
  for (int d = 0; d < nd; ++d)
    for (int side = 0; side < 2; ++side) {
      std::vector<double> a(nd, off), n(nd, 0.0);
      a[d] = side ? x_max : x_min;
      n[d] = side ? -1.0 : 1.0;
      out.push_back(edge(a, n));
    }
  return out;
}
inline std::vector<hom3g_geometry> cube_boundaries(const Grid& g) { return cube_boundaries(g.root_cell()); }

// utilities.hpp:15-21 append_bcs(solver, rootCell, conditions)
inline void append_bcs(Solver& s, const RootCell& rc, const std::vector<Condition>& conds) {
  const auto geos = cube_boundaries(rc);
  for (size_t k = 0; k < geos.size(); ++k) s.append_bc(geos[k], conds[conds.size() == 1 ? 0 : k]);
}
inline void append_bcs(Solver& s, const std::vector<Condition>& conds, const Grid& g) {
  append_bcs(s, g.root_cell(), conds);
}

// solver.hpp:328-376 write_domain: apply_bcs(lhs), then the VTK file
// <domain_name>_<step>.vtk
inline std::string write_domain(Solver& s) {
  s.apply_bcs();
  return vtk::write_domain(s.context_handle(), s.id(), s.domain_name() + "_" + std::to_string(s.step()),
                           s.quantities.gamma());
}
inline void write_domain(Grid&) {}  // the grid has no output of its own here

// utilities.hpp:48-113
inline bool solver_finished(int64_t max_steps, const Solver& s) {
  return s.time() >= s.final_time() || s.step() >= max_steps;
}
inline bool solution_diverged(const Solver& s) { return s.dt() < 1e-300 || s.dt() > 1e300; }
inline void write_output(int64_t output_interval, Solver& s) {
  if (output_interval > 0 && s.step() % output_interval == 0) write_domain(s);
}
inline void write_timestep(const Solver& s) {
  std::fprintf(stderr, "Solver: %s | step: %lld | time: %g | dt: %g\n", s.domain_name().c_str(),
               (long long)s.step(), s.time(), s.dt());
}
// run_solver(grid, solver, maxNoTimeSteps, outputInterval): "Solution
// diverged!" is thrown as Error(HOM3G_ERR_DIVERGED) instead of TERMINATE;
// output_interval <= 0 writes no files
inline void run_solver(Grid& grid, Solver& s, int64_t max_steps, int64_t output_interval) {
  grid.initialize();
  s.initialize();
  if (output_interval > 0) write_domain(s);
  while (!solver_finished(max_steps, s)) {
    if (output_interval > 0) write_timestep(s);
    s.solve();  // throws HOM3G_ERR_DIVERGED
    write_output(output_interval, s);
    if (solution_diverged(s)) throw Error(HOM3G_ERR_DIVERGED, "Solution diverged!");
  }
  if (output_interval > 0) write_domain(s);
}
inline void run_solver(Grid& grid, Solver& s, int64_t max_steps) { run_solver(grid, s, max_steps, 0); }

// coupling.hpp:47-63 applied each step before cns.solve(); heat.solve()
inline void couple_and_step(double Tr0, Solver& cns, Solver& heat, int nsteps = 1) {
  check(hom3g_couple_cht_step(cns.context_handle(), cns.id(), heat.id(), Tr0, nsteps));
}

}  // namespace fv
}  // namespace hom3g
