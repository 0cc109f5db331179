// The reference-side binding a hom3 maintainer would add (a new header next
// to src/solver/fv/solver.hpp), compiled here against the C++ facade: a
// solver type constructed like the reference's Solver(SolverIdx,
// io::Properties) (solver.hpp:88-95) whose properties are read into
// hom3g::fv::Properties.  In hom3, io::Properties is
// std::unordered_map<String, boost::any> (misc/properties.hpp:17) and the
// tags are those of solver/fv/tags.hpp and euler/tags.hpp; this test build
// uses std::any and local tag types of the same names.
#pragma once
#include <any>
#include <string>
#include <unordered_map>

#include "hom3g/fv_solver.hpp"

namespace hom3_binding {

using Properties = std::unordered_map<std::string, std::any>;

// misc/properties.hpp:33-61 insert / read
template <class T> void insert(Properties& p, const std::string& name, T value) { p[name] = value; }
template <class T> T read(const Properties& p, const std::string& name) {
  const auto it = p.find(name);
  if (it == p.end()) throw hom3g::Error(HOM3G_ERR_ARG, "missing property: " + name);
  return std::any_cast<T>(it->second);
}
template <class T> T read_or(const Properties& p, const std::string& name, T dflt) {
  return p.count(name) ? read<T>(p, name) : dflt;
}

// tags (solver/fv/tags.hpp:20-28, euler/tags.hpp:13-16)
namespace time_integration { struct euler_forward {}; struct runge_kutta_2 {}; struct ssp_rk3 {}; }
namespace flux { struct ausm {}; struct lax_friedrichs {}; struct three_point {}; }

template <class TI> constexpr int ti_of() {
  if constexpr (std::is_same_v<TI, time_integration::euler_forward>) return HOM3G_TI_EULER_FWD;
  else if constexpr (std::is_same_v<TI, time_integration::runge_kutta_2>) return HOM3G_TI_HEUN_RK2;
  else return HOM3G_TI_SSP_RK3;
}
template <class F> constexpr int flux_of() {
  if constexpr (std::is_same_v<F, flux::lax_friedrichs>) return HOM3G_FLUX_LLF;
  else if constexpr (std::is_same_v<F, flux::three_point>) return HOM3G_FLUX_THREE_POINT;
  else return HOM3G_FLUX_AUSM;
}

// the property "grid" holds the grid the solver lives on (a Grid* in hom3)
inline hom3g::Grid& grid_of(const Properties& p) { return *read<hom3g::Grid*>(p, "grid"); }

// the io::Properties keys the hot path reads (solver.hpp:88-95,
// euler/physics.hpp:39-42, euler/quantities.hpp:23-29, cns/physics.hpp:57-61,
// cns/quantities.hpp:23-30, heat/physics.hpp:36-39, heat/quantities.hpp:14-17)
template <class TI, class F> hom3g::fv::Properties props_of(const Properties& p) {
  hom3g::fv::Properties q;
  q.timeEnd = read<double>(p, "timeEnd");
  q.CFL = read<double>(p, "CFL");
  q.maxNoCells = read_or<long long>(p, "maxNoCells", 0);
  q.gamma = read_or<double>(p, "gamma", q.gamma);
  q.Minfty = read_or<double>(p, "Minfty", q.Minfty);
  q.Rspecific = read_or<double>(p, "Rspecific", q.Rspecific);
  q.T0 = read_or<double>(p, "T0", q.T0);
  q.rho0 = read_or<double>(p, "rho0", q.rho0);
  q.CFL_viscous = read_or<double>(p, "CFL_viscous", q.CFL_viscous);
  q.ReInfty = read_or<double>(p, "ReInfty", q.ReInfty);
  q.Pr0 = read_or<double>(p, "Pr0", q.Pr0);
  q.Sutherland0 = read_or<double>(p, "Sutherland0", q.Sutherland0);
  q.diffusivity = read_or<double>(p, "diffusivity", q.diffusivity);
  q.T_ref = read_or<double>(p, "T_ref", q.T_ref);
  q.time_integration = ti_of<TI>();
  q.flux = flux_of<F>();
  return q;
}

// Solver<PhysicsTT, TimeIntegration> (solver.hpp:60-61) on the GPU
template <hom3g::fv::physics Ph, class TI, class F = flux::ausm>
struct Solver : hom3g::fv::Solver {
  Solver(int solver_idx, const Properties& p) : hom3g::fv::Solver(grid_of(p), Ph, props_of<TI, F>(p), solver_idx) {}
};

// euler::Solver<nd, NumFlux, TI> and friends (HOM3_FV_PHYSICS_SOLVER_, solver.hpp:996-1003)
namespace euler { template <int nd, class F, class TI> using Solver = hom3_binding::Solver<hom3g::fv::physics::euler, TI, F>; }
namespace cns { template <int nd, class F, class TI> using Solver = hom3_binding::Solver<hom3g::fv::physics::cns, TI, F>; }
namespace heat { template <int nd, class F, class TI> using Solver = hom3_binding::Solver<hom3g::fv::physics::heat, TI, F>; }

}  // namespace hom3_binding
