// euler_fv_solver_test.cpp:258-286 (TEST(euler_fv_solver, isentropic_vortex_ic))
// written against the C++ facade (include/hom3g/fv_solver.hpp): the 2D
// isentropic vortex on the root cell [-8, 8]^2 at minRefLevel 5 with a lambda
// initial condition, the IsentropicVortex condition on every face, and
// run_solver(grid, solver, maxNoTimeSteps, outputInterval).
//
//   vortex_driver <maxNoTimeSteps> <outputInterval> [host]
//
// "host" replaces the built-in IsentropicVortex condition by the same
// condition written here and evaluated on the host (HOM3G_BC_HOST_CALLBACK).
// Prints the final state (one value per line, column-major) for the parity
// test; VTK files go to the working directory when outputInterval > 0.
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "hom3g/fv_solver.hpp"

using namespace hom3g;

int main(int argc, char** argv) {
  const int64_t maxNoTimeSteps = argc > 1 ? std::atoll(argv[1]) : 10;
  const int64_t outputInterval = argc > 2 ? std::atoll(argv[2]) : 0;
  const bool host_bc = argc > 3 && std::strcmp(argv[3], "host") == 0;
  try {
    Context ctx(0);
    /// Create grid
    static const int nd = 2;
    const auto rootCell_2d = RootCell{{-8.0, -8.0}, {8.0, 8.0}};
    const int minRefLevel = 5;
    auto test_grid_2d = Grid{ctx, rootCell_2d, minRefLevel};

    /// Create solver (euler_properties<nd>(&grid, 2.0))
    fv::Properties p;
    p.timeEnd = 2.0;
    p.gamma = 1.4;
    p.Minfty = 0.1;
    p.CFL = 0.2;
    fv::Solver eulerSolver{test_grid_2d, fv::physics::euler, p};

    /// Define initial condition
    auto isentropic_vortex_ic = [](const fv::NumA& x) { return fv::euler::ic::isentropic_vortex(x, 0); };
    eulerSolver.set_initial_condition(isentropic_vortex_ic);

    /// Create boundary conditions
    auto nBc = fv::Condition::isentropic_vortex();
    if (host_bc) {
      // euler/boundary_conditions.hpp:65-89 on the host: ghost and boundary
      // cell take the analytic vortex at their centres at the solver's time
      nBc = fv::Condition::on_host([](fv::GhostSet& s) {
        for (int64_t i = 0; i < s.n; ++i) {
          const fv::NumA qg = fv::euler::ic::isentropic_vortex({s.xg(i, 0), s.xg(i, 1)}, s.time);
          const fv::NumA qb = fv::euler::ic::isentropic_vortex({s.xb(i, 0), s.xb(i, 1)}, s.time);
          for (int v = 0; v < s.nvars; ++v) {
            s.Qg(i, v) = qg[v];
            s.Qb(i, v) = qb[v];
          }
        }
      });
    }
    fv::append_bcs(eulerSolver, test_grid_2d.root_cell(), fv::make_conditions(nd, nBc));

    fv::run_solver(test_grid_2d, eulerSolver, maxNoTimeSteps, outputInterval);

    const auto q = eulerSolver.lhs();
    const int64_t n = eulerSolver.ncells(), nint = eulerSolver.first_ghost();
    std::printf("%lld %lld %d %.17g %.17g %lld\n", (long long)n, (long long)nint, eulerSolver.nvars(),
                eulerSolver.time(), eulerSolver.dt(), (long long)eulerSolver.step());
    for (int v = 0; v < eulerSolver.nvars(); ++v)
      for (int64_t c = 0; c < nint; ++c) std::printf("%.17g\n", q[(size_t)v * n + c]);
  } catch (const Error& e) {
    std::fprintf(stderr, "hom3g error %d: %s\n", e.code, e.what());
    return e.code == HOM3G_ERR_DIVERGED ? 3 : 1;
  }
  return 0;
}
