// A hom3-style driver written against the C++ facade (include/hom3g/fv_solver.hpp):
// the Sod shock tube of euler_fv_solver_test.cpp:126-156 on a 3D level-3 grid,
// forward Euler, Neumann boundaries.  Prints the state after `steps` steps as
// text (one value per line, column-major) for the parity test.
#include <cstdio>
#include <cstdlib>

#include "hom3g/fv_solver.hpp"

int main(int argc, char** argv) {
  const int steps = argc > 1 ? std::atoi(argv[1]) : 4;
  try {
    hom3g::Context ctx(0);
    hom3g::Grid grid(ctx, 3, {0, 0, 0}, {1, 1, 1}, 3);
    hom3g::fv::Properties p;
    p.CFL = 0.2;
    p.timeEnd = 0.2;
    hom3g::fv::Solver s(grid, hom3g::fv::physics::euler, p);
    s.set_initial_condition(HOM3G_IC_SHOCK_TUBE, {0, 0, 0.3, 1.0, 0.75, 1.0, 0.125, 0.0, 0.1});
    hom3g::fv::append_bcs(s, {hom3g::fv::Condition::neumann()}, grid);
    s.initialize();
    for (int i = 0; i < steps; ++i) s.solve();
    const auto q = s.lhs();
    const int64_t n = s.ncells(), nint = s.first_ghost();
    std::printf("%lld %lld %d %.17g %.17g\n", (long long)n, (long long)nint, s.nvars(), s.time(), s.dt());
    for (int v = 0; v < s.nvars(); ++v)
      for (int64_t c = 0; c < nint; ++c) std::printf("%.17g\n", q[(size_t)v * n + c]);
  } catch (const hom3g::Error& e) {
    std::fprintf(stderr, "hom3g error %d: %s\n", e.code, e.what());
    return 1;
  }
  return 0;
}
