// A hom3-style driver through the reference-side binding (hom3_binding.hpp):
// the 3D Sod tube of euler_fv_solver_test.cpp:126-156 with the solver built
// from an io::Properties map as the reference test builds it (level 3,
// AUSM, forward Euler, Neumann boundaries).  Prints the state after `steps`
// steps (one value per line, column-major).
#include <cstdio>
#include <cstdlib>

#include "hom3_binding.hpp"

using namespace hom3_binding;

int main(int argc, char** argv) {
  const int steps = argc > 1 ? std::atoi(argv[1]) : 4;
  try {
    hom3g::Context ctx(0);
    static const int nd = 3;
    hom3g::Grid grid(ctx, hom3g::RootCell{{0, 0, 0}, {1, 1, 1}}, 3);
    Properties p;
    insert<hom3g::Grid*>(p, "grid", &grid);
    insert<long long>(p, "maxNoCells", 0);
    insert<double>(p, "timeEnd", 0.2);
    insert<double>(p, "gamma", 1.4);
    insert<double>(p, "Minfty", 0.1);
    insert<double>(p, "CFL", 0.2);
    insert<double>(p, "Rspecific", 287.058);
    insert<double>(p, "T0", 273.15);
    insert<double>(p, "rho0", 1.225);
    euler::Solver<nd, flux::ausm, time_integration::euler_forward> s(0, p);
    s.set_initial_condition(HOM3G_IC_SHOCK_TUBE, {0, 0, 0.3, 1.0, 0.75, 1.0, 0.125, 0.0, 0.1});
    hom3g::fv::append_bcs(s, grid.root_cell(), hom3g::fv::make_conditions(nd, hom3g::fv::Condition::neumann()));
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
