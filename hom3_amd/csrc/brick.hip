// Brick kernels (uniform Morton-ordered grids) — placeholder until the
// LDS-tiled implementation lands.
#include "internal.hpp"

namespace hom3g {

bool brick_supported(const SolverDev&) { return false; }

void brick_stage(Ctx&, SolverDev&, int, const void*, void*, bool) {
  throw Error(HOM3G_ERR_STATE, "brick kernel not available");
}

}  // namespace hom3g
