// Host check of brick_face_rev (hom3_amd/csrc/face_rev.hpp), the reverse of
// the tiled brick's CNS face flux: for random states (Mach 0.05 to 3, both
// upwind directions, sub- and supersonic), random tangential difference
// sums and a random flux adjoint Fb, every input component of the reverse
// equals <Fb, dF/dx_i> by central differences of the forward face
// (brick_prims + ausm_flux_fast + viscous_flux_fast, the functions the
// brick kernel calls) evaluated in long double.  Prints the worst relative
// difference; exit status 0 when it is below 1e-7.
#include <cstdio>
#include <random>
#include <cmath>
#include "physics.hpp"
namespace hom3g {
template <> inline long double dev_rcp<long double>(long double x) { return 1.0L / x; }
template <> inline long double dev_sqrt<long double>(long double x) { return sqrtl(x); }
template <> inline long double fast_rsq<long double>(long double x) { return 1.0L / sqrtl(x); }
}
#include "face_rev.hpp"
using namespace hom3g;

template <int ND, int D, class T>
void fwd(const PhysConst& pc, const T* qL, const T* qR, const T (*Gt)[ND - 1], T inv_h, T* f) {
  constexpr int NV = ND + 2, RHO = ND;
  T pL, iaL, pR, iaR;
  brick_prims<ND, T>(qL, T(pc.gamma), T(pc.gm1), pL, iaL);
  brick_prims<ND, T>(qR, T(pc.gamma), T(pc.gm1), pR, iaR);
  const T aL = T(pc.gamma) * pL * iaL, aR = T(pc.gamma) * pR * iaR;
  ausm_flux_fast<ND, D, T>(qL, qL[D], pL, aL, iaL, qR, qR[D], pR, aR, iaR, f);
  T Gn[NV], Gf[NV][ND - 1];
  for (int v = 0; v < NV; ++v) { Gn[v] = qR[v] - qL[v]; for (int t = 0; t < ND - 1; ++t) Gf[v][t] = T(1e30); }
  for (int t = 0; t < ND - 1; ++t) {
    const int sd = t < D ? t : t + 1;
    Gf[D][t] = Gt[0][t]; Gf[sd][t] = Gt[1][t]; Gf[RHO][t] = Gt[2][t];
  }
  T fv[NV];
  viscous_flux_fast<ND, D, T>(qL, pL, qR, pR, Gn, Gf, pc, T(pc.q_coef), T(pc.inv_Re0), inv_h, fv);
  for (int v = 0; v < NV; ++v) f[v] += fv[v];
}

template <int ND, int D>
double check(const PhysConst& pc, std::mt19937_64& rng, double mach, int trial, bool verbose) {
  constexpr int NV = ND + 2, RHO = ND, E = ND + 1, NT = ND - 1;
  std::uniform_real_distribution<double> U(-1, 1);
  double qL[NV], qR[NV], Gt[3][NT], Fb[NV];
  auto state = [&](double* q) {
    const double rho = 1.0 + 0.4 * U(rng), p = (1.0 + 0.3 * U(rng)) / pc.gamma;
    double u[3], um2 = 0;
    for (int k = 0; k < ND; ++k) { u[k] = mach * U(rng); um2 += u[k] * u[k]; }
    for (int k = 0; k < ND; ++k) q[k] = rho * u[k];
    q[RHO] = rho;
    q[E] = p / pc.gm1 + 0.5 * rho * um2;
  };
  state(qL); state(qR);
  for (int j = 0; j < 3; ++j) for (int t = 0; t < NT; ++t) Gt[j][t] = 0.3 * U(rng);
  for (int v = 0; v < NV; ++v) Fb[v] = U(rng);
  const double inv_h = 64.0;
  double pL, iaL, pR, iaR;
  brick_prims<ND, double>(qL, pc.gamma, pc.gm1, pL, iaL);
  brick_prims<ND, double>(qR, pc.gamma, pc.gm1, pR, iaR);
  double qLb[NV], qRb[NV], Gtb[3][NT];
  brick_face_rev<ND, D>(pc, qL, qR, pL, pR, iaL, iaR, Gt, inv_h, Fb, qLb, qRb, Gtb);
  // every input direction: <Fb, J e_i> by central differences vs the reverse's component
  double worst = 0;
  const int nin = 2 * NV + 3 * NT;
  for (int i = 0; i < nin; ++i) {
    long double xl[NV], xr[NV], g[3][NT], fp[NV], fm[NV];
    const long double eps = 1e-7L;
    for (int s = 0; s < 2; ++s) {
      for (int v = 0; v < NV; ++v) { xl[v] = qL[v]; xr[v] = qR[v]; }
      for (int j = 0; j < 3; ++j) for (int t = 0; t < NT; ++t) g[j][t] = Gt[j][t];
      const long double h = s == 0 ? eps : -eps;
      if (i < NV) xl[i] += h * (fabsl(xl[i]) + 1);
      else if (i < 2 * NV) xr[i - NV] += h * (fabsl(xr[i - NV]) + 1);
      else g[(i - 2 * NV) / NT][(i - 2 * NV) % NT] += h;
      long double* f = s == 0 ? fp : fm;
      for (int v = 0; v < NV; ++v) f[v] = 0;
      fwd<ND, D, long double>(pc, xl, xr, g, (long double)inv_h, f);
    }
    long double scale = 1;
    if (i < NV) scale = fabsl((long double)qL[i]) + 1;
    else if (i < 2 * NV) scale = fabsl((long double)qR[i - NV]) + 1;
    long double fd = 0;
    for (int v = 0; v < NV; ++v) fd += Fb[v] * (fp[v] - fm[v]) / (2 * eps * scale);
    const double rv = i < NV ? qLb[i] : (i < 2 * NV ? qRb[i - NV] : Gtb[(i - 2 * NV) / NT][(i - 2 * NV) % NT]);
    const double err = fabs((double)fd - rv) / (fabs((double)fd) + 1e-3);
    if (verbose || err > 1e-8) printf("  ND %d D %d trial %d input %2d: fd % .12e rev % .12e err %.2e\n", ND, D, trial, i, (double)fd, rv, err);
    if (err > worst) worst = err;
  }
  return worst;
}

int main() {
  PhysConst pc{};
  pc.gamma = 1.4; pc.gm1 = 0.4; pc.Re0 = 1000; pc.Pr0 = 0.71;
  pc.sOverTref = 110.4 / 273.15; pc.sOverTrefP1 = 1 + pc.sOverTref;
  pc.q_coef = pc.gamma / (pc.gm1 * pc.Pr0); pc.inv_Re0 = 1 / pc.Re0;
  std::mt19937_64 rng(1);
  double worst = 0;
  const double machs[] = {0.05, 0.3, 0.9, 1.6, 3.0};
  for (int trial = 0; trial < 200; ++trial) {
    const double m = machs[trial % 5];
    worst = std::max(worst, check<3, 0>(pc, rng, m, trial, false));
    worst = std::max(worst, check<3, 1>(pc, rng, m, trial, false));
    worst = std::max(worst, check<3, 2>(pc, rng, m, trial, false));
    worst = std::max(worst, check<2, 0>(pc, rng, m, trial, false));
    worst = std::max(worst, check<2, 1>(pc, rng, m, trial, false));
  }
  printf("worst relative error %.3e\n", worst);
  return worst < 1e-7 ? 0 : 1;
}
