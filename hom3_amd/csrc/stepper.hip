// Table-driven ("generic") kernels of the finite-volume step: ghost fill
// (apply_bcs), CFL time step (compute_dt/set_dt), residual + Runge-Kutta
// stage update (num_flux/evolve), initial condition, reductions, coupling.
//
// These kernels follow the reference's cell-centred formulation exactly:
// num_flux (solver.hpp:439-460) evaluates both faces of every cell from the
// neighbour table, slope() (solver.hpp:596-623) recurses to one-sided
// neighbours and calls the boundary condition's slope hook at ghosts.  They
// accept any topology the reference can build (masks, cut boundaries,
// coupled solvers).  Uniform periodic/zero-gradient grids use the brick
// kernels in brick.hip instead.
#include <algorithm>
#include <cmath>
#include <type_traits>

#include "internal.hpp"

namespace hom3g {

namespace {

inline unsigned nblk(int64_t n, int bs = 256) { return (unsigned)((n + bs - 1) / bs); }

struct GenTopo {
  const int32_t* nghbr;
  long long ld;
  const double* len;
  long long n_int;
  long long ncells;
  const int32_t* g_bndry;
  const int32_t* g_hookbc;
  const int32_t* g_image;
  const int32_t* fine;  // 2^(nd-1) cells per face towards finer cells
  // CNS: every cell's slope(c, v, sd) of the stage input at
  // slopes[(sd nv + v) ld + c], formed once per stage (k_cell_slopes); null:
  // the faces walk the slopes themselves
  const double* slopes;
};

struct HookTable {
  int kind[kMaxBC];
};

__device__ __forceinline__ bool wall_hook(int kind) {
  return kind == HOM3G_BC_ADIABATIC_NOSLIP || kind == HOM3G_BC_ISOTHERMAL_NOSLIP || kind == HOM3G_BC_CHT_FLUID;
}

// slope (solver.hpp:596-623) with the boundary-condition hooks
// (boundary_condition.hpp:63-67, cns/boundary_conditions.hpp:143-147).
// TQ: the state's storage type (FP64, or FP32 stage states); the slope in FP64
template <int ND, class TQ>
__device__ double slope(const GenTopo& t, const int* hk, const TQ* Q, int c, int v, int dir) {
  double sign = 1.0;
  for (int it = 0; it < 16; ++it) {
    const int nn = t.nghbr[(2 * dir) * t.ld + c];
    const int np = t.nghbr[(2 * dir + 1) * t.ld + c];
    if (nn >= 0 && np >= 0) {
      const double s = ((double)Q[v * t.ld + np] - (double)Q[v * t.ld + nn]) / (2. * t.len[c]);
      return sign < 0 ? -s : s;
    }
    if (nn >= 0) { c = nn; continue; }
    if (np >= 0) { c = np; continue; }
    const int g = c - (int)t.n_int;
    if (g < 0 || g >= (int)(t.ncells - t.n_int)) break;  // an internal cell without neighbours: undefined
    const int kind = hk[t.g_hookbc[g]];
    if (kind == HOM3G_BC_PERIODIC) { c = t.g_image[g]; continue; }
    if (wall_hook(kind) && v < ND) sign = -sign;
    c = t.g_bndry[g];
  }
  return __builtin_nan("");
}

template <int ND, class TQ, class TC>
__device__ __forceinline__ void load_q(const TQ* Q, long long ld, int c, TC* q) {
#pragma unroll
  for (int v = 0; v < ND + 2; ++v) q[v] = (TC)Q[v * ld + c];
}

// compute_num_flux for the face (l, r) with normal d (euler/cns physics),
// with the two cells' states already loaded (qL, qR).  TC: the arithmetic
// type (double; float for FP32 stage states)
template <int ND, int PHYS, class TQ, class TC>
__device__ __forceinline__ void face_flux_q(const GenTopo& t, const int* hk, const PhysConst& pc, const TQ* Q,
                                            int l, int r, const TC* qL, const TC* qR, int d, double dx,
                                            double dt, TC* f) {
  constexpr int NV = ND + 2;
  TC aL[NV], aR[NV];
#pragma unroll
  for (int v = 0; v < NV; ++v) { aL[v] = qL[v]; aR[v] = qR[v]; }
  if (pc.recon == HOM3G_RECON_MUSCL) {
    // second neighbours; across a periodic ghost they continue at the image
    int ll = t.nghbr[(2 * d) * t.ld + l];
    int rr = t.nghbr[(2 * d + 1) * t.ld + r];
    if (ll < 0 && l >= t.n_int && hk[t.g_hookbc[l - t.n_int]] == HOM3G_BC_PERIODIC)
      ll = t.nghbr[(2 * d) * t.ld + t.g_image[l - t.n_int]];
    if (rr < 0 && r >= t.n_int && hk[t.g_hookbc[r - t.n_int]] == HOM3G_BC_PERIODIC)
      rr = t.nghbr[(2 * d + 1) * t.ld + t.g_image[r - t.n_int]];
#pragma unroll
    for (int v = 0; v < NV; ++v) {
      const TC sl = ll >= 0 ? limit(pc.limiter, qL[v] - (TC)Q[v * t.ld + ll], qR[v] - qL[v]) : TC(0);
      const TC sr = rr >= 0 ? limit(pc.limiter, qR[v] - qL[v], (TC)Q[v * t.ld + rr] - qR[v]) : TC(0);
      aL[v] = qL[v] + TC(0.5) * sl;
      aR[v] = qR[v] - TC(0.5) * sr;
    }
  }
  // Euler: the reciprocal forms (physics.hpp); CNS keeps the reference form
  const TC gamma = TC(pc.gamma), gm1 = TC(pc.gm1);
  const Prim<ND, TC> wL = (PHYS == HOM3G_EULER) ? primitives_rcp<ND>(aL, gamma, gm1) : primitives<ND>(aL, gamma, gm1);
  const Prim<ND, TC> wR = (PHYS == HOM3G_EULER) ? primitives_rcp<ND>(aR, gamma, gm1) : primitives<ND>(aR, gamma, gm1);
  if (pc.flux == HOM3G_FLUX_LLF) llf_flux<ND>(aL, wL, aR, wR, d, TC(dx / dt), f);
  else if constexpr (PHYS == HOM3G_EULER) ausm_flux_rcp_rt<ND>(aL, wL, aR, wR, d, f);
  else ausm_flux_rt<ND>(aL, wL, aR, wR, d, f);
  if constexpr (PHYS == HOM3G_CNS) {
    // surface gradients of all conserved variables (solver.hpp:656-678)
    TC G[NV][ND];
    const TC inv_len_r = TC(1.0 / t.len[r]);
#pragma unroll
    for (int sd = 0; sd < ND; ++sd) {
      if (sd == d) {
#pragma unroll
        for (int v = 0; v < NV; ++v) G[v][sd] = (qR[v] - qL[v]) * inv_len_r;
      } else {
        if (t.slopes) {
          const double* S = t.slopes + (long long)(sd * NV) * t.ld;
#pragma unroll
          for (int v = 0; v < NV; ++v) G[v][sd] = TC(0.5) * (TC(S[v * t.ld + r]) + TC(S[v * t.ld + l]));
        } else {
#pragma unroll
          for (int v = 0; v < NV; ++v)
            G[v][sd] = TC(0.5) * (TC(slope<ND, TQ>(t, hk, Q, r, v, sd)) + TC(slope<ND, TQ>(t, hk, Q, l, v, sd)));
        }
      }
    }
    TC pL = wL.p, pR = wR.p;
    if (pc.recon == HOM3G_RECON_MUSCL) {
      pL = primitives<ND>(qL, gamma, gm1).p;
      pR = primitives<ND>(qR, gamma, gm1).p;
    }
    TC fv[NV];
    if (d == 0) viscous_flux<ND, 0, TC>(qL, pL, qR, pR, G, pc, fv);
    else if (d == 1) viscous_flux<ND, 1, TC>(qL, pL, qR, pR, G, pc, fv);
    else if constexpr (ND == 3) viscous_flux<ND, 2, TC>(qL, pL, qR, pR, G, pc, fv);
#pragma unroll
    for (int v = 0; v < NV; ++v) f[v] = f[v] + fv[v];
  }
}

template <int ND, int PHYS, class TQ, class TC>
__device__ __forceinline__ void face_flux(const GenTopo& t, const int* hk, const PhysConst& pc, const TQ* Q, int l,
                                          int r, int d, double dx, double dt, TC* f) {
  constexpr int NV = ND + 2;
  TC qL[NV], qR[NV];
  load_q<ND>(Q, t.ld, l, qL);
  load_q<ND>(Q, t.ld, r, qR);
  face_flux_q<ND, PHYS, TQ, TC>(t, hk, pc, Q, l, r, qL, qR, d, dx, dt, f);
}

// one stage's buffers: the input and output states (FP64, or FP32 stage
// states: read and written at their type, the fluxes in FP32) and U^n
template <class TIn, class TOut>
struct StageArgsT {
  const TIn* in;
  const double* base;
  TOut* out;
  const double* src;  // per-cell volumetric source (heat) or null
  int kind;           // 0: (in+nf)+src  1: 0.5*((base+in)+nf)  2: 0.75*base+0.25*((in+nf)+src)
                      // 3: (1/3)*base+(2/3)*((in+nf)+src)  4: in+nf
};
using StageArgs = StageArgsT<double, double>;

// flux through one side of cell c: towards neighbour `o` (a cell, or a
// face towards finer cells when o <= -2)
template <int ND, int PHYS, class TQ, class TC>
__device__ __forceinline__ void side_flux(const GenTopo& t, const int* hk, const PhysConst& pc,
                                          const TQ* in, int o, int c, bool minus, int d, double dx, double dt,
                                          TC* f) {
  constexpr int NV = ND + 2;
  if (o >= 0) {
    if (minus) face_flux<ND, PHYS, TQ, TC>(t, hk, pc, in, o, c, d, dx, dt, f);
    else face_flux<ND, PHYS, TQ, TC>(t, hk, pc, in, c, o, d, dx, dt, f);
    return;
  }
  constexpr int K = 1 << (ND - 1);
  const int32_t* fc = t.fine + (long long)(-(o + 2)) * K;
  TC sub[K][NV];
#pragma unroll
  for (int k = 0; k < K; ++k) {
    if (minus) face_flux<ND, PHYS, TQ, TC>(t, hk, pc, in, fc[k], c, d, dx, dt, sub[k]);
    else face_flux<ND, PHYS, TQ, TC>(t, hk, pc, in, c, fc[k], d, dx, dt, sub[k]);
  }
#pragma unroll
  for (int v = 0; v < NV; ++v) {
    const TC sm = (ND == 2) ? sub[0][v] + sub[1][v] : (sub[0][v] + sub[1][v]) + (sub[2 % K][v] + sub[3 % K][v]);
    f[v] = sm * TC((ND == 2) ? 0.5 : 0.25);
  }
}

// one direction of a first-order or MUSCL Euler cell update from preloaded
// states: the same faces and order as the generic loop below
template <int ND, int D, class TQ, class TC>
__device__ __forceinline__ void euler_direction(const GenTopo& t, const int* hk, const PhysConst& pc, const TQ* in,
                                                int c, const int* nbr, const TC* qc,
                                                const TC (*qn)[ND + 2], double dx, double dt, TC* acc) {
  constexpr int NV = ND + 2;
  const int m = nbr[2 * D], p = nbr[2 * D + 1];
  TC fm[NV], fp[NV];
  if (m <= -2 || p <= -2) {
    // 2:1 refinement extension: a face towards finer cells is the mean of
    // its sub-face fluxes, summed pairwise and scaled by 2^(1-nd)
    side_flux<ND, HOM3G_EULER, TQ, TC>(t, hk, pc, in, m, c, true, D, dx, dt, fm);
    side_flux<ND, HOM3G_EULER, TQ, TC>(t, hk, pc, in, p, c, false, D, dx, dt, fp);
  } else if (m < 0 || p < 0) {
    // a face without a neighbour or ghost is undefined in the reference
    // (solver.hpp:448-453); the cell is poisoned instead of reading out of bounds
#pragma unroll
    for (int v = 0; v < NV; ++v) acc[v] = TC(__builtin_nan(""));
    return;
  } else {
    face_flux_q<ND, HOM3G_EULER, TQ, TC>(t, hk, pc, in, m, c, qn[2 * D], qc, D, dx, dt, fm);
    face_flux_q<ND, HOM3G_EULER, TQ, TC>(t, hk, pc, in, c, p, qc, qn[2 * D + 1], D, dx, dt, fp);
  }
#pragma unroll
  for (int v = 0; v < NV; ++v) acc[v] += (fm[v] - fp[v]);
}

// Workgroup ids are dispatched round-robin over the 8 XCDs, each with its
// own L2: map them so that XCD x takes the x-th contiguous eighth of the
// blocks (of the Morton-ordered cells), and a cell's neighbours are mostly
// in its XCD's L2.  A bijection on [0, gridDim.x).
__device__ __forceinline__ long long xcd_block() {
  const long long g = gridDim.x, i = blockIdx.x;
  const long long q = g / 8, r = g % 8, x = i % 8;
  return x * q + (x < r ? x : r) + i / 8;
}

template <int ND, int PHYS, class TIn, class TOut>
__global__ void __launch_bounds__(256) k_stage_generic(GenTopo t, HookTable H, PhysConst pc, StageArgsT<TIn, TOut> a,
                                                       const DevScalars* scal) {
  static_assert(PHYS != HOM3G_HEAT || (sizeof(TIn) == 8 && sizeof(TOut) == 8),
                "FP32 heat stages run k_heat_stage32");
  // the boundary-hook kinds in LDS: indexed per lane (a dynamically indexed
  // kernel argument would be copied to scratch memory by every thread)
  __shared__ int hk[kMaxBC];
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < kMaxBC; ++k) hk[k] = H.kind[k];
  }
  __syncthreads();
  // XCD-aware order for the CNS stage (its faces gather the neighbours'
  // slopes as well as their states: 6% on the channel); the Euler stage on
  // a refined 5 M-leaf grid ran 15% slower with it
  const long long blk = (PHYS == HOM3G_CNS) ? xcd_block() : (long long)blockIdx.x;
  const long long c = blk * blockDim.x + threadIdx.x;
  if (c >= t.n_int) return;
  constexpr int NV = (PHYS == HOM3G_HEAT) ? 1 : ND + 2;
  // the flux arithmetic: FP64, or FP32 with FP32 stage states (the RK
  // combination below stays FP64)
  using TC = std::conditional_t<sizeof(TIn) == 4 || sizeof(TOut) == 4, float, double>;
  const double dt = scal->dt;
  const double dx = t.len[c];
  TC acc[NV];
#pragma unroll
  for (int v = 0; v < NV; ++v) acc[v] = TC(0);
  if constexpr (PHYS == HOM3G_EULER) {
    // the neighbour ids, the cell's and its same-level neighbours' states
    // are all loaded before the first flux (a face towards finer cells or a
    // missing neighbour loads the cell again and takes its own path)
    int nbr[2 * ND];
    TC qc[NV], qn[2 * ND][NV];
#pragma unroll
    for (int k = 0; k < 2 * ND; ++k) nbr[k] = t.nghbr[k * t.ld + c];
    load_q<ND>(a.in, t.ld, (int)c, qc);
#pragma unroll
    for (int k = 0; k < 2 * ND; ++k) load_q<ND>(a.in, t.ld, nbr[k] >= 0 ? nbr[k] : (int)c, qn[k]);
    euler_direction<ND, 0, TIn, TC>(t, hk, pc, a.in, (int)c, nbr, qc, qn, dx, dt, acc);
    euler_direction<ND, 1, TIn, TC>(t, hk, pc, a.in, (int)c, nbr, qc, qn, dx, dt, acc);
    if constexpr (ND == 3) euler_direction<ND, 2, TIn, TC>(t, hk, pc, a.in, (int)c, nbr, qc, qn, dx, dt, acc);
  }
  // CNS and heat: one direction at a time
  for (int d = 0; d < ND && PHYS != HOM3G_EULER; ++d) {
    const int m = t.nghbr[(2 * d) * t.ld + c];
    const int p = t.nghbr[(2 * d + 1) * t.ld + c];
    if (m < 0 || p < 0) {
      // a face without a neighbour or ghost is undefined in the reference
      // (solver.hpp:448-453); the cell is poisoned instead of reading out of bounds
#pragma unroll
      for (int v = 0; v < NV; ++v) acc[v] = TC(__builtin_nan(""));
      continue;
    }
    if constexpr (PHYS == HOM3G_HEAT) {
      // heat/physics.hpp:113-119: (1/dx) * (T_l - T_r), dx of the centre cell
      const double Tc = a.in[c];
      const double fm = 1 / dx * (a.in[m] - Tc);
      const double fp = 1 / dx * (Tc - a.in[p]);
      acc[0] += (fm - fp);
    } else {
      TC fm[NV], fp[NV];
      face_flux<ND, PHYS, TIn, TC>(t, hk, pc, a.in, m, (int)c, d, dx, dt, fm);
      face_flux<ND, PHYS, TIn, TC>(t, hk, pc, a.in, (int)c, p, d, dx, dt, fp);
#pragma unroll
      for (int v = 0; v < NV; ++v) acc[v] += (fm[v] - fp[v]);
    }
  }
  const double s = (a.src != nullptr) ? dt * a.src[c] : 0.0;
#pragma unroll
  for (int v = 0; v < NV; ++v) {
    const double nf = dt / dx * acc[v];
    const double src = (v == 0 && PHYS == HOM3G_HEAT) ? s : 0.0;
    const double q = (double)a.in[v * t.ld + c];
    double o;
    switch (a.kind) {
      case 0: o = q + nf + src; break;
      case 1: o = 0.5 * (a.base[v * t.ld + c] + q + nf); break;
      case 2: o = 0.75 * a.base[v * t.ld + c] + 0.25 * (q + nf + src); break;
      case 3: o = (1. / 3.) * a.base[v * t.ld + c] + (2. / 3.) * (q + nf + src); break;
      default: o = q + nf; break;
    }
    a.out[v * t.ld + c] = (TOut)o;
  }
}

// slope(c, v, sd) of every cell (internal and ghost) for the table-driven
// CNS stage: the faces read each cell's slopes instead of walking them for
// both cells of every face (the same values as slope<ND>)
template <int ND, class TQ>
__global__ void __launch_bounds__(256) k_cell_slopes(GenTopo t, HookTable H, const TQ* Q, double* S) {
  __shared__ int hk[kMaxBC];
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < kMaxBC; ++k) hk[k] = H.kind[k];
  }
  __syncthreads();
  const long long c0 = xcd_block() * blockDim.x + threadIdx.x;
  if (c0 >= t.ncells) return;
  constexpr int NV = ND + 2;
#pragma unroll
  for (int sd = 0; sd < ND; ++sd) {
    // slope's walk (solver.hpp:596-623) once for all variables: it does not
    // depend on the variable but for the wall hooks' sign, which applies to
    // the momenta only
    int c = (int)c0;
    bool flip = false, ok = false;
    int nn = -1, np = -1;
    for (int it = 0; it < 16; ++it) {
      nn = t.nghbr[(2 * sd) * t.ld + c];
      np = t.nghbr[(2 * sd + 1) * t.ld + c];
      if (nn >= 0 && np >= 0) { ok = true; break; }
      if (nn >= 0) { c = nn; continue; }
      if (np >= 0) { c = np; continue; }
      const int g = c - (int)t.n_int;
      if (g < 0 || g >= (int)(t.ncells - t.n_int)) break;  // an internal cell without neighbours: undefined
      const int kind = hk[t.g_hookbc[g]];
      if (kind == HOM3G_BC_PERIODIC) { c = t.g_image[g]; continue; }
      if (wall_hook(kind)) flip = !flip;
      c = t.g_bndry[g];
    }
    const double den = ok ? 2. * t.len[c] : 1.0;
#pragma unroll
    for (int v = 0; v < NV; ++v) {
      double x = __builtin_nan("");
      if (ok) {
        const double sv = ((double)Q[v * t.ld + np] - (double)Q[v * t.ld + nn]) / den;
        x = (flip && v < ND) ? -sv : sv;
      }
      S[(long long)(sd * NV + v) * t.ld + c0] = x;
    }
  }
}

// ---- boundary conditions ---------------------------------------------------
struct GhostArgs {
  void* buf;  // the state buffer (double, or float for FP32 stage states)
  const double* own_lhs;
  long long ld, n_int, ng;
  const int32_t* g_bndry;
  const int32_t* g_pos;
  const int32_t* g_apply;
  const int32_t* g_image;
  const int32_t* g_partner;
  const double* g_dist;
  const double* xc;
  int nd, nvars, physics;
  double gamma;
};

// euler/quantities.hpp:173-186 pv (left-to-right velocity accumulation)
template <int ND>
__device__ void pv_of(const double* cv, double gamma, double* pv) {
  pv[ND] = cv[ND];
  double vm2 = 0;
  for (int d = 0; d < ND; ++d) { pv[d] = cv[d] / cv[ND]; vm2 += pv[d] * pv[d]; }
  pv[ND + 1] = (gamma - 1.0) * (cv[ND + 1] - 0.5 * cv[ND] * vm2);
}
template <int ND>
__device__ void cv_of(const double* pv, double gamma, double* cv) {
  cv[ND] = pv[ND];
  double vm2 = 0;
  for (int d = 0; d < ND; ++d) { cv[d] = pv[ND] * pv[d]; vm2 += pv[d] * pv[d]; }
  cv[ND + 1] = pv[ND + 1] / (gamma - 1.0) + 0.5 * pv[ND] * vm2;
}

// euler/initial_conditions.hpp:19-45 (2D isentropic vortex, beta 5, gamma 1.4)
__device__ void isentropic_vortex(const double* x, double t, double cx, double cy, double* out) {
  const double beta = 5, gamma = 1.4, u = 1, v = 1, pi = 3.141592653589793238462643383279502884;
  const double x_t = x[0] - u * t, y_t = x[1] - v * t;
  const double dx = x_t - cx, dy = y_t - cy;
  const double r2 = dx * dx + dy * dy;
  const double r = sqrt(r2);
  const double rr = r * r;
  const double e1 = exp(1 - rr);
  const double e2 = exp(2 * (1 - rr));
  double pv[4];
  pv[0] = u - beta * e1 * (y_t - cy) / (2 * pi);
  pv[1] = v + beta * e1 * (x_t - cx) / (2 * pi);
  pv[2] = pow((1 - ((gamma - 1) * (beta * beta) * e2) / (16 * gamma * (pi * pi))), 1 / (gamma - 1));
  pv[3] = pow(pv[2], gamma);
  cv_of<2>(pv, 1.4, out);
}

__device__ double fluid_T(const double* Q, long long ld, int c, int nd, double gamma) {
  const double rho = Q[nd * ld + c];
  double u2[3];
  for (int d = 0; d < nd; ++d) u2[d] = Q[d * ld + c] / rho;
  double s = (nd == 2) ? u2[0] * u2[0] + u2[1] * u2[1] : u2[0] * u2[0] + (u2[1] * u2[1] + u2[2] * u2[2]);
  const double p = (gamma - 1.) * (Q[(nd + 1) * ld + c] - 0.5 * rho * s);
  return gamma * p / rho;
}

// T: the state buffer's type; ghost formulas are evaluated in FP64 (an
// FP32 stage state is read, converted, and the ghost rounded once)
template <int ND, class T>
__global__ void k_fill_ghosts(GhostArgs A, BCTable B, const DevScalars* scal) {
  const long long g = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (g >= A.ng) return;
  const int bc = A.g_apply[g];
  if (bc < 0) return;
  const long long ld = A.ld;
  const long long gc = A.n_int + g;
  const int b = A.g_bndry[g];
  const int kind = B.kind[bc];
  const double* pp = B.params[bc];
  T* Q = static_cast<T*>(A.buf);
  const int nv = A.nvars;
  switch (kind) {
    case HOM3G_BC_NEUMANN: {
      const double dx = A.g_dist[g];
      for (int v = 0; v < nv; ++v) {
        const double gv = (B.nparams[bc] > v) ? pp[v] : 0.0;
        Q[v * ld + gc] = Q[v * ld + b] - dx * gv;
      }
      break;
    }
    case HOM3G_BC_DIRICHLET:
      for (int v = 0; v < nv; ++v) Q[v * ld + gc] = 2.0 * pp[v] - Q[v * ld + b];
      break;
    case HOM3G_BC_ISENTROPIC_VORTEX: {
      if constexpr (ND == 2) {
        const double cx = B.nparams[bc] >= 2 ? pp[0] : -5.0;
        const double cy = B.nparams[bc] >= 2 ? pp[1] : -5.0;
        double xg[2] = {A.xc[gc], A.xc[ld + gc]};
        double xb[2] = {A.xc[b], A.xc[ld + b]};
        double qg[4], qb[4];
        isentropic_vortex(xg, scal->time, cx, cy, qg);
        isentropic_vortex(xb, scal->time, cx, cy, qb);
        for (int v = 0; v < 4; ++v) { Q[v * ld + gc] = qg[v]; Q[v * ld + b] = qb[v]; }
      }
      break;
    }
    case HOM3G_BC_INFLOW: case HOM3G_BC_OUTFLOW: case HOM3G_BC_ADIABATIC_NOSLIP: case HOM3G_BC_ADIABATIC_SLIP:
    case HOM3G_BC_ISOTHERMAL_NOSLIP: case HOM3G_BC_CHT_FLUID: {
      constexpr int RHO = ND, P = ND + 1;
      double cvb[ND + 2], pvb[ND + 2], pvg[ND + 2], cvg[ND + 2];
      for (int v = 0; v < ND + 2; ++v) cvb[v] = Q[v * ld + b];
      pv_of<ND>(cvb, A.gamma, pvb);
      if (kind == HOM3G_BC_INFLOW) {
        for (int d = 0; d < ND; ++d) pvg[d] = 2.0 * pp[d] - pvb[d];
        pvg[RHO] = 2.0 * pp[RHO] - pvb[RHO];
        pvg[P] = pvb[P] - 0.0;
      } else if (kind == HOM3G_BC_OUTFLOW) {
        for (int d = 0; d < ND; ++d) pvg[d] = pvb[d] - 0.0;
        pvg[RHO] = pvb[RHO] - 0.0;
        pvg[P] = 2.0 * pp[P] - pvb[P];
      } else if (kind == HOM3G_BC_ADIABATIC_NOSLIP) {
        for (int d = 0; d < ND; ++d) pvg[d] = 0.0 - pvb[d];
        pvg[RHO] = pvb[RHO] - 0.0;
        pvg[P] = pvb[P] - 0.0;
      } else if (kind == HOM3G_BC_ADIABATIC_SLIP) {
        const int normal = (A.g_pos[g] ^ 1) / 2;
        for (int d = 0; d < ND; ++d) pvg[d] = pvb[d] - 0.0;
        pvg[normal] = 0.0 - pvb[normal];
        pvg[RHO] = pvb[RHO] - 0.0;
        pvg[P] = pvb[P] - 0.0;
      } else {
        double Ts;
        if (kind == HOM3G_BC_CHT_FLUID) {
          const double Tf = fluid_T(A.own_lhs, ld, b, ND, A.gamma);
          const double Tsol = B.other_lhs[bc][A.g_partner[g]];
          Ts = (B.kratio[bc] == 1.0) ? 0.5 * (Tf + Tsol) : (Tf + B.kratio[bc] * Tsol) / (1.0 + B.kratio[bc]);
        } else {
          Ts = pp[0];
        }
        for (int d = 0; d < ND; ++d) pvg[d] = 0.0 - pvb[d];
        const double rho_s = pvb[P] * A.gamma / Ts;
        pvg[RHO] = 2.0 * rho_s - pvb[RHO];
        pvg[P] = pvb[P] - 0.0;
      }
      cv_of<ND>(pvg, A.gamma, cvg);
      for (int v = 0; v < ND + 2; ++v) Q[v * ld + gc] = cvg[v];
      break;
    }
    case HOM3G_BC_HEAT_DIRICHLET:
      Q[gc] = 2.0 * pp[0] - Q[b];
      break;
    case HOM3G_BC_CHT_SOLID: {
      const double Tsol = A.own_lhs[b];
      const double Tf = fluid_T(B.other_lhs[bc], B.other_ld[bc], A.g_partner[g], ND, B.other_gamma[bc]);
      const double Ts =
          (B.kratio[bc] == 1.0) ? 0.5 * (Tf + Tsol) : (Tf + B.kratio[bc] * Tsol) / (1.0 + B.kratio[bc]);
      Q[gc] = 2.0 * Ts - Q[b];
      break;
    }
    case HOM3G_BC_HEAT_NEUMANN:
      Q[gc] = Q[b] - A.g_dist[g] * pp[0];
      break;
    case HOM3G_BC_PERIODIC: {
      const int im = A.g_image[g];
      for (int v = 0; v < nv; ++v) Q[v * ld + gc] = Q[v * ld + im];
      break;
    }
    case HOM3G_BC_EULER_COUPLING: {
      // euler/coupling_conditions.hpp:24-31: the other solver's state at the
      // partner cell, from its buffer of the same role
      const double* O = B.other_lhs[bc];
      const long long old = B.other_ld[bc];
      const int pc_ = A.g_partner[g];
      if (pc_ < 0 || O == nullptr) break;
      for (int v = 0; v < nv; ++v) Q[v * ld + gc] = O[v * old + pc_];
      break;
    }
    default:
      break;
  }
}

// ---- time step ---------------------------------------------------------------
__device__ __forceinline__ double wave_min(double v) {
  for (int o = 32; o > 0; o >>= 1) {
    const double w = __shfl_xor(v, o, 64);
    v = ref_min(v, w);
  }
  return v;
}

template <int ND, int PHYS, bool FAST>
__global__ void __launch_bounds__(256) k_dt(const double* Q, long long ld, const double* len, long long n,
                                            PhysConst pc, DevScalars* scal) {
  __shared__ double red[4];
  const long long c = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  double dt = 1.7976931348623157e308;
  if (c < n) {
    const double h = len[c];
    if constexpr (PHYS == HOM3G_HEAT) {
      dt = ref_min(dt, pc.cfl * 0.5 * (h * h));
    } else {
      double q[ND + 2];
      for (int v = 0; v < ND + 2; ++v) q[v] = Q[v * ld + c];
      double cd;
      if (FAST) {
        // the formula the brick kernels fuse into their last stage, so a
        // brick solver's dt does not depend on where the minimum came from
        cd = cell_dt_fast<ND, PHYS == HOM3G_CNS>(q, h, pc);
      } else {
        const Prim<ND, double> w = primitives<ND>(q, pc.gamma, pc.gm1);
        cd = (PHYS == HOM3G_CNS) ? cell_dt_cns<ND, double>(w, h, pc) : cell_dt_euler<ND, double>(w, h, pc.cfl);
      }
      dt = ref_min(dt, cd);
    }
  }
  dt = wave_min(dt);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) red[wid] = dt;
  __syncthreads();
  if (threadIdx.x == 0) {
    double m = red[0];
    for (int k = 1; k < (int)(blockDim.x >> 6); ++k) m = ref_min(m, red[k]);
    atomic_lower(&scal->dtmin, m);
  }
}

// set_dt (solver.hpp:541-551), on the device; resets the running minimum
__global__ void k_set_dt(DevScalars* s, double h) {
  const double dtm = ordered_to_dbl(s->dtmin);
  if (s->time + s->dt < s->tEnd) {
    s->dt = (s->step == s->forced_step) ? s->forced_dt : dtm;
    // solution_diverged (utilities.hpp:73-76), checked by the solve driver;
    // the end-time clamp below (run_solver stops there) is not divergence
    if ((s->dt < 1e-300 || s->dt > 1e300) && s->diverged_step < 0) s->diverged_step = s->step;
  } else {
    s->dt = s->tEnd - s->time;
  }
  s->dtmin = dbl_to_ordered(1.7976931348623157e308);
  // dt / h for the brick stages, once per step instead of once per thread
  s->dtdx = h > 0.0 ? s->dt / h : 0.0;
}

// advance (solver.hpp:502-505)
__global__ void k_advance(DevScalars* s) {
  s->time += s->dt;
  s->step += 1;
}

__global__ void k_reset_min(DevScalars* s) { s->dtmin = dbl_to_ordered(1.7976931348623157e308); }

// coupling::apply(Tr0, cns, heat) (coupling.hpp:47-63): both minima must be
// in dtmin of the two solvers
__global__ void k_couple_cht(DevScalars* cns, DevScalars* heat, double scaling) {
  const double mc = ordered_to_dbl(cns->dtmin);
  const double mh = ordered_to_dbl(heat->dtmin) / scaling;
  const double dt = ref_min(mc, mh);
  cns->forced_step = cns->step;
  cns->forced_dt = dt;
  heat->forced_step = heat->step;
  heat->forced_dt = dt * scaling;
  cns->dtmin = dbl_to_ordered(1.7976931348623157e308);
  heat->dtmin = dbl_to_ordered(1.7976931348623157e308);
}

// coupling::apply(s1, s2) (coupling.hpp:66-73)
__global__ void k_couple_same(DevScalars* a, DevScalars* b) {
  const double dt = ref_min(ordered_to_dbl(a->dtmin), ordered_to_dbl(b->dtmin));
  a->forced_step = a->step;
  a->forced_dt = dt;
  b->forced_step = b->step;
  b->forced_dt = dt;
  a->dtmin = dbl_to_ordered(1.7976931348623157e308);
  b->dtmin = dbl_to_ordered(1.7976931348623157e308);
}

__global__ void k_copy_ghosts(const double* from, double* to, long long ld, long long n_int, long long ng, int nv) {
  const long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (t >= ng * nv) return;
  const long long v = t / ng, g = t % ng;
  to[v * ld + n_int + g] = from[v * ld + n_int + g];
}

// ---- initial condition (solver.hpp:980-992, quadrature.hpp:55-74) -------------
struct ICArgs {
  int kind, nd, nvars, physics;
  double gamma;
  double p[48];
  int np;
};

template <int ND>
__device__ void ic_point(const ICArgs& A, const double* x, double* out) {
  const double* P = A.p;
  constexpr int NV = ND + 2;
  const double pi = 3.141592653589793238462643383279502884;
  switch (A.kind) {
    case HOM3G_IC_CONSTANT_PV: cv_of<ND>(P, A.gamma, out); return;
    case HOM3G_IC_SHOCK_TUBE: {
      const int dir = (int)P[0];
      const double angle = P[1], x0 = P[2];
      const double alpha = angle * (2 * pi) / 360;
      double xr[3] = {0, 0, 0};
      xr[0] = cos(alpha) * x[0] - sin(alpha) * x[1];
      xr[1] = sin(alpha) * x[0] + cos(alpha) * x[1];
      const double pos = (alpha == 0) ? x0 : x0 * sqrt(2.);
      double pv[NV];
      const bool left = xr[dir] < pos;
      const double um = left ? P[4] : P[7];
      for (int d = 0; d < ND; ++d) pv[d] = 0;
      pv[0] = um * cos(alpha);
      pv[1] = um * sin(alpha);
      pv[ND] = left ? P[3] : P[6];
      pv[ND + 1] = left ? P[5] : P[8];
      cv_of<ND>(pv, 1.4, out);
      return;
    }
    case HOM3G_IC_ISENTROPIC_VORTEX: {
      if constexpr (ND == 2) {
        const double cx = A.np > 2 ? P[1] : -5.0, cy = A.np > 2 ? P[2] : -5.0;
        isentropic_vortex(x, P[0], cx, cy, out);
      }
      return;
    }
    case HOM3G_IC_SPHERE: {
      double s = 0;
      double t[3];
      for (int d = 0; d < ND; ++d) { const double u = x[d] - P[d]; t[d] = u * u; }
      s = (ND == 2) ? t[0] + t[1] : t[0] + (t[1] + t[2]);
      const double r = sqrt(s);
      cv_of<ND>((r < P[ND]) ? P + ND + 1 : P + ND + 1 + NV, A.gamma, out);
      return;
    }
    case HOM3G_IC_TGV: {
      if constexpr (ND == 3) {
        const double V0 = P[0], rinf = P[1], pinf = P[2];
        double pv[5];
        pv[0] = V0 * sin(x[0]) * cos(x[1]) * cos(x[2]);
        pv[1] = -V0 * cos(x[0]) * sin(x[1]) * cos(x[2]);
        pv[2] = 0.0;
        pv[3] = rinf;
        pv[4] = pinf + rinf * V0 * V0 / 16.0 * (cos(2 * x[0]) + cos(2 * x[1])) * (cos(2 * x[2]) + 2.0);
        cv_of<3>(pv, A.gamma, out);
      }
      return;
    }
    case HOM3G_IC_FOUR_QUADRANT: {
      const int q = (x[0] >= P[0] ? 1 : 0) + (x[1] >= P[1] ? 2 : 0);
      cv_of<ND>(P + 2 + q * NV, A.gamma, out);
      return;
    }
    case HOM3G_IC_LINEAR_PV: {
      double pv[NV];
      for (int v = 0; v < NV; ++v) {
        double a = P[v];
        for (int d = 0; d < ND; ++d) a += P[NV + v * ND + d] * x[d];
        pv[v] = a;
      }
      cv_of<ND>(pv, A.gamma, out);
      return;
    }
    default: return;
  }
}

template <int ND>
__global__ void k_impose_ic(ICArgs A, const double* xc, const double* len, long long ld, long long n, double* Q) {
  const long long c = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (c >= n) return;
  const double l = len[c];
  const double dx2 = 0.5 * l;
  const double gp = 1 / sqrt(3.0);
  const int nv = A.nvars;
  double res[5] = {0, 0, 0, 0, 0};
  for (int k = 0; k < (1 << ND); ++k) {
    double xg[3];
    for (int d = 0; d < ND; ++d) xg[d] = xc[d * ld + c] + dx2 * (((k >> d) & 1) ? gp : -gp);
    double val[5] = {0, 0, 0, 0, 0};
    if (A.physics == HOM3G_HEAT) {
      if (A.kind == HOM3G_IC_HEAT_CONSTANT) val[0] = A.p[0];
      else if (A.kind == HOM3G_IC_HEAT_SINE) {
        double v = A.p[1];
        for (int d = 0; d < ND; ++d) v *= sin(A.p[2] * (xg[d] - A.p[3]));
        val[0] = A.p[0] + v;
      } else if (A.kind == HOM3G_IC_LINEAR_PV) {
        double a = A.p[0];
        for (int d = 0; d < ND; ++d) a += A.p[1 + d] * xg[d];
        val[0] = a;
      }
    } else {
      ic_point<ND>(A, xg, val);
    }
    for (int v = 0; v < nv; ++v) res[v] = res[v] + val[v];
  }
  const double V = (ND == 2) ? l * l : l * (l * l);
  for (int v = 0; v < nv; ++v) Q[v * ld + c] = res[v] * V / (double)(1 << ND) / V;
}

// ---- reductions ----------------------------------------------------------------
__global__ void __launch_bounds__(256) k_cons_partial(const double* Q, long long ld, const double* len,
                                                      long long n, int nd, int nv, double* partial) {
  __shared__ double sh[256];
  for (int v = 0; v < nv; ++v) {
    double s = 0;
    for (long long c = blockIdx.x * 256ll + threadIdx.x; c < n; c += (long long)gridDim.x * 256) {
      const double h = len[c];
      const double V = (nd == 2) ? h * h : h * (h * h);
      s += Q[v * ld + c] * V;
    }
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
      if ((int)threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
      __syncthreads();
    }
    if (threadIdx.x == 0) partial[v * gridDim.x + blockIdx.x] = sh[0];
    __syncthreads();
  }
}

__global__ void k_cons_final(const double* partial, int nblocks, int nv, double* out) {
  const int v = threadIdx.x;
  if (v >= nv) return;
  double s = 0;
  for (int b = 0; b < nblocks; ++b) s += partial[v * nblocks + b];
  out[v] = s;
}

template <int ND>
__global__ void k_check(const double* Q, long long ld, long long n, int physics, double gamma,
                        unsigned long long* bad) {
  const long long c = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (c >= n) return;
  bool ok;
  if (physics == HOM3G_HEAT) {
    const double T = Q[c];
    ok = isfinite(T);
  } else {
    double q[ND + 2];
    for (int v = 0; v < ND + 2; ++v) q[v] = Q[v * ld + c];
    const Prim<ND, double> w = primitives<ND>(q, gamma, gamma - 1.0);
    ok = isfinite(w.rho) && isfinite(w.p) && w.rho > 0 && w.p > 0;
    for (int v = 0; v < ND + 2; ++v) ok = ok && isfinite(q[v]);
  }
  if (!ok) atomicAdd(bad, 1ull);
}

GenTopo make_topo(SolverDev& S) {
  GenTopo t;
  t.nghbr = S.nghbr.p;
  t.ld = S.cap;
  t.len = S.length.p;
  t.n_int = S.n_int;
  t.ncells = S.ncells;
  t.g_bndry = S.g_bndry.p;
  t.g_hookbc = S.g_hookbc.p;
  t.g_image = S.g_image.p;
  t.fine = S.fine.p;
  t.slopes = nullptr;
  return t;
}

HookTable make_hooks(SolverDev& S) {
  HookTable H;
  for (int k = 0; k < kMaxBC; ++k) H.kind[k] = 0;
  for (size_t b = 0; b < S.bcs.size(); ++b) {
    H.kind[b] = S.bcs[b].kind;
    // a host condition's slope hook in the form of the kind that has it
    if (S.bcs[b].kind == HOM3G_BC_HOST_CALLBACK)
      H.kind[b] = S.bcs[b].hook == HOM3G_SLOPE_WALL ? HOM3G_BC_ADIABATIC_NOSLIP : HOM3G_BC_NEUMANN;
  }
  return H;
}

template <int ND, int PHYS, class TIn, class TOut>
void launch_stage_t(Ctx& C, SolverDev& S, const StageArgsT<TIn, TOut>& a) {
  GenTopo t = make_topo(S);
  C.prof_mark();
  if (PHYS == HOM3G_CNS && S.slopes.p) {
    hipLaunchKernelGGL((k_cell_slopes<ND, TIn>), dim3(nblk(S.ncells)), dim3(256), 0, C.stream, t, make_hooks(S),
                       a.in, S.slopes.p);
    H3_LAUNCH_CHECK();
    t.slopes = S.slopes.p;
  }
  hipLaunchKernelGGL((k_stage_generic<ND, PHYS, TIn, TOut>), dim3(nblk(S.n_int)), dim3(256), 0, C.stream, t,
                     make_hooks(S), S.pc, a, S.scal.p);
  H3_LAUNCH_CHECK();
  C.prof_mark();
}

// Heat stage with FP32 stage states (FP32 flux arithmetic, FP64 RK
// combination; the same statements as k_stage_generic's heat branch):
// in/out are float (stage buffers) or double (lhs)
template <int ND, class TIn, class TOut>
__global__ void __launch_bounds__(256) k_heat_stage32(GenTopo t, const void* inv, const double* base, void* outv,
                                                      const double* src, int kind, const DevScalars* scal) {
  const long long c = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (c >= t.n_int) return;
  const TIn* in = static_cast<const TIn*>(inv);
  const double dt = scal->dt;
  const double dx = t.len[c];
  const float idx = (float)(1 / dx);
  float acc = 0.0f;
  for (int d = 0; d < ND; ++d) {
    const int m = t.nghbr[(2 * d) * t.ld + c];
    const int p = t.nghbr[(2 * d + 1) * t.ld + c];
    if (m < 0 || p < 0) {
      acc = __builtin_nanf("");
      continue;
    }
    const float Tc = (float)in[c];
    const float fm = idx * ((float)in[m] - Tc);
    const float fp = idx * (Tc - (float)in[p]);
    acc += (fm - fp);
  }
  const double sv = (src != nullptr) ? dt * src[c] : 0.0;
  const double nf = dt / dx * (double)acc;
  const double q = (double)in[c];
  double o;
  switch (kind) {
    case 0: o = q + nf + sv; break;
    case 1: o = 0.5 * (base[c] + q + nf); break;
    case 2: o = 0.75 * base[c] + 0.25 * (q + nf + sv); break;
    case 3: o = (1. / 3.) * base[c] + (2. / 3.) * (q + nf + sv); break;
    default: o = q + nf; break;
  }
  static_cast<TOut*>(outv)[c] = (TOut)o;
}

void launch_heat_stage32(Ctx& C, SolverDev& S, const void* in, bool in32, void* out, bool out32, int kind,
                         const double* src) {
  auto go = [&](auto kern) {
    C.prof_mark();
    hipLaunchKernelGGL(kern, dim3(nblk(S.n_int)), dim3(256), 0, C.stream, make_topo(S), in,
                       (const double*)S.lhs.p, out, src, kind, (const DevScalars*)S.scal.p);
    H3_LAUNCH_CHECK();
    C.prof_mark();
  };
  if (S.nd == 2) {
    if (in32 && out32) go(k_heat_stage32<2, float, float>);
    else if (in32) go(k_heat_stage32<2, float, double>);
    else if (out32) go(k_heat_stage32<2, double, float>);
    else go(k_heat_stage32<2, double, double>);
  } else {
    if (in32 && out32) go(k_heat_stage32<3, float, float>);
    else if (in32) go(k_heat_stage32<3, float, double>);
    else if (out32) go(k_heat_stage32<3, double, float>);
    else go(k_heat_stage32<3, double, double>);
  }
}

void launch_stage(Ctx& C, SolverDev& S, const StageArgs& a) {
  if (S.nd == 2) {
    if (S.P.physics == HOM3G_EULER) launch_stage_t<2, HOM3G_EULER>(C, S, a);
    else if (S.P.physics == HOM3G_CNS) launch_stage_t<2, HOM3G_CNS>(C, S, a);
    else launch_stage_t<2, HOM3G_HEAT>(C, S, a);
  } else {
    if (S.P.physics == HOM3G_EULER) launch_stage_t<3, HOM3G_EULER>(C, S, a);
    else if (S.P.physics == HOM3G_CNS) launch_stage_t<3, HOM3G_CNS>(C, S, a);
    else launch_stage_t<3, HOM3G_HEAT>(C, S, a);
  }
}

// Euler / CNS stage on the table-driven kernel with FP32 stage states: in
// and out are float (the stage buffers) or double (lhs)
template <class TIn, class TOut>
void launch_stage_fp32_t(Ctx& C, SolverDev& S, const StageArgsT<TIn, TOut>& a) {
  if (S.nd == 2) {
    if (S.P.physics == HOM3G_EULER) launch_stage_t<2, HOM3G_EULER>(C, S, a);
    else launch_stage_t<2, HOM3G_CNS>(C, S, a);
  } else {
    if (S.P.physics == HOM3G_EULER) launch_stage_t<3, HOM3G_EULER>(C, S, a);
    else launch_stage_t<3, HOM3G_CNS>(C, S, a);
  }
}
void launch_stage_fp32(Ctx& C, SolverDev& S, const void* in, bool in32, void* out, bool out32, int kind,
                       const double* src) {
  H3_REQUIRE(S.P.physics == HOM3G_EULER || S.P.physics == HOM3G_CNS, "Euler or CNS stage");
  const double* base = S.lhs.p;
  if (in32 && out32)
    launch_stage_fp32_t(C, S, StageArgsT<float, float>{(const float*)in, base, (float*)out, src, kind});
  else if (in32)
    launch_stage_fp32_t(C, S, StageArgsT<float, double>{(const float*)in, base, (double*)out, src, kind});
  else if (out32)
    launch_stage_fp32_t(C, S, StageArgsT<double, float>{(const double*)in, base, (float*)out, src, kind});
  else
    launch_stage_fp32_t(C, S, StageArgsT<double, double>{(const double*)in, base, (double*)out, src, kind});
}

void launch_dt(Ctx& C, SolverDev& S) {
  const unsigned nb = nblk(S.n_int);
  const bool fast = S.kernel == HOM3G_KERNEL_BRICK;
  auto go = [&](auto kern) {
    hipLaunchKernelGGL(kern, dim3(nb), dim3(256), 0, C.stream, (const double*)S.lhs.p, (long long)S.cap,
                       (const double*)S.length.p, (long long)S.n_int, S.pc, S.scal.p);
    H3_LAUNCH_CHECK();
  };
  if (S.nd == 2) {
    if (S.P.physics == HOM3G_EULER) fast ? go(k_dt<2, HOM3G_EULER, true>) : go(k_dt<2, HOM3G_EULER, false>);
    else if (S.P.physics == HOM3G_CNS) fast ? go(k_dt<2, HOM3G_CNS, true>) : go(k_dt<2, HOM3G_CNS, false>);
    else go(k_dt<2, HOM3G_HEAT, false>);
  } else {
    if (S.P.physics == HOM3G_EULER) fast ? go(k_dt<3, HOM3G_EULER, true>) : go(k_dt<3, HOM3G_EULER, false>);
    else if (S.P.physics == HOM3G_CNS) fast ? go(k_dt<3, HOM3G_CNS, true>) : go(k_dt<3, HOM3G_CNS, false>);
    else go(k_dt<3, HOM3G_HEAT, false>);
  }
}

__global__ void k_copy(const double4* __restrict__ src, double4* __restrict__ dst, long long n4) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n4; i += (long long)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

// copy variants of the bandwidth sweep (32-byte elements)
typedef double dvec4 __attribute__((ext_vector_type(4)));
// 1: grid-stride, non-temporal loads and stores
__global__ void k_copy_nt(const dvec4* __restrict__ src, dvec4* __restrict__ dst, long long n4) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n4; i += (long long)gridDim.x * blockDim.x)
    __builtin_nontemporal_store(__builtin_nontemporal_load(&src[i]), &dst[i]);
}
// 2: grid-stride, four loads in flight per thread, non-temporal stores
__global__ void k_copy_x4(const dvec4* __restrict__ src, dvec4* __restrict__ dst, long long n4) {
  const long long S = (long long)gridDim.x * blockDim.x;
  long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  for (; i + 3 * S < n4; i += 4 * S) {
    const dvec4 a = src[i], b = src[i + S], c = src[i + 2 * S], d = src[i + 3 * S];
    __builtin_nontemporal_store(a, &dst[i]);
    __builtin_nontemporal_store(b, &dst[i + S]);
    __builtin_nontemporal_store(c, &dst[i + 2 * S]);
    __builtin_nontemporal_store(d, &dst[i + 3 * S]);
  }
  for (; i < n4; i += S) __builtin_nontemporal_store(src[i], &dst[i]);
}
// 3: one element per thread (the grid covers the array), non-temporal store
__global__ void k_copy_flat(const dvec4* __restrict__ src, dvec4* __restrict__ dst, long long n4) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i < n4) __builtin_nontemporal_store(src[i], &dst[i]);
}

}  // namespace

// Plain device copy of `bytes` (for calibrating HBM byte counters and the
// achievable bandwidth); returns the kernel time in ms.
float copy_calibration(Ctx& C, int64_t bytes) {
  const long long n4 = bytes / 32;
  DevBuf<double4> a, b;
  a.alloc(n4);
  b.alloc(n4);
  H3_HIP(hipMemsetAsync(a.p, 0, n4 * 32, C.stream));
  H3_HIP(hipMemsetAsync(b.p, 0, n4 * 32, C.stream));
  hipLaunchKernelGGL(k_copy, dim3(4096), dim3(256), 0, C.stream, (const double4*)a.p, b.p, n4);  // warm
  H3_LAUNCH_CHECK();
  H3_HIP(hipEventRecord(C.ev0, C.stream));
  hipLaunchKernelGGL(k_copy, dim3(4096), dim3(256), 0, C.stream, (const double4*)a.p, b.p, n4);
  H3_LAUNCH_CHECK();
  H3_HIP(hipEventRecord(C.ev1, C.stream));
  H3_HIP(hipEventSynchronize(C.ev1));
  float ms = 0;
  H3_HIP(hipEventElapsedTime(&ms, C.ev0, C.ev1));
  return ms;
}

// Mean time of `reps` launches of copy variant `variant` (0 grid-stride,
// 1 grid-stride non-temporal, 2 grid-stride with four loads in flight,
// 3 one element per thread) of `bytes`, with `grid` x `block` threads
// (grid ignored by variant 3).
float copy_bandwidth(Ctx& C, int64_t bytes, int variant, int grid, int block, int reps) {
  const long long n4 = bytes / 32;
  DevBuf<double4> a, b;
  a.alloc(n4);
  b.alloc(n4);
  H3_HIP(hipMemsetAsync(a.p, 0, n4 * 32, C.stream));
  H3_HIP(hipMemsetAsync(b.p, 0, n4 * 32, C.stream));
  const dvec4* src = reinterpret_cast<const dvec4*>(a.p);
  dvec4* dst = reinterpret_cast<dvec4*>(b.p);
  auto go = [&]() {
    switch (variant) {
      case 0: hipLaunchKernelGGL(k_copy, dim3(grid), dim3(block), 0, C.stream, (const double4*)a.p, b.p, n4); break;
      case 1: hipLaunchKernelGGL(k_copy_nt, dim3(grid), dim3(block), 0, C.stream, src, dst, n4); break;
      case 2: hipLaunchKernelGGL(k_copy_x4, dim3(grid), dim3(block), 0, C.stream, src, dst, n4); break;
      default:
        hipLaunchKernelGGL(k_copy_flat, dim3((unsigned)((n4 + block - 1) / block)), dim3(block), 0, C.stream, src,
                           dst, n4);
        break;
    }
    H3_LAUNCH_CHECK();
  };
  go();  // warm
  H3_HIP(hipEventRecord(C.ev0, C.stream));
  for (int r = 0; r < reps; ++r) go();
  H3_HIP(hipEventRecord(C.ev1, C.stream));
  H3_HIP(hipEventSynchronize(C.ev1));
  float ms = 0;
  H3_HIP(hipEventElapsedTime(&ms, C.ev0, C.ev1));
  return ms / reps;
}

// role of the buffer being filled: 0 lhs, 1 rhs, 2 stage (SSP-RK3); the
// Euler coupling reads the other solver's buffer of the same role
BCTable make_bc_table(Ctx& C, SolverDev& S, int role) {
  BCTable B;
  B.nbc = (int)S.bcs.size();
  for (int b = 0; b < kMaxBC; ++b) {
    B.kind[b] = 0;
    B.nparams[b] = 0;
    B.kratio[b] = 1.0;
    B.other_lhs[b] = nullptr;
    B.other_ld[b] = 0;
    B.other_gamma[b] = 0;
    B.other_physics[b] = 0;
    for (int k = 0; k < 8; ++k) B.params[b][k] = 0;
  }
  for (size_t b = 0; b < S.bcs.size(); ++b) {
    const BCDesc& d = S.bcs[b];
    B.kind[b] = d.kind;
    B.nparams[b] = d.nparams;
    B.kratio[b] = d.kratio;
    for (int k = 0; k < 8; ++k) B.params[b][k] = d.params[k];
    if (d.other_sid >= 0 && d.other_sid < (int)C.solvers.size() && C.solvers[d.other_sid]) {
      SolverDev& O = *C.solvers[d.other_sid];
      B.other_lhs[b] = O.lhs.p;
      if (d.kind == HOM3G_BC_EULER_COUPLING) B.other_lhs[b] = role == 0 ? O.lhs.p : (role == 1 ? O.rhs.p : O.stg.p);
      B.other_ld[b] = O.cap;
      B.other_gamma[b] = O.P.gamma;
      B.other_physics[b] = O.P.physics;
    }
  }
  return B;
}

// ---- host-callback conditions ------------------------------------------------
// state of m ghosts (slots [0, nv m)) and of their boundary cells (slots
// [nv m, 2 nv m)) packed for the host, and written back after the callback
__global__ void k_pack_host_bc(const double* buf, long long ld, int nv, long long m, const int32_t* g,
                               const int32_t* b, double* out) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i >= m) return;
  for (int v = 0; v < nv; ++v) {
    out[v * m + i] = buf[v * ld + g[i]];
    out[(nv + v) * m + i] = buf[v * ld + b[i]];
  }
}
__global__ void k_unpack_host_bc(double* buf, long long ld, int nv, long long m, const int32_t* g,
                                 const int32_t* b, const double* in, const double* before) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i >= m) return;
  for (int v = 0; v < nv; ++v) {
    buf[v * ld + g[i]] = in[v * m + i];
    // boundary cells only where the condition overwrote them
    const double qb = in[(nv + v) * m + i];
    if (qb != before[(nv + v) * m + i]) buf[v * ld + b[i]] = qb;
  }
}

static void prepare_host_bcs(Ctx& C, SolverDev& S) {
  const int64_t ng = S.n_ghost, ld = S.cap;
  std::vector<int32_t> gapp(ng), gb(ng);
  std::vector<double> xc((size_t)S.nd * ld);
  S.g_applybc.download(gapp.data(), ng, C.stream);
  S.g_bndry.download(gb.data(), ng, C.stream);
  S.x_center.download(xc.data(), xc.size(), C.stream);
  H3_HIP(hipStreamSynchronize(C.stream));
  S.host_bc_runs.clear();
  S.host_bc_runs.reserve(S.bcs.size());
  size_t mmax = 0;
  for (size_t b = 0; b < S.bcs.size(); ++b) {
    if (S.bcs[b].kind != HOM3G_BC_HOST_CALLBACK) continue;
    S.host_bc_runs.push_back(std::make_unique<HostBCRun>());
    HostBCRun& R = *S.host_bc_runs.back();
    R.bc = (int)b;
    for (int64_t g = 0; g < ng; ++g)
      if (gapp[g] == (int)b) { R.ghosts.push_back(S.n_int + g); R.bndry.push_back(gb[g]); }
    const size_t m = R.ghosts.size();
    mmax = std::max(mmax, m);
    R.xg.resize((size_t)S.nd * m);
    R.xb.resize((size_t)S.nd * m);
    std::vector<int32_t> g32(m), b32(m);
    for (size_t i = 0; i < m; ++i) {
      g32[i] = (int32_t)R.ghosts[i];
      b32[i] = (int32_t)R.bndry[i];
      for (int d = 0; d < S.nd; ++d) {
        R.xg[d * m + i] = xc[(size_t)d * ld + R.ghosts[i]];
        R.xb[d * m + i] = xc[(size_t)d * ld + R.bndry[i]];
      }
    }
    if (m) {
      R.dg.alloc(m);
      R.db.alloc(m);
      R.dg.upload(g32.data(), m, C.stream);
      R.db.upload(b32.data(), m, C.stream);
    }
  }
  const size_t slots = std::max<size_t>(1, 2 * (size_t)S.nvars * mmax);
  S.hbc_dev.alloc(2 * slots);
  S.hbc_pin.alloc(2 * slots);
  S.hbc_scal.alloc(1);
  H3_HIP(hipStreamSynchronize(C.stream));
  S.host_bc_ready = true;
}

// After the device fill: each host condition gets its ghosts' boundary
// cells (and the solver's time) on the host, fills the ghosts, and they go
// back to the device
void fill_host_ghosts(Ctx& C, SolverDev& S, double* buf, int role) {
  if (!S.host_bc_ready) prepare_host_bcs(C, S);
  const int nv = S.nvars;
  S.scal.download(S.hbc_scal.p, 1, C.stream);
  for (auto& Rp : S.host_bc_runs) {
    HostBCRun& R = *Rp;
    const long long m = (long long)R.ghosts.size();
    if (m == 0) continue;
    const BCDesc& d = S.bcs[R.bc];
    const size_t n2 = 2 * (size_t)nv * m;
    double* before = S.hbc_dev.p + n2;
    hipLaunchKernelGGL(k_pack_host_bc, dim3(nblk(m)), dim3(256), 0, C.stream, (const double*)buf, (long long)S.cap,
                       nv, m, (const int32_t*)R.dg.p, (const int32_t*)R.db.p, S.hbc_dev.p);
    H3_LAUNCH_CHECK();
    H3_HIP(hipMemcpyAsync(before, S.hbc_dev.p, n2 * sizeof(double), hipMemcpyDeviceToDevice, C.stream));
    H3_HIP(hipMemcpyAsync(S.hbc_pin.p, S.hbc_dev.p, n2 * sizeof(double), hipMemcpyDeviceToHost, C.stream));
    H3_HIP(hipStreamSynchronize(C.stream));
    d.cb(d.user, role, S.hbc_scal.p->time, m, R.ghosts.data(), R.bndry.data(), R.xg.data(), R.xb.data(),
         S.hbc_pin.p, S.hbc_pin.p + (size_t)nv * m);
    H3_HIP(hipMemcpyAsync(S.hbc_dev.p, S.hbc_pin.p, n2 * sizeof(double), hipMemcpyHostToDevice, C.stream));
    hipLaunchKernelGGL(k_unpack_host_bc, dim3(nblk(m)), dim3(256), 0, C.stream, buf, (long long)S.cap, nv, m,
                       (const int32_t*)R.dg.p, (const int32_t*)R.db.p, (const double*)S.hbc_dev.p,
                       (const double*)before);
    H3_LAUNCH_CHECK();
    // the pinned staging is reused by the next condition: wait for the copy
    H3_HIP(hipStreamSynchronize(C.stream));
  }
}

void solver_impose_ic_host(Ctx& C, SolverDev& S, hom3g_ic_callback f, void* user) {
  H3_REQUIRE(f != nullptr, "null initial condition");
  const int64_t n = S.n_int, ld = S.cap;
  const int nd = S.nd, nv = S.nvars, ngp = 1 << nd;
  std::vector<double> xc((size_t)nd * ld), len(n), q((size_t)nv * (n > 0 ? n : 1));
  S.x_center.download(xc.data(), xc.size(), C.stream);
  S.length.download(len.data(), n, C.stream);
  H3_HIP(hipStreamSynchronize(C.stream));
  // quadrature/quadrature.hpp:14-74: 2-point Gauss per axis; the cell
  // average ((sum_gp f(x_gp)) V / 2^nd) / V in this order
  const double gpv = 1 / std::sqrt(3);
  for (int64_t c = 0; c < n; ++c) {
    const double dx2 = 0.5 * len[c];
    double result[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int gp = 0; gp < ngp; ++gp) {
      double x[3] = {0, 0, 0}, val[8] = {0, 0, 0, 0, 0, 0, 0, 0};
      for (int d = 0; d < nd; ++d) x[d] = xc[(size_t)d * ld + c] + dx2 * (((gp >> d) & 1) ? gpv : -1 / std::sqrt(3));
      f(user, x, val);
      for (int v = 0; v < nv; ++v) result[v] = result[v] + val[v];
    }
    const double V = (nd == 2) ? len[c] * len[c] : len[c] * (len[c] * len[c]);
    for (int v = 0; v < nv; ++v) q[(size_t)v * n + c] = (result[v] * V / (double)ngp) / V;
  }
  for (int v = 0; v < nv && n > 0; ++v) S.lhs.upload(q.data() + (size_t)v * n, n, C.stream, (size_t)v * ld);
  H3_HIP(hipStreamSynchronize(C.stream));
}

template <class T>
static void fill_ghosts_t(Ctx& C, SolverDev& S, T* buf, int role) {
  if (S.n_ghost == 0) return;
  if (!S.partners_built)
    for (const BCDesc& b : S.bcs)
      H3_REQUIRE(!(b.kind == HOM3G_BC_CHT_FLUID || b.kind == HOM3G_BC_CHT_SOLID || b.kind == HOM3G_BC_EULER_COUPLING),
                 "coupled boundary used before its partner solver is built");
  GhostArgs A;
  A.buf = buf;
  A.own_lhs = S.lhs.p;
  A.ld = S.cap;
  A.n_int = S.n_int;
  A.ng = S.n_ghost;
  A.g_bndry = S.g_bndry.p;
  A.g_pos = S.g_pos.p;
  A.g_apply = S.g_applybc.p;
  A.g_image = S.g_image.p;
  A.g_partner = S.g_partner.p;
  A.g_dist = S.g_dist.p;
  A.xc = S.x_center.p;
  A.nd = S.nd;
  A.nvars = S.nvars;
  A.physics = S.P.physics;
  A.gamma = S.P.gamma;
  const BCTable B = make_bc_table(C, S, role);
  if (S.nd == 2)
    hipLaunchKernelGGL((k_fill_ghosts<2, T>), dim3(nblk(S.n_ghost)), dim3(256), 0, C.stream, A, B,
                       (const DevScalars*)S.scal.p);
  else
    hipLaunchKernelGGL((k_fill_ghosts<3, T>), dim3(nblk(S.n_ghost)), dim3(256), 0, C.stream, A, B,
                       (const DevScalars*)S.scal.p);
  H3_LAUNCH_CHECK();
}

void solver_fill_ghosts(Ctx& C, SolverDev& S, double* buf) {
  const int role = buf == S.lhs.p ? 0 : (buf == S.rhs.p ? 1 : 2);
  fill_ghosts_t<double>(C, S, buf, role);
  if (S.n_ghost > 0 && S.has_host_bc) fill_host_ghosts(C, S, buf, role);
}

// FP32 stage states (rhs32 / stg32): no host-callback or Euler-coupling
// conditions (their ghosts are exchanged in FP64)
void solver_fill_ghosts(Ctx& C, SolverDev& S, float* buf) {
  H3_REQUIRE(!S.has_host_bc, "host-callback conditions need the FP64 state");
  for (const BCDesc& b : S.bcs)
    H3_REQUIRE(b.kind != HOM3G_BC_EULER_COUPLING, "Euler-Euler coupling needs the FP64 state");
  fill_ghosts_t<float>(C, S, buf, buf == S.rhs32.p ? 1 : 2);
}

void solver_alloc_state(Ctx& C, SolverDev& S) {
  const size_t n = (size_t)S.nvars * S.cap;
  S.lhs.alloc(n);
  S.rhs.alloc(n);
  if (S.P.ti == HOM3G_TI_SSP_RK3) S.stg.alloc(n);
  if (S.P.physics == HOM3G_CNS && S.kernel == HOM3G_KERNEL_GENERIC) S.slopes.alloc((size_t)S.nd * n);
  if (S.P.precision == HOM3G_FP32_STATE) {
    S.rhs32.alloc(n);
    if (S.P.ti == HOM3G_TI_SSP_RK3) S.stg32.alloc(n);
  }
  // NaN-initialised like reset_cell (container.hpp:157-172)
  std::vector<double> nan(n, std::nan(""));
  S.lhs.upload(nan.data(), n, C.stream);
  S.rhs.upload(nan.data(), n, C.stream);
  if (S.stg.p) S.stg.upload(nan.data(), n, C.stream);
  S.scal.alloc(1);
  S.host_scal = DevScalars{};
  S.host_scal.tEnd = S.P.time_end;
  S.host_scal.forced_step = -1;
  S.host_scal.dtmin = 0;
  S.host_scal.diverged_step = -1;
  solver_push_scalars(C, S);
  hipLaunchKernelGGL(k_reset_min, dim3(1), dim3(1), 0, C.stream, S.scal.p);
  H3_LAUNCH_CHECK();
  H3_HIP(hipStreamSynchronize(C.stream));
}

void solver_push_scalars(Ctx& C, SolverDev& S) {
  S.scal.upload(&S.host_scal, 1, C.stream);
}
void solver_pull_scalars(Ctx& C, SolverDev& S) {
  S.scal.download(&S.host_scal, 1, C.stream);
  H3_HIP(hipStreamSynchronize(C.stream));
}

void solver_impose_ic(Ctx& C, SolverDev& S, int kind, const double* params, int np) {
  ICArgs A;
  A.kind = kind;
  A.nd = S.nd;
  A.nvars = S.nvars;
  A.physics = S.P.physics;
  A.gamma = S.P.gamma;
  H3_REQUIRE(np <= 48, "too many IC parameters");
  for (int k = 0; k < 48; ++k) A.p[k] = k < np ? params[k] : 0.0;
  A.np = np;
  if (S.nd == 2)
    hipLaunchKernelGGL(k_impose_ic<2>, dim3(nblk(S.n_int)), dim3(256), 0, C.stream, A, (const double*)S.x_center.p,
                       (const double*)S.length.p, (long long)S.cap, (long long)S.n_int, S.lhs.p);
  else
    hipLaunchKernelGGL(k_impose_ic<3>, dim3(nblk(S.n_int)), dim3(256), 0, C.stream, A, (const double*)S.x_center.p,
                       (const double*)S.length.p, (long long)S.cap, (long long)S.n_int, S.lhs.p);
  H3_LAUNCH_CHECK();
}

// Seeded density perturbation (extension): rho *= 1 + amp (2U - 1) at fixed
// velocity and pressure, U = U[0,1) of SplitMix64 keyed by (seed, level,
// Morton index of the cell's node at its level), independent of cell order.
__device__ inline uint64_t splitmix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}

template <int ND>
__global__ void k_perturb(long long n_int, long long ld, const int64_t* node_idx, const int64_t* parent,
                          const int64_t* first_child, uint64_t seed, double amp, double gamma, double* q) {
  const long long c = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (c >= n_int) return;
  constexpr int NV = ND + 2;
  int64_t n = node_idx[c];
  uint64_t m = 0;
  int level = 0;
  while (parent[n] != INT64_MAX) {
    const int64_t p = parent[n];
    m |= (uint64_t)(n - first_child[p]) << (ND * level);
    ++level;
    n = p;
  }
  const uint64_t k = splitmix64(seed ^ splitmix64((uint64_t)level ^ splitmix64(m)));
  const double U = (double)(k >> 11) * (1.0 / 9007199254740992.0);
  const double eps = amp * (2 * U - 1);
  double cvs[NV], pvs[NV];
#pragma unroll
  for (int v = 0; v < NV; ++v) cvs[v] = q[v * ld + c];
  pv_of<ND>(cvs, gamma, pvs);
  pvs[ND] = pvs[ND] * (1 + eps);
  cv_of<ND>(pvs, gamma, cvs);
#pragma unroll
  for (int v = 0; v < NV; ++v) q[v * ld + c] = cvs[v];
}

void solver_apply_perturbation(Ctx& C, SolverDev& S) {
  if (S.perturb_amp == 0 || S.P.physics == HOM3G_HEAT) return;
  GridDev& G = *C.grids[S.grid_id];
  auto go = [&](auto kern) {
    hipLaunchKernelGGL(kern, dim3(nblk(S.n_int)), dim3(256), 0, C.stream, (long long)S.n_int, (long long)S.cap,
                       (const int64_t*)S.node_idx.p, (const int64_t*)G.parent.p, (const int64_t*)G.first_child.p,
                       S.perturb_seed, S.perturb_amp, S.P.gamma, S.lhs.p);
    H3_LAUNCH_CHECK();
  };
  if (S.nd == 2) go(k_perturb<2>);
  else go(k_perturb<3>);
}

// initialize(): dt_ = 0, time_ = 0, step_ = 0, forced step invalid, then set_dt()
void solver_init_time(Ctx& C, SolverDev& S) {
  S.dtmin_current = false;
  S.host_scal.dt = 0;
  S.host_scal.time = 0;
  S.host_scal.step = 0;
  S.host_scal.tEnd = S.P.time_end;
  S.host_scal.forced_dt = 0;
  S.host_scal.forced_step = -1;
  S.host_scal.dtmin = 0;
  S.host_scal.diverged_step = -1;
  solver_push_scalars(C, S);
  hipLaunchKernelGGL(k_reset_min, dim3(1), dim3(1), 0, C.stream, S.scal.p);
  launch_dt(C, S);
  hipLaunchKernelGGL(k_set_dt, dim3(1), dim3(1), 0, C.stream, S.scal.p,
                     S.kernel == HOM3G_KERNEL_BRICK ? S.h_uniform : 0.0);
  H3_LAUNCH_CHECK();
}

void solver_min_dt_device(Ctx& C, SolverDev& S, unsigned long long*) {
  launch_dt(C, S);
}

double solver_min_dt(Ctx& C, SolverDev& S) {
  S.dtmin_current = false;
  hipLaunchKernelGGL(k_reset_min, dim3(1), dim3(1), 0, C.stream, S.scal.p);
  launch_dt(C, S);
  solver_pull_scalars(C, S);
  const double r = ordered_to_dbl(S.host_scal.dtmin);
  hipLaunchKernelGGL(k_reset_min, dim3(1), dim3(1), 0, C.stream, S.scal.p);
  H3_LAUNCH_CHECK();
  return r;
}

// one solve() (solver.hpp:140-145); skip_dt_reduction when force_dt was just
// applied on the device for this step
static void solve_one(Ctx& C, SolverDev& S, bool skip_dt_reduction) {
  const bool brick = (S.kernel == HOM3G_KERNEL_BRICK);
  if (!brick || S.brick_table) solver_fill_ghosts(C, S, S.lhs.p);
  if (!skip_dt_reduction && !(brick && S.dtmin_current)) {
    hipLaunchKernelGGL(k_reset_min, dim3(1), dim3(1), 0, C.stream, S.scal.p);
    launch_dt(C, S);
  }
  hipLaunchKernelGGL(k_set_dt, dim3(1), dim3(1), 0, C.stream, S.scal.p, brick ? S.h_uniform : 0.0);
  H3_LAUNCH_CHECK();
  S.dtmin_current = false;
  const double* src = S.source.p;
  if (brick) {
    // the next step's CFL minimum is fused into the last stage when the
    // boundary conditions leave the internal cells untouched
    const bool fuse = !S.uses_vortex_bc;
    const bool f32 = S.P.precision == HOM3G_FP32_STATE;
    // tiled bricks read the ghosts of each stage's input (the lhs ghosts
    // were filled above)
    auto ghosts = [&](void* buf) {
      if (!S.brick_table) return;
      if (buf == (void*)S.rhs32.p || buf == (void*)S.stg32.p) solver_fill_ghosts(C, S, static_cast<float*>(buf));
      else solver_fill_ghosts(C, S, static_cast<double*>(buf));
    };
    if (S.P.ti == HOM3G_TI_EULER_FWD) {
      brick_stage(C, S, 0, S.lhs.p, S.rhs.p, fuse, true);
      // lhs = rhs on the internal cells; the ghost slots keep this step's
      // apply_bcs(lhs)
      if (S.brick_table && S.n_ghost) {
        hipLaunchKernelGGL(k_copy_ghosts, dim3(nblk(S.n_ghost * S.nvars)), dim3(256), 0, C.stream,
                           (const double*)S.lhs.p, S.rhs.p, (long long)S.cap, (long long)S.n_int,
                           (long long)S.n_ghost, S.nvars);
        H3_LAUNCH_CHECK();
      }
      std::swap(S.lhs.p, S.rhs.p);
    } else if (S.P.ti == HOM3G_TI_HEUN_RK2) {
      void* s1 = f32 ? (void*)S.rhs32.p : (void*)S.rhs.p;
      brick_stage(C, S, 4, S.lhs.p, s1, false);
      ghosts(s1);
      brick_stage(C, S, 1, s1, S.lhs.p, fuse, true);
    } else {
      void* s1 = f32 ? (void*)S.rhs32.p : (void*)S.rhs.p;
      void* s2 = f32 ? (void*)S.stg32.p : (void*)S.stg.p;
      brick_stage(C, S, 0, S.lhs.p, s1, false);
      ghosts(s1);
      brick_stage(C, S, 2, s1, s2, false);
      ghosts(s2);
      brick_stage(C, S, 3, s2, S.lhs.p, fuse, true);
    }
    S.dtmin_current = fuse;
  } else if (S.P.precision == HOM3G_FP32_STATE && S.P.physics != HOM3G_HEAT) {
    // table-driven Euler / CNS in FP32 state: FP32 stage buffers read and
    // written at FP32, the fluxes in FP32, the RK combination in FP64, U^n (lhs)
    // FP64; the ghosts of the stage buffers are filled in FP32
    if (S.P.ti == HOM3G_TI_EULER_FWD) {
      launch_stage(C, S, StageArgs{S.lhs.p, S.lhs.p, S.rhs.p, src, 0});
      if (S.n_ghost) {
        hipLaunchKernelGGL(k_copy_ghosts, dim3(nblk(S.n_ghost * S.nvars)), dim3(256), 0, C.stream,
                           (const double*)S.lhs.p, S.rhs.p, (long long)S.cap, (long long)S.n_int,
                           (long long)S.n_ghost, S.nvars);
        H3_LAUNCH_CHECK();
      }
      std::swap(S.lhs.p, S.rhs.p);
    } else if (S.P.ti == HOM3G_TI_HEUN_RK2) {
      launch_stage_fp32(C, S, S.lhs.p, false, S.rhs32.p, true, 4, nullptr);
      solver_fill_ghosts(C, S, S.rhs32.p);
      launch_stage_fp32(C, S, S.rhs32.p, true, S.lhs.p, false, 1, nullptr);
    } else {
      launch_stage_fp32(C, S, S.lhs.p, false, S.rhs32.p, true, 0, src);
      solver_fill_ghosts(C, S, S.rhs32.p);
      launch_stage_fp32(C, S, S.rhs32.p, true, S.stg32.p, true, 2, src);
      solver_fill_ghosts(C, S, S.stg32.p);
      launch_stage_fp32(C, S, S.stg32.p, true, S.lhs.p, false, 3, src);
    }
  } else if (S.P.precision == HOM3G_FP32_STATE) {
    // table-driven heat in FP32 state: FP32 stage buffers, FP64 lhs; the
    // interface ghosts (CHT) are formed from both solvers' FP64 lhs
    if (S.P.ti == HOM3G_TI_EULER_FWD) {
      launch_heat_stage32(C, S, S.lhs.p, false, S.rhs.p, false, 0, src);
      if (S.n_ghost) {
        hipLaunchKernelGGL(k_copy_ghosts, dim3(nblk(S.n_ghost * S.nvars)), dim3(256), 0, C.stream,
                           (const double*)S.lhs.p, S.rhs.p, (long long)S.cap, (long long)S.n_int,
                           (long long)S.n_ghost, S.nvars);
        H3_LAUNCH_CHECK();
      }
      std::swap(S.lhs.p, S.rhs.p);
    } else if (S.P.ti == HOM3G_TI_HEUN_RK2) {
      launch_heat_stage32(C, S, S.lhs.p, false, S.rhs32.p, true, 4, nullptr);
      solver_fill_ghosts(C, S, S.rhs32.p);
      launch_heat_stage32(C, S, S.rhs32.p, true, S.lhs.p, false, 1, nullptr);
    } else {
      launch_heat_stage32(C, S, S.lhs.p, false, S.rhs32.p, true, 0, src);
      solver_fill_ghosts(C, S, S.rhs32.p);
      launch_heat_stage32(C, S, S.rhs32.p, true, S.stg32.p, true, 2, src);
      solver_fill_ghosts(C, S, S.stg32.p);
      launch_heat_stage32(C, S, S.stg32.p, true, S.lhs.p, false, 3, src);
    }
  } else if (S.P.ti == HOM3G_TI_EULER_FWD) {
    launch_stage(C, S, StageArgs{S.lhs.p, S.lhs.p, S.rhs.p, src, 0});
    // lhs = rhs on the internal cells; the ghost slots keep the values of
    // this step's apply_bcs(lhs)
    if (S.n_ghost) {
      hipLaunchKernelGGL(k_copy_ghosts, dim3(nblk(S.n_ghost * S.nvars)), dim3(256), 0, C.stream,
                         (const double*)S.lhs.p, S.rhs.p, (long long)S.cap, (long long)S.n_int,
                         (long long)S.n_ghost, S.nvars);
      H3_LAUNCH_CHECK();
    }
    std::swap(S.lhs.p, S.rhs.p);
  } else if (S.P.ti == HOM3G_TI_HEUN_RK2) {
    launch_stage(C, S, StageArgs{S.lhs.p, S.lhs.p, S.rhs.p, nullptr, 4});
    solver_fill_ghosts(C, S, S.rhs.p);
    launch_stage(C, S, StageArgs{S.rhs.p, S.lhs.p, S.lhs.p, nullptr, 1});
  } else {
    launch_stage(C, S, StageArgs{S.lhs.p, S.lhs.p, S.rhs.p, src, 0});
    solver_fill_ghosts(C, S, S.rhs.p);
    launch_stage(C, S, StageArgs{S.rhs.p, S.lhs.p, S.stg.p, src, 2});
    solver_fill_ghosts(C, S, S.stg.p);
    launch_stage(C, S, StageArgs{S.stg.p, S.lhs.p, S.lhs.p, src, 3});
  }
  if (!brick) {  // the brick kernels advance in their last stage
    hipLaunchKernelGGL(k_advance, dim3(1), dim3(1), 0, C.stream, S.scal.p);
    H3_LAUNCH_CHECK();
  }
}

void solver_step(Ctx& C, SolverDev& S, bool skip_dt_reduction) { solve_one(C, S, skip_dt_reduction); }

// FNV-1a over the bytes a captured step depends on
static uint64_t fnv_bytes(uint64_t h, const void* p, size_t n) {
  const unsigned char* b = static_cast<const unsigned char*>(p);
  for (size_t i = 0; i < n; ++i) h = (h ^ b[i]) * 1099511628211ull;
  return h;
}

static uint64_t step_key(const SolverDev& S) {
  uint64_t h = 1469598103934665603ull;
  const void* ptrs[] = {S.lhs.p, S.rhs.p, S.stg.p, S.rhs32.p, S.stg32.p, S.scal.p, S.source.p, S.nghbr.p};
  h = fnv_bytes(h, ptrs, sizeof ptrs);
  const int64_t ints[] = {S.kernel, S.n_int, S.n_ghost, S.cap, S.uses_vortex_bc ? 1 : 0, S.dtmin_current ? 1 : 0};
  h = fnv_bytes(h, ints, sizeof ints);
  h = fnv_bytes(h, &S.P, sizeof S.P);
  h = fnv_bytes(h, &S.pc, sizeof S.pc);
  if (!S.bcs.empty()) h = fnv_bytes(h, S.bcs.data(), S.bcs.size() * sizeof(BCDesc));
  return h | 1;  // 0 = no graph
}

// Steps whose launches are the same every time can be replayed from one
// captured graph: not while timing single launches, not forward Euler
// (which swaps the state buffers), not with coupling conditions (their
// partner's buffers are stepped by the coupled drivers)
static bool step_graph_ok(const Ctx& C, const SolverDev& S) {
  if (C.prof_on || S.P.ti == HOM3G_TI_EULER_FWD) return false;
  for (const BCDesc& b : S.bcs)
    if (b.kind == HOM3G_BC_CHT_FLUID || b.kind == HOM3G_BC_CHT_SOLID || b.kind == HOM3G_BC_EULER_COUPLING ||
        b.kind == HOM3G_BC_HOST_CALLBACK)
      return false;
  return true;
}

void solver_solve(Ctx& C, SolverDev& S, int nsteps) {
  if (nsteps < 3 || !step_graph_ok(C, S)) {
    for (int i = 0; i < nsteps; ++i) solve_one(C, S, false);
    return;
  }
  // one eager step brings the per-step flags to their steady values; the
  // same step is then captured once (capture does not run it) and replayed
  solve_one(C, S, false);
  const bool dc = S.dtmin_current;
  const uint64_t key = step_key(S);
  StepGraph& G = S.step_graph;
  if (!G.exec || G.key != key) {
    G.drop();
    hipGraph_t g = nullptr;
    H3_HIP(hipStreamBeginCapture(C.stream, hipStreamCaptureModeThreadLocal));
    try {
      solve_one(C, S, false);
    } catch (...) {
      (void)hipStreamEndCapture(C.stream, &g);
      if (g) (void)hipGraphDestroy(g);
      S.dtmin_current = dc;
      throw;
    }
    H3_HIP(hipStreamEndCapture(C.stream, &g));
    G.graph = g;
    H3_HIP(hipGraphInstantiate(&G.exec, G.graph, nullptr, nullptr, 0));
    G.key = key;
    H3_REQUIRE(S.dtmin_current == dc, "a captured step must leave the step flags as it found them");
  }
  for (int i = 1; i < nsteps; ++i) H3_HIP(hipGraphLaunch(G.exec, C.stream));
}

// the CFL minimum of a solver's current lhs: the one a brick solver's last
// stage fused in, or a reduction
static void current_min_dt(Ctx& C, SolverDev& S) {
  if (S.dtmin_current) return;
  hipLaunchKernelGGL(k_reset_min, dim3(1), dim3(1), 0, C.stream, S.scal.p);
  launch_dt(C, S);
}

void couple_cht_step(Ctx& C, SolverDev& cns, SolverDev& heat, double Tr0) {
  // coupling::apply: min_dt of both solvers (on their current lhs)
  current_min_dt(C, cns);
  current_min_dt(C, heat);
  cns.dtmin_current = heat.dtmin_current = false;
  const double scaling = Tr0 * cns.pc.Pr0 * cns.pc.Re0;
  hipLaunchKernelGGL(k_couple_cht, dim3(1), dim3(1), 0, C.stream, cns.scal.p, heat.scal.p, scaling);
  H3_LAUNCH_CHECK();
  solve_one(C, cns, true);
  solve_one(C, heat, true);
}

void couple_same_step(Ctx& C, SolverDev& a, SolverDev& b) {
  current_min_dt(C, a);
  current_min_dt(C, b);
  a.dtmin_current = b.dtmin_current = false;
  hipLaunchKernelGGL(k_couple_same, dim3(1), dim3(1), 0, C.stream, a.scal.p, b.scal.p);
  H3_LAUNCH_CHECK();
  solve_one(C, a, true);
  solve_one(C, b, true);
}

void solver_conservation_sums(Ctx& C, SolverDev& S, double* out) {
  const int nb = 256;
  DevBuf<double> part, res;
  part.alloc((size_t)nb * S.nvars);
  res.alloc(S.nvars);
  hipLaunchKernelGGL(k_cons_partial, dim3(nb), dim3(256), 0, C.stream, (const double*)S.lhs.p, (long long)S.cap,
                     (const double*)S.length.p, (long long)S.n_int, S.nd, S.nvars, part.p);
  hipLaunchKernelGGL(k_cons_final, dim3(1), dim3(64), 0, C.stream, (const double*)part.p, nb, S.nvars, res.p);
  H3_LAUNCH_CHECK();
  res.download(out, S.nvars, C.stream);
  H3_HIP(hipStreamSynchronize(C.stream));
}

// ---- allocation-free conservation sums --------------------------------------
// Scratch of solver_conservation_sums_noalloc: kNoAllocBlocks partial sums
// per variable, for at most kNoAllocMaxVars variables (16 KiB), in a
// file-scope device array.  Every GPU the library runs on holds its own copy
// until the process ends; every call on a GPU shares it (each call rewrites
// all the partials it reads, so nothing needs clearing; calls that overlap
// on one GPU race).  Host code never uses the array's name: it uses the
// address hipGetSymbolAddress returns on that device (Ctx::scratch).
constexpr int kNoAllocBlocks = 256, kNoAllocMaxVars = 8;
static __device__ __align__(256) double scratch[kNoAllocBlocks * kNoAllocMaxVars];

void noalloc_init(Ctx& C) {
  void* p = nullptr;
  H3_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  C.scratch = static_cast<double*>(p);
}

void solver_conservation_sums_noalloc(Ctx& C, SolverDev& S, double* sums_dev) {
  H3_REQUIRE(C.scratch != nullptr, "the allocation-free scratch is not set up on this device");
  H3_REQUIRE(sums_dev != nullptr, "null output");
  if (S.nvars > kNoAllocMaxVars)
    throw Error(HOM3G_ERR_CAPACITY, "conservation sums without allocation: at most 8 variables per solver");
  hipLaunchKernelGGL(k_cons_partial, dim3(kNoAllocBlocks), dim3(256), 0, C.stream, (const double*)S.lhs.p,
                     (long long)S.cap, (const double*)S.length.p, (long long)S.n_int, S.nd, S.nvars, C.scratch);
  hipLaunchKernelGGL(k_cons_final, dim3(1), dim3(64), 0, C.stream, (const double*)C.scratch, kNoAllocBlocks, S.nvars,
                     sums_dev);
  H3_LAUNCH_CHECK();
}

int64_t solver_check_state(Ctx& C, SolverDev& S) {
  DevBuf<unsigned long long> bad;
  bad.alloc(1);
  unsigned long long z = 0;
  bad.upload(&z, 1, C.stream);
  if (S.nd == 2)
    hipLaunchKernelGGL(k_check<2>, dim3(nblk(S.n_int)), dim3(256), 0, C.stream, (const double*)S.lhs.p,
                       (long long)S.cap, (long long)S.n_int, S.P.physics, S.P.gamma, bad.p);
  else
    hipLaunchKernelGGL(k_check<3>, dim3(nblk(S.n_int)), dim3(256), 0, C.stream, (const double*)S.lhs.p,
                       (long long)S.cap, (long long)S.n_int, S.P.physics, S.P.gamma, bad.p);
  H3_LAUNCH_CHECK();
  bad.download(&z, 1, C.stream);
  H3_HIP(hipStreamSynchronize(C.stream));
  return (int64_t)z;
}

}  // namespace hom3g
