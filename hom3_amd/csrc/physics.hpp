// Device physics for the hom3 finite-volume hot path on gfx950.
//
// One set of __device__ functions serves the table-driven (generic) kernels
// and the brick kernels.  Formulas follow the reference's physics:
//   euler/physics.hpp:83-149   primitive helpers (u, p, a, T)
//   euler/physics.hpp:276-325  AUSM (Liou-Steffen), theta_ = a * (rho u, rho, rho E + p)
//   euler/physics.hpp:98-113, 262-267  local Lax-Friedrichs
//   cns/physics.hpp:221-338    three-point viscous flux (chain-rule gradients,
//                              Sutherland viscosity from face-averaged p, rho)
//   heat/physics.hpp:113-119   three-point heat flux
//   euler/physics.hpp:182-196, cns/physics.hpp:167-187, heat/physics.hpp:82-85  cell dt
// The values are the reference's; the arithmetic is reorganised for the GPU
// (each shared sub-expression evaluated once, x*x for pow(x, 2.), t*sqrt(t)
// for pow(t, 1.5), FMA contraction), which moves results by a few ULP.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace hom3g {

struct PhysConst {
  int physics, nd, nvars, flux, recon, limiter;
  double gamma, gm1, cfl, cfl_v, Re0, Pr0, sOverTref, sOverTrefP1;
  // 1 / (2 (1 + S/T0) max(1/Re0, 1/(Re0 Pr0 (g-1)))): the viscous CFL bound
  // cfl_v h^2 rho (T + S/T0) / (T a) times this (the fast bounds below)
  double visc_dt;
  double q_coef, inv_Re0;  // g / ((g-1) Pr0) and 1 / Re0 (CNS; 0 otherwise)
};

template <class T> __device__ __forceinline__ T dev_sqrt(T x);
template <> __device__ __forceinline__ double dev_sqrt<double>(double x) { return __builtin_sqrt(x); }
template <> __device__ __forceinline__ float dev_sqrt<float>(float x) { return __builtin_sqrtf(x); }

// Reciprocal: IEEE-correct in FP64; the hardware reciprocal (1 ULP) in FP32,
// whose state error budget (1e-5 relative) is far above it.
template <class T> __device__ __forceinline__ T dev_rcp(T x);
// v_rcp_f64 refined by two Newton steps: within an ulp of 1/x for the
// normal, finite arguments it gets (densities, sound speeds, temperatures),
// at half the instructions of the IEEE division sequence
template <> __device__ __forceinline__ double dev_rcp<double>(double x) {
  double r = __builtin_amdgcn_rcp(x);
  double e = __builtin_fma(-x, r, 1.0);
  r = __builtin_fma(r, e, r);
  e = __builtin_fma(-x, r, 1.0);
  return __builtin_fma(r, e, r);
}
template <> __device__ __forceinline__ float dev_rcp<float>(float x) { return __builtin_amdgcn_rcpf(x); }
template <class T> __device__ __forceinline__ T fast_sqrt(T x) { return dev_sqrt(x); }
// v_rsq_f64 refined by two Goldschmidt steps and a final residual
// correction: within an ulp of sqrt(x) for normal positive x, without the
// range scaling of the general sequence
template <> __device__ __forceinline__ double fast_sqrt<double>(double x) {
  const double y = __builtin_amdgcn_rsq(x);
  double g = x * y, h = 0.5 * y;
  double r = __builtin_fma(-g, h, 0.5);
  g = __builtin_fma(g, r, g);
  h = __builtin_fma(h, r, h);
  r = __builtin_fma(-g, h, 0.5);
  g = __builtin_fma(g, r, g);
  h = __builtin_fma(h, r, h);
  const double d = __builtin_fma(-g, g, x);
  return __builtin_fma(d, h, g);
}
template <> __device__ __forceinline__ float fast_sqrt<float>(float x) { return __builtin_amdgcn_sqrtf(x); }
// 1/sqrt(x) for normal positive x: v_rsq_f64 refined by two Newton steps,
// y += y (1 - x y^2) / 2 (within an ulp); the hardware rsq (1 ULP) in FP32
template <class T> __device__ __forceinline__ T fast_rsq(T x);
template <> __device__ __forceinline__ double fast_rsq<double>(double x) {
  double y = __builtin_amdgcn_rsq(x);
  double e = __builtin_fma(-(x * y), y, 1.0);
  y = __builtin_fma(0.5 * y, e, y);
  e = __builtin_fma(-(x * y), y, 1.0);
  return __builtin_fma(0.5 * y, e, y);
}
template <> __device__ __forceinline__ float fast_rsq<float>(float x) { return __builtin_amdgcn_rsqf(x); }

// Sutherland's law (cns/quantities.hpp:113-115): T^1.5 (1 + S/T0) / (T + S/T0)
template <class T>
__device__ __forceinline__ T sutherland(T Tt, T s, T s1) {
  return Tt * dev_sqrt(Tt) * s1 / (Tt + s);
}

// Eigen fixed-size reduction order for the squared norm (a0 + (a1 + a2))
template <int ND, class T>
__device__ __forceinline__ T esum_sq(const T* u) {
  if constexpr (ND == 2) return u[0] * u[0] + u[1] * u[1];
  else return u[0] * u[0] + (u[1] * u[1] + u[2] * u[2]);
}

// p = (g-1)(rho E - rho |u|^2 / 2) and 1/(rho a) = (g p rho)^(-1/2),
// a = sqrt(g p / rho), of a conserved state: the per-position values of the
// brick kernels (forward and reverse)
template <int ND, class T>
__device__ __forceinline__ void brick_prims(const T* q, T gamma, T gm1, T& p, T& ia) {
  const T ir = dev_rcp(q[ND]);
  T u[ND];
#pragma unroll
  for (int d = 0; d < ND; ++d) u[d] = q[d] * ir;
  p = gm1 * (q[ND + 1] - T(0.5) * q[ND] * esum_sq<ND>(u));
  ia = fast_rsq(gamma * p * q[ND]);
}

// Cell primitives: u_d = rho u_d / rho, p = (g-1)(rho E - 0.5 rho |u|^2), a = sqrt(g p / rho)
template <int ND, class T>
struct Prim {
  T u[ND];
  T rho, p, a;
};

template <int ND, class T>
__device__ __forceinline__ Prim<ND, T> primitives(const T* q, T gamma, T gm1) {
  Prim<ND, T> w;
  w.rho = q[ND];
  for (int d = 0; d < ND; ++d) w.u[d] = q[d] / w.rho;
  w.p = gm1 * (q[ND + 1] - T(0.5) * w.rho * esum_sq<ND>(w.u));
  w.a = dev_sqrt(gamma * w.p / w.rho);
  return w;
}

// AUSM split Mach / pressure polynomials (euler/physics.hpp:297-310)
template <int S, class T>
__device__ __forceinline__ T m_int(T M) {
  const T s = T(S);
  const T aM = M < 0 ? -M : M;
  const T mp = M + s;
  return aM > T(1) ? T(0.5) * (M + s * aM) : s * T(0.25) * (mp * mp);
}
template <int S, class T>
__device__ __forceinline__ T p_int(T M) {
  const T s = T(S);
  const T aM = M < 0 ? -M : M;
  const T mp = M + s;
  return aM > T(1) ? T(0.5) * (M + s * aM) / M : T(0.25) * (mp * mp) * (T(2) - s * M);
}

// AUSM flux along direction D between left/right conserved states with their
// primitives precomputed.  f[v] for v < nvars.
template <int ND, int D, class T>
__device__ __forceinline__ void ausm_flux(const T* qL, const Prim<ND, T>& wL, const T* qR,
                                          const Prim<ND, T>& wR, T* f) {
  const T ML = wL.u[D] / wL.a;
  const T MR = wR.u[D] / wR.a;
  const T m_i = m_int<+1>(ML) + m_int<-1>(MR);
  const T p_i = p_int<+1>(ML) * wL.p + p_int<-1>(MR) * wR.p;
  const bool left = m_i >= T(0);
  const T* q = left ? qL : qR;
  const T ac = left ? wL.a : wR.a;
  const T pc = left ? wL.p : wR.p;
#pragma unroll
  for (int k = 0; k < ND; ++k) f[k] = (q[k] * ac) * m_i;
  f[ND] = (q[ND] * ac) * m_i;
  f[ND + 1] = ((q[ND + 1] + pc) * ac) * m_i;
  f[D] += p_i;
}

// First-order and MUSCL Euler paths (table-driven, batched and brick
// kernels): the primitives with one refined reciprocal of rho instead of
// nd + 1 divisions, and AUSM with the Mach numbers from refined reciprocals
// of a and the supersonic pressure weights in closed form (exactly 1 or 0).
// Values agree with primitives() / ausm_flux() to rounding.
template <int ND, class T>
__device__ __forceinline__ Prim<ND, T> primitives_rcp(const T* q, T gamma, T gm1) {
  Prim<ND, T> w;
  w.rho = q[ND];
  const T ir = dev_rcp(w.rho);
#pragma unroll
  for (int d = 0; d < ND; ++d) w.u[d] = q[d] * ir;
  w.p = gm1 * (q[ND + 1] - T(0.5) * w.rho * esum_sq<ND>(w.u));
  w.a = fast_sqrt(gamma * w.p * ir);
  return w;
}
template <int S, class T> __device__ __forceinline__ T p_int_fast(T M);
template <class T>
__device__ __forceinline__ void ausm_split(T ML, T MR, T pL, T pR, T& m_i, T& p_i);
template <int ND, int D, class T>
__device__ __forceinline__ void ausm_flux_rcp(const T* qL, const Prim<ND, T>& wL, const T* qR,
                                              const Prim<ND, T>& wR, T* f) {
  const T ML = wL.u[D] * dev_rcp(wL.a);
  const T MR = wR.u[D] * dev_rcp(wR.a);
  T m_i, p_i;
  ausm_split(ML, MR, wL.p, wR.p, m_i, p_i);
  const bool left = m_i >= T(0);
  const T* q = left ? qL : qR;
  const T ac = left ? wL.a : wR.a;
  const T pc = left ? wL.p : wR.p;
#pragma unroll
  for (int k = 0; k < ND; ++k) f[k] = (q[k] * ac) * m_i;
  f[ND] = (q[ND] * ac) * m_i;
  f[ND + 1] = ((q[ND + 1] + pc) * ac) * m_i;
  f[D] += p_i;
}
template <int ND, class T>
__device__ __forceinline__ void ausm_flux_rcp_rt(const T* qL, const Prim<ND, T>& wL, const T* qR,
                                                 const Prim<ND, T>& wR, int d, T* f) {
  if (d == 0) ausm_flux_rcp<ND, 0, T>(qL, wL, qR, wR, f);
  else if (d == 1) ausm_flux_rcp<ND, 1, T>(qL, wL, qR, wR, f);
  else if constexpr (ND == 3) ausm_flux_rcp<ND, 2, T>(qL, wL, qR, wR, f);
}

// Runtime-direction variant (generic kernels)
template <int ND, class T>
__device__ __forceinline__ void ausm_flux_rt(const T* qL, const Prim<ND, T>& wL, const T* qR,
                                             const Prim<ND, T>& wR, int d, T* f) {
  if (d == 0) ausm_flux<ND, 0, T>(qL, wL, qR, wR, f);
  else if (d == 1) ausm_flux<ND, 1, T>(qL, wL, qR, wR, f);
  else if constexpr (ND == 3) ausm_flux<ND, 2, T>(qL, wL, qR, wR, f);
}

// Local Lax-Friedrichs (euler/physics.hpp:262-267), normal D
template <int ND, int D, class T>
__device__ __forceinline__ void llf_flux_d(const T* qL, const Prim<ND, T>& wL, const T* qR,
                                           const Prim<ND, T>& wR, T dx_over_dt, T* f) {
  T fl[ND + 2], fr[ND + 2];
  fl[ND] = qL[D];
  fr[ND] = qR[D];
  for (int i = 0; i < ND; ++i) { fl[i] = qL[i] * wL.u[D]; fr[i] = qR[i] * wR.u[D]; }
  fl[D] += wL.p;
  fr[D] += wR.p;
  fl[ND + 1] = wL.u[D] * (qL[ND + 1] + wL.p);
  fr[ND + 1] = wR.u[D] * (qR[ND + 1] + wR.p);
  for (int v = 0; v < ND + 2; ++v) f[v] = T(0.5) * (fl[v] + fr[v] + dx_over_dt * (qL[v] - qR[v]));
}

// runtime direction (a compile-time normal inside, so no local array is
// indexed at run time)
template <int ND, class T>
__device__ __forceinline__ void llf_flux(const T* qL, const Prim<ND, T>& wL, const T* qR,
                                         const Prim<ND, T>& wR, int d, T dx_over_dt, T* f) {
  if (d == 0) llf_flux_d<ND, 0, T>(qL, wL, qR, wR, dx_over_dt, f);
  else if (d == 1) llf_flux_d<ND, 1, T>(qL, wL, qR, wR, dx_over_dt, f);
  else if constexpr (ND == 3) llf_flux_d<ND, 2, T>(qL, wL, qR, wR, dx_over_dt, f);
}

// MUSCL limiter (extension).  LIM_ZERO returns 0 -> first order exactly.
template <class T>
__device__ __forceinline__ T limit(int lim, T a, T b) {
  if (lim == 0) return T(0);
  if (a * b <= T(0)) return T(0);
  if (lim == 1) return (a < 0 ? -a : a) < (b < 0 ? -b : b) ? a : b;
  return T(2) * a * b / (a + b);
}

// Three-point viscous flux (cns/physics.hpp:305-338) for a face with normal
// D.  Inputs: the two cell states, their pressures, and the surface gradient
// of every conserved variable G[v][sd] (surface_slope, solver.hpp:656-678).
template <int ND, int D, class T>
__device__ __forceinline__ void viscous_flux(const T* qL, T pL, const T* qR, T pR,
                                             const T (*G)[ND], const PhysConst& pc, T* f) {
  const int RHO = ND, E = ND + 1;
  const T gamma = T(pc.gamma), gm1 = T(pc.gm1);
  const T rho_s = T(0.5) * (qL[RHO] + qR[RHO]);
  const T inv_rho_s = T(1) / rho_s;
  T u_s[ND], rhou_s[ND];
  T umag2 = T(0);
#pragma unroll
  for (int k = 0; k < ND; ++k) {
    rhou_s[k] = T(0.5) * (qL[k] + qR[k]);
    u_s[k] = rhou_s[k] * inv_rho_s;
    umag2 += u_s[k] * u_s[k];
  }
  const T rhoE_s = T(0.5) * (qL[E] + qR[E]);
  const T p_s = gm1 * (rhoE_s - T(0.5) * rho_s * umag2);
  // grad_u_srfc(c, sd) = (G[c][sd] rho_s - G[rho][sd] rhou_s[c]) / rho_s^2
  T gu[ND][ND];
  const T inv_rho2 = inv_rho_s * inv_rho_s;
#pragma unroll
  for (int c = 0; c < ND; ++c)
#pragma unroll
    for (int sd = 0; sd < ND; ++sd) gu[c][sd] = (G[c][sd] * rho_s - G[RHO][sd] * rhou_s[c]) * inv_rho2;
  // grad_p_srfc(D)
  T grho_umag2 = G[RHO][D] * umag2;
#pragma unroll
  for (int k = 0; k < ND; ++k) grho_umag2 += rho_s * T(2) * u_s[k] * gu[k][D];
  const T gradP = gm1 * (G[E][D] - T(0.5) * grho_umag2);
  // Sutherland viscosity from the face averages of the cell p and rho
  const T Tm = gamma * (T(0.5) * (pL + pR)) / (T(0.5) * (qL[RHO] + qR[RHO]));
  const T mu = sutherland(Tm, T(pc.sOverTref), T(pc.sOverTrefP1));
  // heat flux (cns/physics.hpp:101-105)
  const T q = -mu * gamma / (gm1 * T(pc.Pr0) * rho_s) * (gradP - p_s * inv_rho_s * G[RHO][D]);
  // shear stress column D (cns/physics.hpp:110-123): nabla_u(i,j) = gu[j][i]
  T div;
  if constexpr (ND == 2) div = gu[0][0] + gu[1][1];
  else div = gu[0][0] + (gu[1][1] + gu[2][2]);
  div = div * T(2) / T(3);
  const T inv_Re0 = T(1) / T(pc.Re0);
  T fE = T(0);
#pragma unroll
  for (int i = 0; i < ND; ++i) {
    T t = gu[D][i] + gu[i][D];
    if (i == D) t -= div;
    t = -mu * t;
    f[i] = inv_Re0 * t;
    fE += u_s[i] * t;
  }
  f[RHO] = T(0);
  f[E] = inv_Re0 * (fE + q);
}


// ---------------------------------------------------------------------------
// Brick-kernel forms.  Same formulas with fewer instructions: reciprocals of
// rho and a precomputed per cell, the supersonic branch of p_int written in
// its exact closed form ((M + s|M|) / (2M) is 1 or 0), and the viscous
// gradients built from 1/rho_s.  Results differ from the reference forms by a
// few ULP.
template <int S, class T>
__device__ __forceinline__ T p_int_fast(T M) {
  const T aM = M < 0 ? -M : M;
  const T mp = M + T(S);
  if (aM > T(1)) return (S > 0) ? (M > T(0) ? T(1) : T(0)) : (M < T(0) ? T(1) : T(0));
  return T(0.25) * (mp * mp) * (T(2) - T(S) * M);
}

// The AUSM interface Mach number and pressure, m_int<+1>(ML) + m_int<-1>(MR)
// and p_int_fast<+1>(ML) pL + p_int_fast<-1>(MR) pR.  A wave whose faces are
// all subsonic on both sides (|M| <= 1) takes the polynomial branches with
// no selects (the same formulas; the compiler may contract them into other
// fused multiply-adds, a ULP-level difference).
template <class T>
__device__ __forceinline__ void ausm_split(T ML, T MR, T pL, T pR, T& m_i, T& p_i) {
  const T aML = ML < T(0) ? -ML : ML;
  const T aMR = MR < T(0) ? -MR : MR;
  if (__all(aML <= T(1) && aMR <= T(1))) {
    const T mpL = ML + T(1), mpR = MR - T(1);
    m_i = T(0.25) * (mpL * mpL) + T(-0.25) * (mpR * mpR);
    p_i = (T(0.25) * (mpL * mpL) * (T(2) - ML)) * pL + (T(0.25) * (mpR * mpR) * (T(2) + MR)) * pR;
  } else {
    m_i = m_int<+1>(ML) + m_int<-1>(MR);
    p_i = p_int_fast<+1>(ML) * pL + p_int_fast<-1>(MR) * pR;
  }
}

// AUSM along D with p and a of both sides precomputed; the Mach numbers are
// uL * iaL and uR * iaR (u and 1/a, or rho u and 1/(rho a))
template <int ND, int D, class T>
__device__ __forceinline__ void ausm_flux_fast(const T* qL, T uL, T pL, T aL, T iaL, const T* qR, T uR, T pR,
                                               T aR, T iaR, T* f) {
  const T ML = uL * iaL;
  const T MR = uR * iaR;
  T m_i, p_i;
  ausm_split(ML, MR, pL, pR, m_i, p_i);
  const bool left = m_i >= T(0);
  const T am = (left ? aL : aR) * m_i;
#pragma unroll
  for (int k = 0; k < ND + 1; ++k) f[k] = (left ? qL[k] : qR[k]) * am;
  f[ND + 1] = ((left ? qL[ND + 1] : qR[ND + 1]) + (left ? pL : pR)) * am;
  f[D] += p_i;
}

// Viscous flux for the face with normal D from unscaled differences:
// dn[v] = q_R - q_L (h times the normal gradient), dt[v][t] = the sum of the
// two cells' central differences along tangential direction t (4h times
// the tangential surface gradient; t over the two other directions in
// increasing order).  Same formulas as viscous_flux with the gradient
// scalings folded into 1/rho_s and mu / h, and the face averages
// rho_s = S / 2, u_s = (rho u)_sum / S, T = gamma (p_L + p_R) / S
// (S = rho_L + rho_R); results differ from the reference form by a few ULP.
template <int ND, int D, class T>
__device__ __forceinline__ void viscous_flux_fast(const T* qL, T pL, const T* qR, T pR, const T* dn,
                                                  const T (*dt)[ND - 1], const PhysConst& pc, T q_coef,
                                                  T inv_Re0, T inv_h, T* f) {
  const int RHO = ND, E = ND + 1;
  const T S = qL[RHO] + qR[RHO];
  const T iS = dev_rcp(S);
  const T rho_s = T(0.5) * S;
  const T irs = T(2) * iS;     // 1 / rho_s
  const T irs4 = T(0.5) * iS;  // 1 / (4 rho_s), exact
  T u_s[ND];
  T umag2 = T(0);
#pragma unroll
  for (int k = 0; k < ND; ++k) {
    u_s[k] = (qL[k] + qR[k]) * iS;
    umag2 += u_s[k] * u_s[k];
  }
  // p_s = (g-1)(rhoE_s - rho_s |u_s|^2 / 2)
  const T p_s = T(0.5 * pc.gm1) * ((qL[E] + qR[E]) - rho_s * umag2);
  // h times the velocity gradients du_c/dx_sd = (G[c][sd] - G[rho][sd] u_s[c]) / rho_s:
  // full row c = D (du_D/dx_*), column sd = D (du_*/dx_D) and the diagonal
  T gcol[ND];  // du_k / dx_D
#pragma unroll
  for (int k = 0; k < ND; ++k) gcol[k] = (dn[k] - dn[RHO] * u_s[k]) * irs;
  T grow[ND];  // du_D / dx_k
  T diag[ND];  // du_k / dx_k
#pragma unroll
  for (int k = 0; k < ND; ++k) {
    if (k == D) {
      grow[k] = gcol[D];
      diag[k] = gcol[D];
    } else {
      const int t = (k < D) ? k : k - 1;
      grow[k] = (dt[D][t] - dt[RHO][t] * u_s[D]) * irs4;
      diag[k] = (dt[k][t] - dt[RHO][t] * u_s[k]) * irs4;
    }
  }
  // h dp/dx_D = (g-1)(dn[E] - (dn[rho] |u_s|^2 + sum_k 2 rho_s u_s[k] gcol[k]) / 2)
  T grho_umag2 = dn[RHO] * umag2;
#pragma unroll
  for (int k = 0; k < ND; ++k) grho_umag2 += S * u_s[k] * gcol[k];
  const T gradP = T(pc.gm1) * (dn[E] - T(0.5) * grho_umag2);
  // Sutherland viscosity, T = gamma * avg(p) / avg(rho)
  const T Tm = T(pc.gamma) * (pL + pR) * iS;
  const T mu = Tm * fast_sqrt(Tm) * T(pc.sOverTrefP1) * dev_rcp(Tm + T(pc.sOverTref));
  const T mh = mu * inv_h;
  // heat flux: -mu gamma/((gamma-1) Pr rho_s) (dp - p/rho drho)
  const T q = -mh * q_coef * irs * (gradP - p_s * irs * dn[RHO]);
  T div;
  if constexpr (ND == 2) div = diag[0] + diag[1];
  else div = diag[0] + (diag[1] + diag[2]);
  div = div * T(2.0 / 3.0);
  T fE = T(0);
#pragma unroll
  for (int i = 0; i < ND; ++i) {
    T t = grow[i] + gcol[i];
    if (i == D) t -= div;
    t = -mh * t;
    f[i] = inv_Re0 * t;
    fE += u_s[i] * t;
  }
  f[RHO] = T(0);
  f[E] = inv_Re0 * (fE + q);
}

// Cell time step (CFL), reference formulas
template <class T> struct Lowest;
template <> struct Lowest<double> { static __device__ __forceinline__ double v() { return -1.7976931348623157e308; } };
template <> struct Lowest<float> { static __device__ __forceinline__ float v() { return -3.40282347e38f; } };

// std::max / std::min semantics ((a < b) ? b : a), so a NaN candidate is ignored
template <class T> __device__ __forceinline__ T ref_max(T a, T b) { return (a < b) ? b : a; }
template <class T> __device__ __forceinline__ T ref_min(T a, T b) { return (b < a) ? b : a; }

template <int ND, class T>
__device__ __forceinline__ T cell_dt_euler(const Prim<ND, T>& w, T h, T cfl) {
  T S = Lowest<T>::v();
  for (int d = 0; d < ND; ++d) S = ref_max(S, (w.u[d] < 0 ? -w.u[d] : w.u[d]) + w.a);
  return cfl * h / S;
}

template <int ND, class T>
__device__ __forceinline__ T cell_dt_cns(const Prim<ND, T>& w, T h, const PhysConst& pc) {
  const T dt_inv = cell_dt_euler<ND, T>(w, h, T(pc.cfl));
  const T Tc = T(pc.gamma) * w.p / w.rho;
  const T mu = sutherland(Tc, T(pc.sOverTref), T(pc.sOverTrefP1));
  const T S_coeff = mu / (w.rho * h);
  const T S_visc = ref_max(S_coeff / T(pc.Re0), S_coeff / (T(pc.Re0) * T(pc.Pr0) * T(pc.gm1)));
  const T dt_v = T(pc.cfl_v) * h / (T(2) * S_visc);
  return ref_min(dt_inv, dt_v);
}

// CFL bound of one cell (euler/physics.hpp:182-196, cns/physics.hpp:167-187)
// with reciprocals instead of divisions (brick kernels)
template <int ND, bool VISC>
__device__ __forceinline__ double cell_dt_fast(const double* q, double h, const PhysConst& pc) {
  const double ir = dev_rcp(q[ND]);
  double u[ND];
#pragma unroll
  for (int d = 0; d < ND; ++d) u[d] = q[d] * ir;
  const double p = pc.gm1 * (q[ND + 1] - 0.5 * q[ND] * esum_sq<ND>(u));
  const double Tc = pc.gamma * p * ir;
  const double a = fast_sqrt(Tc);
  double S = Lowest<double>::v();
#pragma unroll
  for (int d = 0; d < ND; ++d) S = ref_max(S, (u[d] < 0 ? -u[d] : u[d]) + a);
  const double dt_inv = pc.cfl * h * dev_rcp(S);
  if constexpr (!VISC) return dt_inv;
  // cfl_v h / (2 S_visc), S_visc = mu / (rho h) max(1/Re0, 1/(Re0 Pr0 (g-1))),
  // mu = T^1.5 (1 + S/T0) / (T + S/T0) with T = a^2: one reciprocal
  const double dt_v = pc.cfl_v * (h * h) * pc.visc_dt * q[ND] * (Tc + pc.sOverTref) * dev_rcp(Tc * a);
  return ref_min(dt_inv, dt_v);
}

// The same bound in FP32 arithmetic for the FP32-state brick kernels (the
// state they advance is FP32 to ~1e-7 anyway; the RK accumulation stays
// FP64): single-instruction reciprocals and square roots.
template <int ND, bool VISC>
__device__ __forceinline__ float cell_dt_fast_f32(const double* qd, float h, const PhysConst& pc) {
  float q[ND + 2];
#pragma unroll
  for (int v = 0; v < ND + 2; ++v) q[v] = (float)qd[v];
  const float ir = dev_rcp(q[ND]);
  float u[ND];
#pragma unroll
  for (int d = 0; d < ND; ++d) u[d] = q[d] * ir;
  const float gamma = (float)pc.gamma;
  const float p = (float)pc.gm1 * (q[ND + 1] - 0.5f * q[ND] * esum_sq<ND>(u));
  const float Tc = gamma * p * ir;
  const float a = fast_sqrt(Tc);
  float S = Lowest<float>::v();
#pragma unroll
  for (int d = 0; d < ND; ++d) S = ref_max(S, (u[d] < 0 ? -u[d] : u[d]) + a);
  const float dt_inv = (float)pc.cfl * h * dev_rcp(S);
  if constexpr (!VISC) return dt_inv;
  // the viscous bound as in cell_dt_fast, one reciprocal
  const float dt_v = (float)(pc.cfl_v * pc.visc_dt) * (h * h) * q[ND] * (Tc + (float)pc.sOverTref) * dev_rcp(Tc * a);
  return ref_min(dt_inv, dt_v);
}

// Ordered-integer mapping for exact, order-independent atomic min of doubles
__device__ __forceinline__ unsigned long long dbl_to_ordered(double x) {
  unsigned long long b = __double_as_longlong(x);
  return (b & 0x8000000000000000ull) ? ~b : (b | 0x8000000000000000ull);
}
__host__ __device__ __forceinline__ double ordered_to_dbl(unsigned long long o) {
  unsigned long long b = (o & 0x8000000000000000ull) ? (o & 0x7fffffffffffffffull) : ~o;
  double r;
  __builtin_memcpy(&r, &b, 8);
  return r;
}

// Lowers the running minimum *dst (ordered encoding) to m.  Every block of a
// launch reduces into the same address, and same-address atomics serialise
// in one L2 channel (about 12 ns each here: 65,536 blocks cost 0.76 ms).  The
// minimum only falls while a launch runs (it is reset by an earlier kernel),
// so a block whose value is not below a relaxed read of it cannot lower it
// and skips the atomic; the result is the same minimum.
__device__ __forceinline__ void atomic_lower(unsigned long long* dst, double m) {
  const unsigned long long o = dbl_to_ordered(m);
  if (o < __hip_atomic_load(dst, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) atomicMin(dst, o);
}

}  // namespace hom3g
