// Discrete tangent and adjoint of the coupled fluid/solid step (no reference
// counterpart; the forward map is the reference's: coupling::apply, then
// cns.solve(), then heat.solve(), coupling.hpp:47-63,
// coupled_cns_heat_test.cpp:235-248) on the table-driven FP64 kernels.
//
// Objective: the time-integrated wall heat flux through the conjugate
// interface, J = sum_n dt_heat^n Q(state^{n+1}), Q = sum over the solid's
// interface ghosts of h^(nd-2) 2 (T_b - T_s), T_s the interface temperature
// (heat/coupling_conditions.hpp:24-37).  Parameters: the solid's per-cell
// volumetric source S.  The dt schedule is frozen (a fixed sequence of
// fluid dts, heat dt = dt Tr0 Pr0 Re0), so the map from S to J is smooth.
//
// Tangent: every stage and ghost fill re-evaluated with a forward-mode
// number (value + derivative along dS).
//
// Adjoint: reverse sweep over the steps with two-level checkpointing
// (states every `ckpt_every` steps in HBM; each segment is recomputed once,
// storing the stage buffers of its steps).  Each stage is reversed face by
// face: the adjoint of a face flux is pulled back through its inputs (the
// two cell states and the tangential surface slopes) by a hand-written
// reverse of the flux, and scattered with atomic adds to the cells, to the
// slope stencils (solver.hpp:596-678,
// including the boundary hooks) and through the ghost fills
// (boundary_condition.hpp, cns/ and heat/ boundary and coupling conditions)
// to the boundary cells and the coupled solver.
#include <algorithm>
#include <cstdio>
#include <cstring>

#include "dual.hpp"
#include "internal.hpp"

namespace hom3g {

namespace {

inline unsigned nblk(int64_t n, int bs = 256) { return (unsigned)((n + bs - 1) / bs); }

struct AdjTopo {
  const int32_t* nghbr;
  long long ld, n_int, ng;
  const double* len;
  const int32_t* g_bndry;
  const int32_t* g_hookbc;
  const int32_t* g_image;
  const int32_t* g_apply;
  const int32_t* g_pos;
  const int32_t* g_partner;
  const double* g_dist;
  int nd, nvars;
  double gamma;          // this solver's gamma (fluid) or unused (heat)
  double other_gamma;    // the coupled solver's gamma
  long long other_ld;
  int kind[kMaxBC];
  int nparams[kMaxBC];
  double params[kMaxBC][8];
  double kratio[kMaxBC];
};

AdjTopo make_adj_topo(SolverDev& S, const SolverDev* other) {
  AdjTopo t;
  t.nghbr = S.nghbr.p;
  t.ld = S.cap;
  t.n_int = S.n_int;
  t.ng = S.n_ghost;
  t.len = S.length.p;
  t.g_bndry = S.g_bndry.p;
  t.g_hookbc = S.g_hookbc.p;
  t.g_image = S.g_image.p;
  t.g_apply = S.g_applybc.p;
  t.g_pos = S.g_pos.p;
  t.g_partner = S.g_partner.p;
  t.g_dist = S.g_dist.p;
  t.nd = S.nd;
  t.nvars = S.nvars;
  t.gamma = S.P.gamma;
  t.other_gamma = other ? other->P.gamma : 0.0;
  t.other_ld = other ? other->cap : 0;
  for (int b = 0; b < kMaxBC; ++b) {
    t.kind[b] = -1;
    t.nparams[b] = 0;
    t.kratio[b] = 1.0;
    for (int k = 0; k < 8; ++k) t.params[b][k] = 0;
  }
  for (size_t b = 0; b < S.bcs.size(); ++b) {
    t.kind[b] = S.bcs[b].kind;
    t.nparams[b] = S.bcs[b].nparams;
    t.kratio[b] = S.bcs[b].kratio;
    for (int k = 0; k < 8; ++k) t.params[b][k] = S.bcs[b].params[k];
  }
  return t;
}

__device__ __forceinline__ bool wall_hook_k(int kind) {
  return kind == HOM3G_BC_ADIABATIC_NOSLIP || kind == HOM3G_BC_ISOTHERMAL_NOSLIP || kind == HOM3G_BC_CHT_FLUID;
}

// slope(c, v, dir) = sign (Q[np] - Q[nn]) / len2 after the walk of
// solver.hpp:596-623 (one-sided cells recurse into the neighbour, ghosts
// without a neighbour use the BC hook: wall hooks negate velocities,
// periodic ghosts continue at the image).  np < 0: undefined.
struct SlopeT {
  int np, nn;
  double len2, sign;
  double ilen2;  // 1 / len2
};

__device__ SlopeT slope_walk(const AdjTopo& t, int c, int v, int dir) {
  double sign = 1.0;
  for (int it = 0; it < 16; ++it) {
    const int nn = t.nghbr[(2 * dir) * t.ld + c];
    const int np = t.nghbr[(2 * dir + 1) * t.ld + c];
    if (nn >= 0 && np >= 0) return SlopeT{np, nn, 2. * t.len[c], sign, 1.0 / (2. * t.len[c])};
    if (nn >= 0) { c = nn; continue; }
    if (np >= 0) { c = np; continue; }
    const int g = c - (int)t.n_int;
    if (g < 0 || g >= t.ng) break;  // an internal cell without neighbours: undefined
    const int kind = t.kind[t.g_hookbc[g]];
    if (kind == HOM3G_BC_PERIODIC) { c = t.g_image[g]; continue; }
    if (wall_hook_k(kind) && v < t.nd) sign = -sign;
    c = t.g_bndry[g];
  }
  return SlopeT{-1, -1, 1.0, 0.0, 1.0};
}

template <class T>
__device__ __forceinline__ T slope_value(const SlopeT& s, T qp, T qn) {
  const T d = (qp - qn) * s.ilen2;
  return s.sign < 0 ? -d : d;
}

// ---- templated forward pieces ------------------------------------------------

// CNS/Euler face flux of the table-driven kernel (AUSM + three-point viscous)
// with the tangential surface gradients given and the normal one formed
// from the two states.
template <int ND, int PHYS, int D, class T>
__device__ void flux_face(const PhysConst& pc, const T* qL, const T* qR, T (*G)[ND], double inv_len_r, T* f) {
  constexpr int NV = ND + 2;
  const Prim<ND, T> wL = primitives<ND>(qL, T(pc.gamma), T(pc.gm1));
  const Prim<ND, T> wR = primitives<ND>(qR, T(pc.gamma), T(pc.gm1));
  ausm_flux<ND, D, T>(qL, wL, qR, wR, f);
  if constexpr (PHYS == HOM3G_CNS) {
#pragma unroll
    for (int v = 0; v < NV; ++v) G[v][D] = (qR[v] - qL[v]) * inv_len_r;
    T fv[NV];
    viscous_flux<ND, D, T>(qL, wL.p, qR, wR.p, G, pc, fv);
#pragma unroll
    for (int v = 0; v < NV; ++v) f[v] = f[v] + fv[v];
  }
}

// ---- reverse of the face flux ------------------------------------------------
//
// Hand-written reverse mode of flux_face: one forward evaluation of the
// intermediates, then their adjoints in reverse order.  Piecewise formulas
// take the branch of the forward values, as the forward-mode number does.

// d m_int<S>/dM, d p_int<S>/dM (AUSM split polynomials, euler/physics.hpp:297-310)
template <int S>
__device__ __forceinline__ double dm_int(double M) {
  const double aM = M < 0 ? -M : M;
  if (aM > 1.0) return (S > 0) ? (M > 0 ? 1.0 : 0.0) : (M < 0 ? 1.0 : 0.0);
  return S * 0.5 * (M + S);
}
template <int S>
__device__ __forceinline__ double dp_int(double M) {
  const double aM = M < 0 ? -M : M;
  if (aM > 1.0) return 0.0;
  const double mp = M + S;
  return 0.5 * mp * (2.0 - S * M) - S * 0.25 * (mp * mp);
}

// primitives() with one refined reciprocal of rho instead of nd + 1
// divisions (the reverse sweep differentiates the same function; its
// intermediates agree with the primal's to rounding)
template <int ND>
__device__ __forceinline__ Prim<ND, double> prim_rev_fwd(const double* q, double gamma, double gm1) {
  Prim<ND, double> w;
  w.rho = q[ND];
  const double ir = dev_rcp(w.rho);
#pragma unroll
  for (int d = 0; d < ND; ++d) w.u[d] = q[d] * ir;
  w.p = gm1 * (q[ND + 1] - 0.5 * w.rho * esum_sq<ND>(w.u));
  w.a = fast_sqrt(gamma * w.p * ir);
  return w;
}

// Adjoint of primitives(): (ub, pb, ab) -> qb (accumulated)
template <int ND>
__device__ __forceinline__ void primitives_rev(const double* q, const Prim<ND, double>& w, double gamma, double gm1,
                                               double* ub, double pb, double ab, double* qb) {
  const double irho = dev_rcp(w.rho);
  double rhob = 0.0;
  // a = sqrt(gamma p / rho)
  const double zb = ab * 0.5 * dev_rcp(w.a);
  pb += zb * gamma * irho;
  rhob -= zb * (gamma * w.p * irho) * irho;
  // p = (gamma - 1)(rho E - 0.5 rho |u|^2)
  qb[ND + 1] += gm1 * pb;
  const double s = esum_sq<ND>(w.u);
  rhob -= 0.5 * gm1 * s * pb;
  const double sb = -0.5 * gm1 * w.rho * pb;
#pragma unroll
  for (int k = 0; k < ND; ++k) {
    const double u_b = ub[k] + 2.0 * w.u[k] * sb;
    qb[k] += u_b * irho;
    rhob -= u_b * w.u[k] * irho;
  }
  qb[ND] += rhob;
}

// Adjoint of viscous_flux along D: Fb -> (qLb, pLb, qRb, pRb, Gb), accumulated
template <int ND, int D>
__device__ void viscous_flux_rev(const PhysConst& pc, const double* qL, double pL, const double* qR, double pR,
                                 const double (*G)[ND], const double* Fb, double* qLb, double& pLb, double* qRb,
                                 double& pRb, double (*Gb)[ND]) {
  constexpr int RHO = ND, E = ND + 1;
  const double gamma = pc.gamma, gm1 = pc.gm1;
  // forward intermediates
  const double rho_s = 0.5 * (qL[RHO] + qR[RHO]);
  const double irs = dev_rcp(rho_s);
  double rhou_s[ND], u_s[ND], umag2 = 0.0;
#pragma unroll
  for (int k = 0; k < ND; ++k) {
    rhou_s[k] = 0.5 * (qL[k] + qR[k]);
    u_s[k] = rhou_s[k] * irs;
    umag2 += u_s[k] * u_s[k];
  }
  const double p_s = gm1 * (0.5 * (qL[E] + qR[E]) - 0.5 * rho_s * umag2);
  const double irs2 = irs * irs;
  // only the velocity gradients the flux reads: row D, column D and the
  // diagonal (the tangential cross terms do not enter the face flux)
  auto used = [](int c, int sd) { return c == D || sd == D || c == sd; };
  double gu[ND][ND];
#pragma unroll
  for (int c = 0; c < ND; ++c)
#pragma unroll
    for (int sd = 0; sd < ND; ++sd)
      gu[c][sd] = used(c, sd) ? (G[c][sd] * rho_s - G[RHO][sd] * rhou_s[c]) * irs2 : 0.0;
  double grho_umag2 = G[RHO][D] * umag2;
#pragma unroll
  for (int k = 0; k < ND; ++k) grho_umag2 += rho_s * 2.0 * u_s[k] * gu[k][D];
  const double gradP = gm1 * (G[E][D] - 0.5 * grho_umag2);
  const double Tm = gamma * (0.5 * (pL + pR)) * irs;
  const double iTs = dev_rcp(Tm + pc.sOverTref);
  const double mu = Tm * fast_sqrt(Tm) * pc.sOverTrefP1 * iTs;
  const double K = -mu * pc.q_coef * irs;
  const double Y = gradP - p_s * irs * G[RHO][D];
  double div = 0.0;
#pragma unroll
  for (int k = 0; k < ND; ++k) div += gu[k][k];
  div = div * 2.0 / 3.0;
  const double iRe = pc.inv_Re0;
  // reverse
  double u_sb[ND], rhou_sb[ND], gub[ND][ND];
#pragma unroll
  for (int k = 0; k < ND; ++k) {
    u_sb[k] = 0.0;
    rhou_sb[k] = 0.0;
#pragma unroll
    for (int j = 0; j < ND; ++j) gub[k][j] = 0.0;
  }
  double mub = 0.0, rho_sb = 0.0, irsb = 0.0, divb = 0.0;
  const double fEb = iRe * Fb[E];
#pragma unroll
  for (int i = 0; i < ND; ++i) {
    const double Si = gu[D][i] + gu[i][D] - (i == D ? div : 0.0);
    const double tb = iRe * Fb[i] + fEb * u_s[i];
    u_sb[i] += fEb * (-mu * Si);
    mub -= tb * Si;
    const double Sb = -mu * tb;
    gub[D][i] += Sb;
    gub[i][D] += Sb;
    if (i == D) divb -= Sb;
  }
#pragma unroll
  for (int k = 0; k < ND; ++k) gub[k][k] += divb * (2.0 / 3.0);
  // heat flux q = K Y
  const double Kb = fEb * Y, Yb = fEb * K;
  mub += Kb * (-pc.q_coef * irs);
  rho_sb -= Kb * K * irs;
  double p_sb = -Yb * irs * G[RHO][D];
  irsb -= Yb * p_s * G[RHO][D];
  Gb[RHO][D] -= Yb * p_s * irs;
  // Sutherland viscosity of T = gamma avg(p) / avg(rho)
  const double Tmb = mub * mu * (1.5 * dev_rcp(Tm) - iTs);
  const double pab = Tmb * gamma * irs;
  rho_sb -= Tmb * Tm * irs;
  pLb += 0.5 * pab;
  pRb += 0.5 * pab;
  // gradP = (gamma - 1)(G[E][D] - 0.5 grho_umag2)
  Gb[E][D] += gm1 * Yb;
  const double grb = -0.5 * gm1 * Yb;
  Gb[RHO][D] += grb * umag2;
  double umag2b = grb * G[RHO][D];
#pragma unroll
  for (int k = 0; k < ND; ++k) {
    rho_sb += grb * 2.0 * u_s[k] * gu[k][D];
    u_sb[k] += grb * rho_s * 2.0 * gu[k][D];
    gub[k][D] += grb * rho_s * 2.0 * u_s[k];
  }
  // velocity gradients
  double irs2b = 0.0;
#pragma unroll
  for (int c = 0; c < ND; ++c)
#pragma unroll
    for (int sd = 0; sd < ND; ++sd) {
      if (!used(c, sd)) continue;
      const double g = gub[c][sd];
      Gb[c][sd] += g * rho_s * irs2;
      Gb[RHO][sd] -= g * rhou_s[c] * irs2;
      rho_sb += g * G[c][sd] * irs2;
      rhou_sb[c] -= g * G[RHO][sd] * irs2;
      irs2b += g * (G[c][sd] * rho_s - G[RHO][sd] * rhou_s[c]);
    }
  irsb += 2.0 * irs * irs2b;
  // p_s = (gamma - 1)(rhoE_s - 0.5 rho_s |u_s|^2)
  const double rhoE_sb = gm1 * p_sb;
  rho_sb -= 0.5 * gm1 * umag2 * p_sb;
  umag2b -= 0.5 * gm1 * rho_s * p_sb;
#pragma unroll
  for (int k = 0; k < ND; ++k) {
    u_sb[k] += 2.0 * u_s[k] * umag2b;
    rhou_sb[k] += u_sb[k] * irs;
    irsb += u_sb[k] * rhou_s[k];
  }
  rho_sb -= irsb * irs * irs;
  qLb[RHO] += 0.5 * rho_sb;
  qRb[RHO] += 0.5 * rho_sb;
  qLb[E] += 0.5 * rhoE_sb;
  qRb[E] += 0.5 * rhoE_sb;
#pragma unroll
  for (int k = 0; k < ND; ++k) {
    qLb[k] += 0.5 * rhou_sb[k];
    qRb[k] += 0.5 * rhou_sb[k];
  }
}

// Adjoint of flux_face: Fb -> qLb, qRb (set) and the tangential gradient
// adjoints Gtb[v][sd], sd != D (set).
template <int ND, int PHYS, int D>
__device__ void flux_face_rev(const PhysConst& pc, const double* qL, const double* qR, const double (*Gt)[ND],
                              double inv_len_r, const double* Fb, double* qLb, double* qRb, double (*Gtb)[ND]) {
  constexpr int NV = ND + 2, E = ND + 1;
  const double gamma = pc.gamma, gm1 = pc.gm1;
  const Prim<ND, double> wL = prim_rev_fwd<ND>(qL, gamma, gm1);
  const Prim<ND, double> wR = prim_rev_fwd<ND>(qR, gamma, gm1);
  const double iaL = dev_rcp(wL.a), iaR = dev_rcp(wR.a);
#pragma unroll
  for (int v = 0; v < NV; ++v) { qLb[v] = 0.0; qRb[v] = 0.0; }
  double uLb[ND], uRb[ND];
#pragma unroll
  for (int k = 0; k < ND; ++k) { uLb[k] = 0.0; uRb[k] = 0.0; }
  double pLb = 0.0, pRb = 0.0, aLb = 0.0, aRb = 0.0;
  // AUSM: f[k] = q_up[k] a_up m_i (k <= ND), f[E] = (q_up[E] + p_up) a_up m_i, f[D] += p_i
  const double ML = wL.u[D] * iaL, MR = wR.u[D] * iaR;
  const double m_i = m_int<+1>(ML) + m_int<-1>(MR);
  const bool left = m_i >= 0.0;
  const double* q = left ? qL : qR;
  double* qb = left ? qLb : qRb;
  const double ac = left ? wL.a : wR.a;
  const double pu = left ? wL.p : wR.p;
  double m_ib = 0.0, acb = 0.0;
#pragma unroll
  for (int k = 0; k < ND + 1; ++k) {
    m_ib += Fb[k] * (q[k] * ac);
    acb += Fb[k] * q[k] * m_i;
    qb[k] += Fb[k] * ac * m_i;
  }
  const double qe = q[E] + pu;
  m_ib += Fb[E] * (qe * ac);
  acb += Fb[E] * qe * m_i;
  const double qeb = Fb[E] * ac * m_i;
  qb[E] += qeb;
  if (left) { pLb += qeb; aLb += acb; } else { pRb += qeb; aRb += acb; }
  const double pib = Fb[D];
  pLb += pib * p_int<+1>(ML);
  pRb += pib * p_int<-1>(MR);
  const double MLb = pib * wL.p * dp_int<+1>(ML) + m_ib * dm_int<+1>(ML);
  const double MRb = pib * wR.p * dp_int<-1>(MR) + m_ib * dm_int<-1>(MR);
  uLb[D] += MLb * iaL;
  aLb -= MLb * ML * iaL;
  uRb[D] += MRb * iaR;
  aRb -= MRb * MR * iaR;
  if constexpr (PHYS == HOM3G_CNS) {
    double G[NV][ND], Gb[NV][ND];
#pragma unroll
    for (int v = 0; v < NV; ++v)
#pragma unroll
      for (int sd = 0; sd < ND; ++sd) {
        G[v][sd] = (sd == D) ? (qR[v] - qL[v]) * inv_len_r : Gt[v][sd];
        Gb[v][sd] = 0.0;
      }
    viscous_flux_rev<ND, D>(pc, qL, wL.p, qR, wR.p, G, Fb, qLb, pLb, qRb, pRb, Gb);
#pragma unroll
    for (int v = 0; v < NV; ++v) {
      qRb[v] += Gb[v][D] * inv_len_r;
      qLb[v] -= Gb[v][D] * inv_len_r;
#pragma unroll
      for (int sd = 0; sd < ND; ++sd) Gtb[v][sd] = (sd == D) ? 0.0 : Gb[v][sd];
    }
  }
  primitives_rev<ND>(qL, wL, gamma, gm1, uLb, pLb, aLb, qLb);
  primitives_rev<ND>(qR, wR, gamma, gm1, uRb, pRb, aRb, qRb);
}

// euler/quantities.hpp pv / cv in the order of the table-driven ghost fill
template <int ND, class T>
__device__ void pv_of_t(const T* cv, double gamma, T* pv) {
  pv[ND] = cv[ND];
  T vm2 = T(0);
  for (int d = 0; d < ND; ++d) { pv[d] = cv[d] / cv[ND]; vm2 += pv[d] * pv[d]; }
  pv[ND + 1] = (gamma - 1.0) * (cv[ND + 1] - 0.5 * cv[ND] * vm2);
}
template <int ND, class T>
__device__ void cv_of_t(const T* pv, double gamma, T* cv) {
  cv[ND] = pv[ND];
  T vm2 = T(0);
  for (int d = 0; d < ND; ++d) { cv[d] = pv[ND] * pv[d]; vm2 += pv[d] * pv[d]; }
  cv[ND + 1] = pv[ND + 1] / (gamma - 1.0) + 0.5 * pv[ND] * vm2;
}
template <int ND, class T>
__device__ T fluid_T_t(const T* cv, double gamma) {
  const T rho = cv[ND];
  T u[ND];
  for (int d = 0; d < ND; ++d) u[d] = cv[d] / rho;
  const T s = esum_sq<ND>(u);
  const T p = (gamma - 1.) * (cv[ND + 1] - 0.5 * rho * s);
  return gamma * p / rho;
}

// Ghost value of ghost g for its applied condition (the table-driven ghost
// fill, all kinds but the periodic image, which is a copy).  cvb: buffer at
// the boundary cell; own: this solver's lhs at the boundary cell (CHT);
// oth: the coupled solver's lhs at the partner cell (CHT).
template <int ND, class T>
__device__ void ghost_value(const AdjTopo& t, int bc, int g, const T* cvb, const T* own, const T* oth, T* out) {
  const int kind = t.kind[bc];
  const double* pp = t.params[bc];
  const int nv = t.nvars;
  switch (kind) {
    case HOM3G_BC_NEUMANN: {
      const double dx = t.g_dist[g];
      for (int v = 0; v < ND + 2; ++v) {
        if (v >= nv) break;
        const double gv = (t.nparams[bc] > v) ? pp[v] : 0.0;
        out[v] = cvb[v] - dx * gv;
      }
      return;
    }
    case HOM3G_BC_DIRICHLET:
      for (int v = 0; v < ND + 2; ++v)
        if (v < nv) out[v] = 2.0 * pp[v] - cvb[v];
      return;
    case HOM3G_BC_INFLOW: case HOM3G_BC_OUTFLOW: case HOM3G_BC_ADIABATIC_NOSLIP: case HOM3G_BC_ADIABATIC_SLIP:
    case HOM3G_BC_ISOTHERMAL_NOSLIP: case HOM3G_BC_CHT_FLUID: {
      constexpr int RHO = ND, P = ND + 1;
      T pvb[ND + 2], pvg[ND + 2];
      pv_of_t<ND>(cvb, t.gamma, pvb);
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
        const int normal = (t.g_pos[g] ^ 1) / 2;
        for (int d = 0; d < ND; ++d) pvg[d] = (d == normal) ? 0.0 - pvb[d] : pvb[d] - 0.0;
        pvg[RHO] = pvb[RHO] - 0.0;
        pvg[P] = pvb[P] - 0.0;
      } else {
        T Ts;
        if (kind == HOM3G_BC_CHT_FLUID) {
          const T Tf = fluid_T_t<ND>(own, t.gamma);
          const T Tsol = oth[0];
          Ts = (t.kratio[bc] == 1.0) ? 0.5 * (Tf + Tsol) : (Tf + t.kratio[bc] * Tsol) / (1.0 + t.kratio[bc]);
        } else {
          Ts = T(pp[0]);
        }
        for (int d = 0; d < ND; ++d) pvg[d] = 0.0 - pvb[d];
        const T rho_s = pvb[P] * t.gamma / Ts;
        pvg[RHO] = 2.0 * rho_s - pvb[RHO];
        pvg[P] = pvb[P] - 0.0;
      }
      cv_of_t<ND>(pvg, t.gamma, out);
      return;
    }
    case HOM3G_BC_HEAT_DIRICHLET:
      out[0] = 2.0 * pp[0] - cvb[0];
      return;
    case HOM3G_BC_CHT_SOLID: {
      const T Tsol = own[0];
      const T Tf = fluid_T_t<ND>(oth, t.other_gamma);
      const T Ts = (t.kratio[bc] == 1.0) ? 0.5 * (Tf + Tsol) : (Tf + t.kratio[bc] * Tsol) / (1.0 + t.kratio[bc]);
      out[0] = 2.0 * Ts - cvb[0];
      return;
    }
    case HOM3G_BC_HEAT_NEUMANN:
      out[0] = cvb[0] - t.g_dist[g] * pp[0];
      return;
    default:
      for (int v = 0; v < ND + 2; ++v)
        if (v < nv) out[v] = cvb[v];
      return;
  }
}

__device__ __forceinline__ bool uses_own(int kind) {
  return kind == HOM3G_BC_CHT_FLUID || kind == HOM3G_BC_CHT_SOLID;
}

// ---- tangent kernels ---------------------------------------------------------

template <int ND>
__global__ void __launch_bounds__(256) k_tan_ghosts(AdjTopo t, double* bv, double* bt, const double* ownv, const double* ownt,
                             const double* othv, const double* otht) {
  const long long g = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (g >= t.ng) return;
  const int bc = t.g_apply[g];
  if (bc < 0) return;
  const long long ld = t.ld, gc = t.n_int + g;
  const int nv = t.nvars;
  const int kind = t.kind[bc];
  if (kind == HOM3G_BC_PERIODIC) {
    const int im = t.g_image[g];
    for (int v = 0; v < nv; ++v) { bv[v * ld + gc] = bv[v * ld + im]; bt[v * ld + gc] = bt[v * ld + im]; }
    return;
  }
  const int b = t.g_bndry[g];
  Dual cvb[ND + 2], own[ND + 2], oth[ND + 2], out[ND + 2];
  for (int v = 0; v < nv; ++v) {
    cvb[v] = Dual(bv[v * ld + b], bt[v * ld + b]);
    own[v] = Dual(ownv[v * ld + b], ownt[v * ld + b]);
  }
  if (uses_own(kind)) {
    const int pc_ = t.g_partner[g];
    const int onv = (kind == HOM3G_BC_CHT_FLUID) ? 1 : ND + 2;
    for (int v = 0; v < onv; ++v) oth[v] = Dual(othv[v * t.other_ld + pc_], otht[v * t.other_ld + pc_]);
  }
  ghost_value<ND, Dual>(t, bc, (int)g, cvb, own, oth, out);
  for (int v = 0; v < nv; ++v) { bv[v * ld + gc] = out[v].v; bt[v * ld + gc] = out[v].d; }
}

template <int ND, int PHYS>
__device__ void face_dual(const AdjTopo& t, const PhysConst& pc, const double* Qv, const double* Qt, int l, int r,
                          int d, Dual* f) {
  constexpr int NV = ND + 2;
  Dual qL[NV], qR[NV], G[NV][ND];
  for (int v = 0; v < NV; ++v) {
    qL[v] = Dual(Qv[v * t.ld + l], Qt[v * t.ld + l]);
    qR[v] = Dual(Qv[v * t.ld + r], Qt[v * t.ld + r]);
  }
  if constexpr (PHYS == HOM3G_CNS) {
    for (int sd = 0; sd < ND; ++sd) {
      if (sd == d) continue;
      for (int v = 0; v < NV; ++v) {
        const SlopeT sr = slope_walk(t, r, v, sd), sl = slope_walk(t, l, v, sd);
        const Dual s_r = (sr.np < 0) ? Dual(__builtin_nan(""))
                                     : slope_value(sr, Dual(Qv[v * t.ld + sr.np], Qt[v * t.ld + sr.np]),
                                                   Dual(Qv[v * t.ld + sr.nn], Qt[v * t.ld + sr.nn]));
        const Dual s_l = (sl.np < 0) ? Dual(__builtin_nan(""))
                                     : slope_value(sl, Dual(Qv[v * t.ld + sl.np], Qt[v * t.ld + sl.np]),
                                                   Dual(Qv[v * t.ld + sl.nn], Qt[v * t.ld + sl.nn]));
        G[v][sd] = 0.5 * (s_r + s_l);
      }
    }
  }
  const double ilr = 1.0 / t.len[r];
  if (d == 0) flux_face<ND, PHYS, 0, Dual>(pc, qL, qR, G, ilr, f);
  else if (d == 1) flux_face<ND, PHYS, 1, Dual>(pc, qL, qR, G, ilr, f);
  else if constexpr (ND == 3) flux_face<ND, PHYS, 2, Dual>(pc, qL, qR, G, ilr, f);
}

struct TanStage {
  const double* inv; const double* intn;
  const double* basev; const double* baset;
  double* outv; double* outt;
  const double* src; const double* dsrc;
  int kind;
};

template <int ND, int PHYS>
__global__ void __launch_bounds__(256) k_tan_stage(AdjTopo t, PhysConst pc, TanStage a, const DevScalars* scal) {
  const long long c = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (c >= t.n_int) return;
  constexpr int NV = (PHYS == HOM3G_HEAT) ? 1 : ND + 2;
  const double dt = scal->dt;
  const double dx = t.len[c];
  Dual acc[NV];
  for (int v = 0; v < NV; ++v) acc[v] = Dual(0.0);
  for (int d = 0; d < ND; ++d) {
    const int m = t.nghbr[(2 * d) * t.ld + c];
    const int p = t.nghbr[(2 * d + 1) * t.ld + c];
    if (m < 0 || p < 0) {
      for (int v = 0; v < NV; ++v) acc[v] = Dual(__builtin_nan(""));
      continue;
    }
    if constexpr (PHYS == HOM3G_HEAT) {
      const Dual Tc(a.inv[c], a.intn[c]);
      const Dual fm = 1 / dx * (Dual(a.inv[m], a.intn[m]) - Tc);
      const Dual fp = 1 / dx * (Tc - Dual(a.inv[p], a.intn[p]));
      acc[0] += (fm - fp);
    } else {
      Dual fm[NV], fp[NV];
      face_dual<ND, PHYS>(t, pc, a.inv, a.intn, m, (int)c, d, fm);
      face_dual<ND, PHYS>(t, pc, a.inv, a.intn, (int)c, p, d, fp);
      for (int v = 0; v < NV; ++v) acc[v] += (fm[v] - fp[v]);
    }
  }
  const Dual s = (a.src != nullptr) ? Dual(dt * a.src[c], dt * a.dsrc[c]) : Dual(0.0);
  for (int v = 0; v < NV; ++v) {
    const Dual nf = dt / dx * acc[v];
    const Dual sv = (v == 0 && PHYS == HOM3G_HEAT) ? s : Dual(0.0);
    const Dual q(a.inv[v * t.ld + c], a.intn[v * t.ld + c]);
    Dual o;
    switch (a.kind) {
      case 0: o = q + nf + sv; break;
      case 2: o = 0.75 * Dual(a.basev[v * t.ld + c], a.baset[v * t.ld + c]) + 0.25 * (q + nf + sv); break;
      case 3: o = (1. / 3.) * Dual(a.basev[v * t.ld + c], a.baset[v * t.ld + c]) + (2. / 3.) * (q + nf + sv); break;
      default: o = q + nf; break;
    }
    a.outv[v * t.ld + c] = o.v;
    a.outt[v * t.ld + c] = o.d;
  }
}

// ---- adjoint kernels ---------------------------------------------------------

// One face (L, R, d) per thread: faces (m, c) of every internal cell c and
// faces (c, p) towards a ghost p.  Fbar = w dt/len_R obar_R [R internal]
// - w dt/len_L obar_L [L internal]; its pull-back through the face flux
// (flux_face_rev) is scattered into Ibar (cells, ghosts, slope stencils).
template <int ND, int PHYS>
__global__ void __launch_bounds__(128) __attribute__((amdgpu_waves_per_eu(2))) k_adj_faces(AdjTopo t, PhysConst pc, const double* Q, const double* obar, double w, double dt,
                            double* Ibar) {
  constexpr int NV = ND + 2;
  // grid.y = 2 d + side: a wave handles one kind of face for consecutive
  // cells (coalesced state reads; the plus faces, mostly internal and
  // skipped, do not idle half of every wave)
  const long long c = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (c >= t.n_int) return;
  const int d = (int)(blockIdx.y >> 1);
  const int side = (int)(blockIdx.y & 1);
  int L, R;
  if (side == 0) {
    L = t.nghbr[(2 * d) * t.ld + c];
    R = (int)c;
  } else {
    L = (int)c;
    R = t.nghbr[(2 * d + 1) * t.ld + c];
    if (R < t.n_int) return;  // an internal face, handled from R's side
  }
  if (L < 0 || R < 0) return;
  double Fb[NV];
  bool any = false;
  for (int v = 0; v < NV; ++v) {
    double fb = 0;
    if (R < t.n_int) fb += w * (dt / t.len[R]) * obar[v * t.ld + R];
    if (L < t.n_int) fb -= w * (dt / t.len[L]) * obar[v * t.ld + L];
    Fb[v] = fb;
    any = any || fb != 0.0;
  }
  if (!any) return;
  double qL[NV], qR[NV], Gt[NV][ND];
  for (int v = 0; v < NV; ++v) {
    qL[v] = Q[v * t.ld + L];
    qR[v] = Q[v * t.ld + R];
  }
  // slope walks per tangential direction: the stencil does not depend on
  // the variable; wall hooks flip the sign of the momentum slopes only.
  // (Walked again for the scatter below rather than kept in registers.)
  if constexpr (PHYS == HOM3G_CNS) {
    for (int sd = 0; sd < ND; ++sd) {
      if (sd == d) continue;
      const SlopeT sr = slope_walk(t, R, 0, sd), sl = slope_walk(t, L, 0, sd);
      for (int v = 0; v < NV; ++v) {
        SlopeT a = sr, b = sl;
        if (v >= ND) { a.sign = 1.0; b.sign = 1.0; }
        const double ga = a.np < 0 ? __builtin_nan("") : slope_value(a, Q[v * t.ld + a.np], Q[v * t.ld + a.nn]);
        const double gb = b.np < 0 ? __builtin_nan("") : slope_value(b, Q[v * t.ld + b.np], Q[v * t.ld + b.nn]);
        Gt[v][sd] = 0.5 * (ga + gb);
      }
    }
  }
  const double ilr = 1.0 / t.len[R];
  double qLb[NV], qRb[NV], Gtb[NV][ND];
  if (d == 0) flux_face_rev<ND, PHYS, 0>(pc, qL, qR, Gt, ilr, Fb, qLb, qRb, Gtb);
  else if (d == 1) flux_face_rev<ND, PHYS, 1>(pc, qL, qR, Gt, ilr, Fb, qLb, qRb, Gtb);
  else if constexpr (ND == 3) flux_face_rev<ND, PHYS, 2>(pc, qL, qR, Gt, ilr, Fb, qLb, qRb, Gtb);
  for (int v = 0; v < NV; ++v) {
    if (qLb[v] != 0.0) atomicAdd(&Ibar[v * t.ld + L], qLb[v]);
    if (qRb[v] != 0.0) atomicAdd(&Ibar[v * t.ld + R], qRb[v]);
  }
  if constexpr (PHYS == HOM3G_CNS) {
    // G = 0.5 (slope(R) + slope(L)), slope = sign (Q[np] - Q[nn]) / len2
    for (int sd = 0; sd < ND; ++sd) {
      if (sd == d) continue;
      for (int h = 0; h < 2; ++h) {
        const SlopeT s = slope_walk(t, h == 0 ? R : L, 0, sd);
        if (s.np < 0) continue;
        for (int v = 0; v < NV; ++v) {
          const double xb = Gtb[v][sd];
          if (xb == 0.0) continue;
          const double sg = (v < ND && s.sign < 0) ? -1.0 : 1.0;
          const double cf = 0.5 * xb * sg * s.ilen2;
          atomicAdd(&Ibar[v * t.ld + s.np], cf);
          atomicAdd(&Ibar[v * t.ld + s.nn], -cf);
        }
      }
    }
  }
}

// heat: acc = sum_d (1/dx)(T_m - T_c) - (1/dx)(T_c - T_p), linear
__global__ void k_adj_heat(AdjTopo t, const double* obar, double w, double dt, double* Ibar) {
  const long long c = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (c >= t.n_int) return;
  const double dx = t.len[c];
  const double ab = w * (dt / dx) * obar[c];
  if (ab == 0.0) return;
  const double k = (1 / dx) * ab;
  for (int d = 0; d < t.nd; ++d) {
    const int m = t.nghbr[(2 * d) * t.ld + c];
    const int p = t.nghbr[(2 * d + 1) * t.ld + c];
    if (m < 0 || p < 0) continue;
    atomicAdd(&Ibar[m], k);
    atomicAdd(&Ibar[p], k);
    atomicAdd(&Ibar[c], -2.0 * k);
  }
}

// direct terms of out = w_base base + w (in + nf + dt S)
__global__ void k_adj_cell(long long n_int, long long ld, int nv, const double* obar, double w, double w_base,
                           double dt, double* Ibar, double* Ubar, double* Sbar) {
  const long long c = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (c >= n_int) return;
  for (int v = 0; v < nv; ++v) {
    const double o = obar[v * ld + c];
    Ibar[v * ld + c] += w * o;
    if (w_base != 0.0) Ubar[v * ld + c] += w_base * o;
  }
  if (Sbar) Sbar[c] += w * dt * obar[c];
}

// pull-back of the ghost fill: ghost adjoints (Ibar ghost slots) into the
// boundary cell of the buffer, this solver's lhs (CHT) and the coupled
// solver (CHT partner).  blockIdx.y = k selects the input whose adjoint a
// thread forms (k < nv: buffer, then own, then partner variables), so the
// forward-mode sweeps of one ghost run in parallel threads instead of in
// sequence (the ghosts alone leave most SIMDs idle)
template <int ND>
__global__ void __launch_bounds__(256) k_adj_ghosts(AdjTopo t, const double* bufv, const double* ownv, const double* othv, double* Ibar,
                             double* Ownbar, double* Othbar) {
  const long long g = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  const int k = blockIdx.y;
  if (g >= t.ng) return;
  const int bc = t.g_apply[g];
  if (bc < 0) return;
  const long long ld = t.ld, gc = t.n_int + g;
  const int nv = t.nvars;
  double gb[ND + 2];
  bool any = false;
  for (int v = 0; v < ND + 2; ++v) {
    gb[v] = (v < nv) ? Ibar[v * ld + gc] : 0.0;
    any = any || gb[v] != 0.0;
  }
  if (!any) return;
  const int kind = t.kind[bc];
  if (kind == HOM3G_BC_PERIODIC) {
    const int im = t.g_image[g];
    if (k == 0)
      for (int v = 0; v < ND + 2; ++v)
        if (v < nv) atomicAdd(&Ibar[v * ld + im], gb[v]);
    return;
  }
  const bool cht = uses_own(kind);
  const int onv = (kind == HOM3G_BC_CHT_FLUID) ? 1 : ND + 2;
  const int nin = nv + (cht ? nv + onv : 0);
  if (k >= nin) return;
  const int b = t.g_bndry[g];
  const int pc_ = cht ? t.g_partner[g] : 0;
  double cvb0[ND + 2], own0[ND + 2], oth0[ND + 2];
  for (int v = 0; v < ND + 2; ++v) {
    cvb0[v] = (v < nv) ? bufv[v * ld + b] : 0.0;
    own0[v] = (v < nv) ? ownv[v * ld + b] : 0.0;
  }
  for (int v = 0; v < ND + 2; ++v) oth0[v] = (cht && v < onv) ? othv[v * t.other_ld + pc_] : 0.0;
  {
    Dual cvb[ND + 2], own[ND + 2], oth[ND + 2], out[ND + 2];
    for (int v = 0; v < ND + 2; ++v) {
      cvb[v] = Dual(cvb0[v], (v < nv && k == v) ? 1.0 : 0.0);
      own[v] = Dual(own0[v], (v < nv && k == nv + v) ? 1.0 : 0.0);
      oth[v] = Dual(oth0[v], (v < onv && k == 2 * nv + v) ? 1.0 : 0.0);
    }
    ghost_value<ND, Dual>(t, bc, (int)g, cvb, own, oth, out);
    double xb = 0;
    for (int v = 0; v < ND + 2; ++v)
      if (v < nv) xb += gb[v] * out[v].d;
    if (xb == 0.0) return;
    if (k < nv) atomicAdd(&Ibar[k * ld + b], xb);
    else if (k < 2 * nv) atomicAdd(&Ownbar[(k - nv) * ld + b], xb);
    else atomicAdd(&Othbar[(k - 2 * nv) * t.other_ld + pc_], xb);
  }
}

// wall heat flux Q of the solid (its CHT ghosts), times `scale`, into *acc;
// with tangents: *dacc receives the derivative
template <int ND>
__global__ void k_objective(AdjTopo th, const double* Hv, const double* Ht, const double* Cv, const double* Ct,
                            const DevScalars* hscal, double* acc, double* dacc) {
  const long long g = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (g >= th.ng) return;
  const int bc = th.g_apply[g];
  if (bc < 0 || th.kind[bc] != HOM3G_BC_CHT_SOLID) return;
  const int b = th.g_bndry[g], pc_ = th.g_partner[g];
  const Dual Tb(Hv[b], Ht ? Ht[b] : 0.0);
  Dual cv[ND + 2];
  for (int v = 0; v < ND + 2; ++v) cv[v] = Dual(Cv[v * th.other_ld + pc_], Ct ? Ct[v * th.other_ld + pc_] : 0.0);
  const Dual Tf = fluid_T_t<ND>(cv, th.other_gamma);
  const double kr = th.kratio[bc];
  const Dual Ts = (kr == 1.0) ? 0.5 * (Tf + Tb) : (Tf + kr * Tb) / (1.0 + kr);
  const double h = th.len[b];
  const double wgt = (ND == 2) ? 1.0 : h;
  const Dual q = wgt * (2 * (Tb - Ts));
  const double dt = hscal->dt;
  atomicAdd(acc, dt * q.v);
  if (dacc) atomicAdd(dacc, dt * q.d);
}

template <int ND>
__global__ void k_objective_adj(AdjTopo th, const double* Hv, const double* Cv, double scale, double* Hbar,
                                double* Cbar) {
  const long long g = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (g >= th.ng) return;
  const int bc = th.g_apply[g];
  if (bc < 0 || th.kind[bc] != HOM3G_BC_CHT_SOLID) return;
  const int b = th.g_bndry[g], pc_ = th.g_partner[g];
  const double kr = th.kratio[bc];
  const double h = th.len[b];
  const double wgt = (ND == 2) ? 1.0 : h;
  for (int k = 0; k < ND + 3; ++k) {
    const Dual Tb(Hv[b], k == 0 ? 1.0 : 0.0);
    Dual cv[ND + 2];
    for (int v = 0; v < ND + 2; ++v) cv[v] = Dual(Cv[v * th.other_ld + pc_], k == 1 + v ? 1.0 : 0.0);
    const Dual Tf = fluid_T_t<ND>(cv, th.other_gamma);
    const Dual Ts = (kr == 1.0) ? 0.5 * (Tf + Tb) : (Tf + kr * Tb) / (1.0 + kr);
    const Dual q = wgt * (2 * (Tb - Ts));
    const double xb = scale * q.d;
    if (k == 0) atomicAdd(&Hbar[b], xb);
    else atomicAdd(&Cbar[(k - 1) * th.other_ld + pc_], xb);
  }
}

// final-state misfit of the solid: acc += 1/2 sum_c (T_c - obs_c)^2 over the
// internal cells; with a tangent, dacc += sum_c (T_c - obs_c) dT_c
__global__ void __launch_bounds__(256) k_misfit(long long n, const double* T, const double* Tt, const double* obs,
                                                double* acc, double* dacc) {
  __shared__ double red[2][4];
  double a = 0.0, d = 0.0;
  for (long long c = blockIdx.x * (long long)blockDim.x + threadIdx.x; c < n; c += (long long)gridDim.x * blockDim.x) {
    const double r = T[c] - obs[c];
    a += 0.5 * r * r;
    if (Tt) d += r * Tt[c];
  }
  for (int off = 32; off > 0; off >>= 1) {
    a += __shfl_xor(a, off, 64);
    d += __shfl_xor(d, off, 64);
  }
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) { red[0][w] = a; red[1][w] = d; }
  __syncthreads();
  if (threadIdx.x == 0) {
    atomicAdd(acc, (red[0][0] + red[0][1]) + (red[0][2] + red[0][3]));
    if (dacc) atomicAdd(dacc, (red[1][0] + red[1][1]) + (red[1][2] + red[1][3]));
  }
}

// adjoint seed of the misfit: Tbar_c += T_c - obs_c
__global__ void k_misfit_seed(long long n, const double* T, const double* obs, double* Tbar) {
  const long long c = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (c < n) Tbar[c] += T[c] - obs[c];
}

__global__ void k_zero(long long n, double* a) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    a[i] = 0.0;
}

__global__ void k_copyd(long long n, const double* a, double* b) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    b[i] = a[i];
}

// coupling::apply with a prescribed fluid dt (forced on both solvers)
__global__ void k_force_pair(DevScalars* cs, DevScalars* hs, double dt, double scaling) {
  cs->forced_step = cs->step;
  cs->forced_dt = dt;
  hs->forced_step = hs->step;
  hs->forced_dt = dt * scaling;
  cs->dtmin = dbl_to_ordered(1.7976931348623157e308);
  hs->dtmin = dbl_to_ordered(1.7976931348623157e308);
}

// set_dt (solver.hpp:541-551) and advance (:502-505) for the tangent run
__global__ void k_set_dt_a(DevScalars* s) {
  const double dtm = ordered_to_dbl(s->dtmin);
  if (s->time + s->dt < s->tEnd) s->dt = (s->step == s->forced_step) ? s->forced_dt : dtm;
  else s->dt = s->tEnd - s->time;
  s->dtmin = dbl_to_ordered(1.7976931348623157e308);
}
__global__ void k_advance_a(DevScalars* s) {
  s->time += s->dt;
  s->step += 1;
}

void zero(Ctx& C, double* p, long long n) {
  hipLaunchKernelGGL(k_zero, dim3(std::min<long long>(4096, (n + 255) / 256 + 1)), dim3(256), 0, C.stream, n, p);
  H3_LAUNCH_CHECK();
}
void copyd(Ctx& C, const double* a, double* b, long long n) {
  hipLaunchKernelGGL(k_copyd, dim3(std::min<long long>(4096, (n + 255) / 256 + 1)), dim3(256), 0, C.stream, n, a, b);
  H3_LAUNCH_CHECK();
}

void check_pair(SolverDev& cns, SolverDev& heat) {
  H3_REQUIRE(cns.P.physics == HOM3G_CNS && heat.P.physics == HOM3G_HEAT, "needs a CNS and a heat solver");
  H3_REQUIRE(cns.kernel == HOM3G_KERNEL_GENERIC && heat.kernel == HOM3G_KERNEL_GENERIC,
             "the adjoint runs on the table-driven FP64 kernels");
  for (SolverDev* S : {&cns, &heat}) {
    H3_REQUIRE(S->P.ti == HOM3G_TI_EULER_FWD || S->P.ti == HOM3G_TI_SSP_RK3,
               "the adjoint supports forward Euler and SSP-RK3");
    H3_REQUIRE(S->P.recon == HOM3G_RECON_NONE && S->P.flux != HOM3G_FLUX_LLF, "the adjoint supports AUSM without reconstruction");
    H3_REQUIRE(!S->nonuniform, "the adjoint supports uniform grids");
    for (const BCDesc& b : S->bcs)
      H3_REQUIRE(b.kind != HOM3G_BC_ISENTROPIC_VORTEX, "the adjoint does not support the isentropic vortex boundary");
  }
  H3_REQUIRE(cns.P.ti == heat.P.ti, "both solvers must use the same time integration");
}

double scaling_of(double Tr0, const SolverDev& cns) { return Tr0 * cns.pc.Pr0 * cns.pc.Re0; }

template <class F>
void by_nd(int nd, F f) {
  if (nd == 2) f(std::integral_constant<int, 2>());
  else f(std::integral_constant<int, 3>());
}

void objective_launch(Ctx& C, SolverDev& heat, SolverDev& cns, const AdjTopo& th, const double* Ht,
                      const double* Ct, double* acc, double* dacc) {
  if (heat.n_ghost == 0) return;
  by_nd(heat.nd, [&](auto ndc) {
    constexpr int ND = decltype(ndc)::value;
    hipLaunchKernelGGL(k_objective<ND>, dim3(nblk(heat.n_ghost)), dim3(256), 0, C.stream, th,
                       (const double*)heat.lhs.p, Ht, (const double*)cns.lhs.p, Ct, (const DevScalars*)heat.scal.p,
                       acc, dacc);
  });
  H3_LAUNCH_CHECK();
}

void misfit_launch(Ctx& C, SolverDev& heat, const double* obs, const double* Ht, double* acc, double* dacc) {
  const long long n = heat.n_int;
  const unsigned grid = (unsigned)std::min<long long>(nblk(n), 1024);
  hipLaunchKernelGGL(k_misfit, dim3(grid), dim3(256), 0, C.stream, n, (const double*)heat.lhs.p, Ht, obs, acc, dacc);
  H3_LAUNCH_CHECK();
}

// observations of the misfit objective on the device (empty: wall heat flux)
struct Obs {
  DevBuf<double> buf;
  const double* p = nullptr;
  Obs(Ctx& C, const SolverDev& heat, const double* host) {
    if (!host) return;
    buf.alloc(heat.cap);
    zero(C, buf.p, heat.cap);
    buf.upload(host, heat.n_int, C.stream);
    p = buf.p;
  }
};

// ---- tangent step ------------------------------------------------------------
struct TanBufs {
  DevBuf<double> L, R, S;
};

void tan_ghosts(Ctx& C, SolverDev& S, const AdjTopo& t, double* bv, double* bt, const double* ownv,
                const double* ownt, const double* othv, const double* otht) {
  if (S.n_ghost == 0) return;
  by_nd(S.nd, [&](auto ndc) {
    constexpr int ND = decltype(ndc)::value;
    hipLaunchKernelGGL(k_tan_ghosts<ND>, dim3(nblk(S.n_ghost)), dim3(256), 0, C.stream, t, bv, bt, ownv, ownt,
                       othv, otht);
  });
  H3_LAUNCH_CHECK();
}

void tan_stage(Ctx& C, SolverDev& S, const AdjTopo& t, const TanStage& a) {
  by_nd(S.nd, [&](auto ndc) {
    constexpr int ND = decltype(ndc)::value;
    if (S.P.physics == HOM3G_CNS)
      hipLaunchKernelGGL((k_tan_stage<ND, HOM3G_CNS>), dim3(nblk(S.n_int)), dim3(256), 0, C.stream, t, S.pc, a,
                         (const DevScalars*)S.scal.p);
    else if (S.P.physics == HOM3G_EULER)
      hipLaunchKernelGGL((k_tan_stage<ND, HOM3G_EULER>), dim3(nblk(S.n_int)), dim3(256), 0, C.stream, t, S.pc, a,
                         (const DevScalars*)S.scal.p);
    else
      hipLaunchKernelGGL((k_tan_stage<ND, HOM3G_HEAT>), dim3(nblk(S.n_int)), dim3(256), 0, C.stream, t, S.pc, a,
                         (const DevScalars*)S.scal.p);
  });
  H3_LAUNCH_CHECK();
}

void tan_step(Ctx& C, SolverDev& S, const AdjTopo& t, TanBufs& T, const double* othv, const double* otht,
              const double* src, const double* dsrc) {
  hipLaunchKernelGGL(k_set_dt_a, dim3(1), dim3(1), 0, C.stream, S.scal.p);
  H3_LAUNCH_CHECK();
  const long long nv = S.nvars, ld = S.cap;
  tan_ghosts(C, S, t, S.lhs.p, T.L.p, S.lhs.p, T.L.p, othv, otht);
  if (S.P.ti == HOM3G_TI_SSP_RK3) {
    tan_stage(C, S, t, TanStage{S.lhs.p, T.L.p, S.lhs.p, T.L.p, S.rhs.p, T.R.p, src, dsrc, 0});
    tan_ghosts(C, S, t, S.rhs.p, T.R.p, S.lhs.p, T.L.p, othv, otht);
    tan_stage(C, S, t, TanStage{S.rhs.p, T.R.p, S.lhs.p, T.L.p, S.stg.p, T.S.p, src, dsrc, 2});
    tan_ghosts(C, S, t, S.stg.p, T.S.p, S.lhs.p, T.L.p, othv, otht);
    tan_stage(C, S, t, TanStage{S.stg.p, T.S.p, S.lhs.p, T.L.p, S.lhs.p, T.L.p, src, dsrc, 3});
  } else {
    tan_stage(C, S, t, TanStage{S.lhs.p, T.L.p, S.lhs.p, T.L.p, S.rhs.p, T.R.p, src, dsrc, 0});
    for (long long v = 0; v < nv && S.n_ghost > 0; ++v) {
      copyd(C, S.lhs.p + v * ld + S.n_int, S.rhs.p + v * ld + S.n_int, S.n_ghost);
      copyd(C, T.L.p + v * ld + S.n_int, T.R.p + v * ld + S.n_int, S.n_ghost);
    }
    std::swap(S.lhs.p, S.rhs.p);
    std::swap(T.L.p, T.R.p);
  }
  hipLaunchKernelGGL(k_advance_a, dim3(1), dim3(1), 0, C.stream, S.scal.p);
  H3_LAUNCH_CHECK();
}

// ---- reverse step of one solver ------------------------------------------------
struct RevArgs {
  const double* B0;  // U^n with its ghosts
  const double* B1;  // stage 1 with ghosts (SSP-RK3)
  const double* B2;  // stage 2 with ghosts (SSP-RK3)
  const double* oth; // coupled solver's lhs as read by the ghost fills
  double* Unext;     // adjoint of U^{n+1} (internal)
  double* Ucur;      // adjoint of U^n (zeroed by the caller)
  double* Othbar;    // adjoint of the coupled solver's state
  double* Sbar;      // adjoint of the source (heat) or null
  double* I;         // work: stage input adjoint
  double* O;         // work: stage output adjoint
  double dt;
};

void faces_rev(Ctx& C, SolverDev& S, const AdjTopo& t, const double* Q, const double* obar, double w, double dt,
               double* Ibar) {
  by_nd(S.nd, [&](auto ndc) {
    constexpr int ND = decltype(ndc)::value;
    if (S.P.physics == HOM3G_HEAT) {
      hipLaunchKernelGGL(k_adj_heat, dim3(nblk(S.n_int)), dim3(256), 0, C.stream, t, obar, w, dt, Ibar);
    } else {
      const dim3 grid(nblk(S.n_int, 128), 2 * ND);
      if (S.P.physics == HOM3G_CNS)
        hipLaunchKernelGGL((k_adj_faces<ND, HOM3G_CNS>), grid, dim3(128), 0, C.stream, t, S.pc, Q, obar, w, dt,
                           Ibar);
      else
        hipLaunchKernelGGL((k_adj_faces<ND, HOM3G_EULER>), grid, dim3(128), 0, C.stream, t, S.pc, Q, obar, w, dt,
                           Ibar);
    }
  });
  H3_LAUNCH_CHECK();
}

void cell_rev(Ctx& C, SolverDev& S, const double* obar, double w, double wb, double dt, double* Ibar, double* Ubar,
              double* Sbar) {
  hipLaunchKernelGGL(k_adj_cell, dim3(nblk(S.n_int)), dim3(256), 0, C.stream, (long long)S.n_int, (long long)S.cap,
                     S.nvars, obar, w, wb, dt, Ibar, Ubar, Sbar);
  H3_LAUNCH_CHECK();
}

void ghosts_rev(Ctx& C, SolverDev& S, const AdjTopo& t, const double* buf, const double* own, const double* oth,
                double* Ibar, double* Ownbar, double* Othbar) {
  if (S.n_ghost == 0) return;
  by_nd(S.nd, [&](auto ndc) {
    constexpr int ND = decltype(ndc)::value;
    hipLaunchKernelGGL(k_adj_ghosts<ND>, dim3(nblk(S.n_ghost), 3 * (ND + 2)), dim3(256), 0, C.stream, t, buf, own, oth, Ibar,
                       Ownbar, Othbar);
  });
  H3_LAUNCH_CHECK();
}

void rev_solver(Ctx& C, SolverDev& S, const AdjTopo& t, const RevArgs& r) {
  const long long n = (long long)S.nvars * S.cap;
  if (S.P.ti == HOM3G_TI_SSP_RK3) {
    zero(C, r.I, n);
    faces_rev(C, S, t, r.B2, r.Unext, 2. / 3., r.dt, r.I);
    cell_rev(C, S, r.Unext, 2. / 3., 1. / 3., r.dt, r.I, r.Ucur, r.Sbar);
    ghosts_rev(C, S, t, r.B2, r.B0, r.oth, r.I, r.Ucur, r.Othbar);
    copyd(C, r.I, r.O, n);
    zero(C, r.I, n);
    faces_rev(C, S, t, r.B1, r.O, 0.25, r.dt, r.I);
    cell_rev(C, S, r.O, 0.25, 0.75, r.dt, r.I, r.Ucur, r.Sbar);
    ghosts_rev(C, S, t, r.B1, r.B0, r.oth, r.I, r.Ucur, r.Othbar);
    copyd(C, r.I, r.O, n);
    faces_rev(C, S, t, r.B0, r.O, 1.0, r.dt, r.Ucur);
    cell_rev(C, S, r.O, 1.0, 0.0, r.dt, r.Ucur, r.Ucur, r.Sbar);
    ghosts_rev(C, S, t, r.B0, r.B0, r.oth, r.Ucur, r.Ucur, r.Othbar);
  } else {
    faces_rev(C, S, t, r.B0, r.Unext, 1.0, r.dt, r.Ucur);
    cell_rev(C, S, r.Unext, 1.0, 0.0, r.dt, r.Ucur, r.Ucur, r.Sbar);
    ghosts_rev(C, S, t, r.B0, r.B0, r.oth, r.Ucur, r.Ucur, r.Othbar);
  }
}

}  // namespace

// ---- drivers -----------------------------------------------------------------

void cht_objective_run(Ctx& C, SolverDev& cns, SolverDev& heat, double Tr0, int nsteps, const double* dts,
                       double* dts_out, ChtRunStats& st, const double* obs) {
  check_pair(cns, heat);
  const double scaling = scaling_of(Tr0, cns);
  const AdjTopo th = make_adj_topo(heat, &cns);
  const Obs ob(C, heat, obs);
  DevBuf<double> acc, dtrec;
  acc.alloc(1);
  zero(C, acc.p, 1);
  std::vector<double> rec;
  H3_HIP(hipEventRecord(C.ev0, C.stream));
  for (int n = 0; n < nsteps; ++n) {
    if (dts) {
      hipLaunchKernelGGL(k_force_pair, dim3(1), dim3(1), 0, C.stream, cns.scal.p, heat.scal.p, dts[n], scaling);
      H3_LAUNCH_CHECK();
      cns.dtmin_current = heat.dtmin_current = false;
      solver_step(C, cns, true);
      solver_step(C, heat, true);
    } else {
      couple_cht_step(C, cns, heat, Tr0);
    }
    if (!ob.p) objective_launch(C, heat, cns, th, nullptr, nullptr, acc.p, nullptr);
    if (dts_out) {
      DevScalars s;
      H3_HIP(hipMemcpyAsync(&s, cns.scal.p, sizeof(DevScalars), hipMemcpyDeviceToHost, C.stream));
      H3_HIP(hipStreamSynchronize(C.stream));
      dts_out[n] = s.dt;
    }
  }
  if (ob.p) misfit_launch(C, heat, ob.p, nullptr, acc.p, nullptr);
  H3_HIP(hipEventRecord(C.ev1, C.stream));
  H3_HIP(hipEventSynchronize(C.ev1));
  H3_HIP(hipEventElapsedTime(&st.ms_forward, C.ev0, C.ev1));
  st.ms_primal_step = nsteps ? st.ms_forward / nsteps : 0.f;
  acc.download(&st.J, 1, C.stream);
  H3_HIP(hipStreamSynchronize(C.stream));
}

void cht_tangent_run(Ctx& C, SolverDev& cns, SolverDev& heat, double Tr0, int nsteps, const double* dts,
                     const double* dsource, double* J, double* dJ, const double* obs) {
  check_pair(cns, heat);
  H3_REQUIRE(dts != nullptr && dsource != nullptr, "the tangent needs a dt schedule and a source direction");
  const double scaling = scaling_of(Tr0, cns);
  const AdjTopo tc = make_adj_topo(cns, &heat), th = make_adj_topo(heat, &cns);
  TanBufs Tc, Th;
  for (DevBuf<double>* b : {&Tc.L, &Tc.R, &Tc.S}) { b->alloc((size_t)cns.nvars * cns.cap); zero(C, b->p, b->n); }
  for (DevBuf<double>* b : {&Th.L, &Th.R, &Th.S}) { b->alloc((size_t)heat.nvars * heat.cap); zero(C, b->p, b->n); }
  DevBuf<double> ds, zsrc, acc;
  ds.alloc(heat.cap);
  zero(C, ds.p, heat.cap);
  ds.upload(dsource, heat.n_int, C.stream);
  const double* src = heat.source.p;
  if (!src) {
    zsrc.alloc(heat.cap);
    zero(C, zsrc.p, heat.cap);
    src = zsrc.p;
  }
  acc.alloc(2);
  zero(C, acc.p, 2);
  const Obs ob(C, heat, obs);
  for (int n = 0; n < nsteps; ++n) {
    hipLaunchKernelGGL(k_force_pair, dim3(1), dim3(1), 0, C.stream, cns.scal.p, heat.scal.p, dts[n], scaling);
    H3_LAUNCH_CHECK();
    tan_step(C, cns, tc, Tc, heat.lhs.p, Th.L.p, nullptr, nullptr);
    tan_step(C, heat, th, Th, cns.lhs.p, Tc.L.p, src, ds.p);
    if (!ob.p) objective_launch(C, heat, cns, th, Th.L.p, Tc.L.p, acc.p, acc.p + 1);
  }
  if (ob.p) misfit_launch(C, heat, ob.p, Th.L.p, acc.p, acc.p + 1);
  cns.dtmin_current = heat.dtmin_current = false;
  double h[2];
  acc.download(h, 2, C.stream);
  H3_HIP(hipStreamSynchronize(C.stream));
  *J = h[0];
  *dJ = h[1];
}

void cht_adjoint_run(Ctx& C, SolverDev& cns, SolverDev& heat, double Tr0, int nsteps, const double* dts,
                     int ckpt_every, double* grad_source, ChtRunStats& st, const double* obs) {
  check_pair(cns, heat);
  H3_REQUIRE(dts != nullptr, "the adjoint needs a dt schedule");
  H3_REQUIRE(ckpt_every >= 0, "ckpt_every must be 0 (store every step) or positive");
  const double scaling = scaling_of(Tr0, cns);
  const AdjTopo tc = make_adj_topo(cns, &heat), th = make_adj_topo(heat, &cns);
  const bool rk3 = cns.P.ti == HOM3G_TI_SSP_RK3;
  const long long nc = (long long)cns.nvars * cns.cap, nh = (long long)heat.nvars * heat.cap;
  // ckpt_every == 0: the forward sweep keeps every step's stage inputs in
  // HBM and the reverse sweep recomputes nothing
  const bool store_all = ckpt_every == 0;
  const int K = store_all ? std::max(nsteps, 1) : ckpt_every, nseg = (nsteps + K - 1) / K;
  const int KS = std::min(K, nsteps);
  const int per_step = rk3 ? 3 : 1;
  if (store_all) {
    size_t free_b = 0, total_b = 0;
    H3_HIP(hipMemGetInfo(&free_b, &total_b));
    const double need = (double)KS * per_step * (double)(nc + nh) * sizeof(double);
    if (need > 0.9 * (double)free_b) {
      char msg[256];
      snprintf(msg, sizeof msg, "storing every step needs %.1f GiB of device memory, %.1f GiB free; use checkpoints",
               need / (1 << 30), free_b / double(1 << 30));
      throw Error(HOM3G_ERR_CAPACITY, msg);
    }
  }
  std::vector<DevBuf<double>> C0(KS), C1(rk3 ? KS : 0), C2(rk3 ? KS : 0), H0(KS), H1(rk3 ? KS : 0), H2(rk3 ? KS : 0);
  for (int j = 0; j < KS; ++j) {
    C0[j].alloc(nc);
    H0[j].alloc(nh);
    if (rk3) { C1[j].alloc(nc); C2[j].alloc(nc); H1[j].alloc(nh); H2[j].alloc(nh); }
  }
  // one coupled step that keeps the stage inputs of both solvers in slot j
  // (the ghost fills are those solver_step performs; repeating them changes
  // nothing)
  auto record_step = [&](int n, int j) {
    hipLaunchKernelGGL(k_force_pair, dim3(1), dim3(1), 0, C.stream, cns.scal.p, heat.scal.p, dts[n], scaling);
    H3_LAUNCH_CHECK();
    solver_fill_ghosts(C, cns, cns.lhs.p);
    copyd(C, cns.lhs.p, C0[j].p, nc);
    solver_step(C, cns, true);
    if (rk3) { copyd(C, cns.rhs.p, C1[j].p, nc); copyd(C, cns.stg.p, C2[j].p, nc); }
    solver_fill_ghosts(C, heat, heat.lhs.p);
    copyd(C, heat.lhs.p, H0[j].p, nh);
    solver_step(C, heat, true);
    if (rk3) { copyd(C, heat.rhs.p, H1[j].p, nh); copyd(C, heat.stg.p, H2[j].p, nh); }
  };

  // forward sweep: checkpoints every K steps (in device memory, or in
  // pinned host memory when asked or when device memory is short), or every
  // stage input
  bool host_ck = false;
  if (!store_all) {
    size_t free_b = 0, total_b = 0;
    H3_HIP(hipMemGetInfo(&free_b, &total_b));
    const double need = (double)nseg * (double)(nc + nh) * sizeof(double);
    host_ck = C.ckpt_memory == HOM3G_CKPT_HOST || (C.ckpt_memory == HOM3G_CKPT_AUTO && need > 0.8 * (double)free_b);
  }
  std::vector<DevBuf<double>> ckC(store_all || host_ck ? 0 : nseg), ckH(store_all || host_ck ? 0 : nseg);
  std::vector<PinnedBuf<double>> hkC(host_ck ? nseg : 0), hkH(host_ck ? nseg : 0);
  DevBuf<double> acc;
  acc.alloc(1);
  zero(C, acc.p, 1);
  const Obs ob(C, heat, obs);
  H3_HIP(hipEventRecord(C.ev0, C.stream));
  for (int n = 0; n < nsteps; ++n) {
    if (store_all) {
      record_step(n, n);
    } else {
      if (n % K == 0) {
        if (host_ck) {
          hkC[n / K].alloc(nc);
          hkH[n / K].alloc(nh);
          H3_HIP(hipMemcpyAsync(hkC[n / K].p, cns.lhs.p, nc * sizeof(double), hipMemcpyDeviceToHost, C.stream));
          H3_HIP(hipMemcpyAsync(hkH[n / K].p, heat.lhs.p, nh * sizeof(double), hipMemcpyDeviceToHost, C.stream));
        } else {
          ckC[n / K].alloc(nc);
          ckH[n / K].alloc(nh);
          copyd(C, cns.lhs.p, ckC[n / K].p, nc);
          copyd(C, heat.lhs.p, ckH[n / K].p, nh);
        }
      }
      hipLaunchKernelGGL(k_force_pair, dim3(1), dim3(1), 0, C.stream, cns.scal.p, heat.scal.p, dts[n], scaling);
      H3_LAUNCH_CHECK();
      solver_step(C, cns, true);
      solver_step(C, heat, true);
    }
    if (!ob.p) objective_launch(C, heat, cns, th, nullptr, nullptr, acc.p, nullptr);
  }
  if (ob.p) misfit_launch(C, heat, ob.p, nullptr, acc.p, nullptr);
  H3_HIP(hipEventRecord(C.ev1, C.stream));
  H3_HIP(hipEventSynchronize(C.ev1));
  H3_HIP(hipEventElapsedTime(&st.ms_forward, C.ev0, C.ev1));
  st.ms_primal_step = nsteps ? st.ms_forward / nsteps : 0.f;
  acc.download(&st.J, 1, C.stream);
  // the solvers end at step N; keep their final state and scalars
  DevBuf<double> finC, finH;
  finC.alloc(nc);
  finH.alloc(nh);
  copyd(C, cns.lhs.p, finC.p, nc);
  copyd(C, heat.lhs.p, finH.p, nh);
  DevScalars fsc, fsh;
  H3_HIP(hipMemcpyAsync(&fsc, cns.scal.p, sizeof(DevScalars), hipMemcpyDeviceToHost, C.stream));
  H3_HIP(hipMemcpyAsync(&fsh, heat.scal.p, sizeof(DevScalars), hipMemcpyDeviceToHost, C.stream));
  H3_HIP(hipStreamSynchronize(C.stream));

  // reverse sweep
  DevBuf<double> Cn1, Cn, Hn1, Hn, Ic, Oc, Ih, Oh, Sbar, Cend, Hend;
  for (DevBuf<double>* b : {&Cn1, &Cn, &Ic, &Oc}) b->alloc(nc);
  for (DevBuf<double>* b : {&Hn1, &Hn, &Ih, &Oh, &Sbar}) b->alloc(nh);
  if (!store_all) {
    Cend.alloc(nc);
    Hend.alloc(nh);
  }
  const double* CendP = store_all ? finC.p : Cend.p;
  const double* HendP = store_all ? finH.p : Hend.p;
  zero(C, Cn1.p, nc);
  zero(C, Hn1.p, nh);
  zero(C, Sbar.p, nh);
  if (ob.p) {
    // misfit: the adjoint of the final solid state is T_N - obs
    hipLaunchKernelGGL(k_misfit_seed, dim3(nblk(heat.n_int)), dim3(256), 0, C.stream, (long long)heat.n_int,
                       (const double*)finH.p, ob.p, Hn1.p);
    H3_LAUNCH_CHECK();
  }
  H3_HIP(hipEventRecord(C.ev0, C.stream));
  for (int sg = nseg - 1; sg >= 0; --sg) {
    const int n0 = sg * K, n1 = std::min(nsteps, n0 + K), len = n1 - n0;
    if (!store_all) {
      if (host_ck) {
        H3_HIP(hipMemcpyAsync(cns.lhs.p, hkC[sg].p, nc * sizeof(double), hipMemcpyHostToDevice, C.stream));
        H3_HIP(hipMemcpyAsync(heat.lhs.p, hkH[sg].p, nh * sizeof(double), hipMemcpyHostToDevice, C.stream));
      } else {
        copyd(C, ckC[sg].p, cns.lhs.p, nc);
        copyd(C, ckH[sg].p, heat.lhs.p, nh);
      }
      for (int j = 0; j < len; ++j) record_step(n0 + j, j);
      copyd(C, cns.lhs.p, Cend.p, nc);
      copyd(C, heat.lhs.p, Hend.p, nh);
    }
    for (int j = len - 1; j >= 0; --j) {
      const int n = n0 + j;
      const double dtc = dts[n], dth = dts[n] * scaling;
      const double* Cnext = (j + 1 < len) ? C0[j + 1].p : CendP;
      const double* Hnext = (j + 1 < len) ? H0[j + 1].p : HendP;
      if (heat.n_ghost > 0 && !ob.p) {
        by_nd(heat.nd, [&](auto ndc) {
          constexpr int ND = decltype(ndc)::value;
          hipLaunchKernelGGL(k_objective_adj<ND>, dim3(nblk(heat.n_ghost)), dim3(256), 0, C.stream, th, Hnext, Cnext,
                             dth, Hn1.p, Cn1.p);
        });
        H3_LAUNCH_CHECK();
      }
      zero(C, Hn.p, nh);
      RevArgs rh{H0[j].p, rk3 ? H1[j].p : nullptr, rk3 ? H2[j].p : nullptr, Cnext, Hn1.p, Hn.p, Cn1.p, Sbar.p,
                 Ih.p, Oh.p, dth};
      rev_solver(C, heat, th, rh);
      zero(C, Cn.p, nc);
      RevArgs rc{C0[j].p, rk3 ? C1[j].p : nullptr, rk3 ? C2[j].p : nullptr, H0[j].p, Cn1.p, Cn.p, Hn.p, nullptr,
                 Ic.p, Oc.p, dtc};
      rev_solver(C, cns, tc, rc);
      std::swap(Cn1.p, Cn.p);
      std::swap(Hn1.p, Hn.p);
    }
  }
  H3_HIP(hipEventRecord(C.ev1, C.stream));
  H3_HIP(hipEventSynchronize(C.ev1));
  H3_HIP(hipEventElapsedTime(&st.ms_reverse, C.ev0, C.ev1));
  if (grad_source) Sbar.download(grad_source, heat.n_int, C.stream);
  copyd(C, finC.p, cns.lhs.p, nc);
  copyd(C, finH.p, heat.lhs.p, nh);
  H3_HIP(hipMemcpyAsync(cns.scal.p, &fsc, sizeof(DevScalars), hipMemcpyHostToDevice, C.stream));
  H3_HIP(hipMemcpyAsync(heat.scal.p, &fsh, sizeof(DevScalars), hipMemcpyHostToDevice, C.stream));
  H3_HIP(hipStreamSynchronize(C.stream));
  cns.dtmin_current = heat.dtmin_current = false;
}

}  // namespace hom3g
