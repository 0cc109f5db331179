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
// reverse of the flux, to the cells, to the slope stencils
// (solver.hpp:596-678, including the boundary hooks) and through the ghost
// fills (boundary_condition.hpp, cns/ and heat/ boundary and coupling
// conditions) to the boundary cells and the coupled solver.
//
// Determinism: no adjoint is accumulated with atomics.  Every face, ghost
// and objective term writes its contributions to slots of its own; per-cell
// gather kernels then sum the slots that reach a cell in a fixed order
// (faces by direction, slope stencils and ghost pull-backs through index
// lists built once from the topology, sorted by source).  Sums over the
// objective's terms are ordered single-block reductions.  Two runs give
// bit-identical gradients, whatever the checkpoint spacing or placement.
#include <algorithm>
#include <functional>
#include <cstdio>
#include <cstring>

#include "brick_geo.hpp"
#include "dual.hpp"
#include "face_rev.hpp"
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
__device__ __forceinline__ void pv_of_t(const T* cv, double gamma, T* pv) {
  pv[ND] = cv[ND];
  T vm2 = T(0);
  for (int d = 0; d < ND; ++d) { pv[d] = cv[d] / cv[ND]; vm2 += pv[d] * pv[d]; }
  pv[ND + 1] = (gamma - 1.0) * (cv[ND + 1] - 0.5 * cv[ND] * vm2);
}
template <int ND, class T>
__device__ __forceinline__ void cv_of_t(const T* pv, double gamma, T* cv) {
  cv[ND] = pv[ND];
  T vm2 = T(0);
  for (int d = 0; d < ND; ++d) { cv[d] = pv[ND] * pv[d]; vm2 += pv[d] * pv[d]; }
  cv[ND + 1] = pv[ND + 1] / (gamma - 1.0) + 0.5 * pv[ND] * vm2;
}
template <int ND, class T>
__device__ __forceinline__ T fluid_T_t(const T* cv, double gamma) {
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
__device__ __forceinline__ void ghost_value(const AdjTopo& t, int bc, int g, const T* cvb, const T* own, const T* oth, T* out) {
  const int kind = t.kind[bc];
  const double* pp = t.params[bc];
  const int nv = t.nvars;
  switch (kind) {
    case HOM3G_BC_NEUMANN: {
      const double dx = t.g_dist[g];
#pragma unroll
      for (int v = 0; v < ND + 2; ++v) {
        if (v < nv) {
          const double gv = (t.nparams[bc] > v) ? pp[v] : 0.0;
          out[v] = cvb[v] - dx * gv;
        }
      }
      return;
    }
    case HOM3G_BC_DIRICHLET:
#pragma unroll
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
#pragma unroll
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

// ---- tiled-brick reverse of the face fluxes ------------------------------------
//
// The reverse of a stage's face fluxes for a solver on tiled bricks (the
// forward's halo table): one workgroup per brick, the stage input of the
// brick and its halo in LDS.  Each face is reversed once, with its flux
// adjoint Fbar = w dt/h (obar_R - obar_L) (internal cells only): the faces
// below the brick's cells by the brick, a face between two bricks by the
// upper one, a face towards a ghost above the brick by the brick.  Per
// direction every face below a brick cell is reversed by one thread
// (brick_face_rev, the reverse of the forward brick's own face arithmetic)
// into face slots in LDS (the adjoints reaching L and those of the face's
// tangential difference sums; the adjoint reaching R, the thread's own
// cell, is added to its accumulator directly), and every target (the
// thread's cell and its halo entry) gathers from the slots of the faces
// that read it, in a fixed order; the faces towards ghosts above the brick
// follow in one more pass.  The
// brick's cells go to Ibar, its halo cells to per-entry slots, stored in
// the order of the cells they read: k_adj_tiled_gather adds them to their
// internal cells (or periodic images) and the ghost pull-back sums them for
// the ghosts, in entry order.  No atomics: the sums do not depend on
// scheduling.  The
// slope hook of a ghost face (the ghost's tangential slopes are its boundary
// cell's, velocities negated at walls) routes its tangential adjoint to the
// boundary cell's stencil only, with twice the weight (zero for wall
// momenta).

struct AdjBrickArgs {
  const double* Q;       // stage input, ghosts filled
  const double* obar;    // adjoint of the stage output (internal cells)
  double* Ibar;          // += the brick's cells
  double* HS;            // halo entry slots: hpos[brick * NH + j] * NV + v
  const int32_t* hpos;   // the slot of each halo entry, in the order of the cells they read (-1: not read)
  const uint32_t* halo;  // the forward's halo table
  long long ld, n_int;
  double fbs;            // w dt / h
  double inv_h;
  double w, wb;          // direct terms: Ibar += w obar, Ubar += wb obar (Ubar null: not here)
  double* Ubar;
  int assign;            // Ibar holds no earlier terms at the brick's cells: written, not added to
  int ubar_assign;       // the same for Ubar
};

template <int ND>
struct AdjBrickSmem {
  using G = Geo<ND, HOM3G_CNS, false>;
  static constexpr int NT = BrickCfg<ND>::NT;
  // slots per direction: the faces below the brick's cells in a main pass
  // (their R adjoints go straight to the cells' accumulators), the ND TANG
  // upper-boundary faces in the tail pass
  static constexpr int NF = NT;
  static constexpr int NTAIL = ND * G::TANG;
  static_assert(NTAIL <= NF, "tail faces fit the slots");
  double q[G::NV][G::NPOS];
  double p[G::NPOS], ia[G::NPOS];  // p and 1/(rho a), as the forward forms them
  // face slots in two sets, pass P using set P % 2 and the tail pass set
  // ND % 2, so a pass writes its slots while slower waves still gather from
  // the previous pass's: the adjoints reaching L, those of the face's
  // tangential difference sums of the variables the flux reads (j = 0 the
  // momentum along the face normal, 1 the momentum along the tangential
  // direction, 2 rho), and for the tail faces the adjoints reaching R
  struct {
    double fl[2][G::NV][NF];
    double fg[2][3][ND - 1][NF];
    double fr[G::NV][NTAIL];
  } f;
  unsigned long long gmask[(2 * ND * G::TANG + 63) / 64][2];
};

template <int ND>
__device__ __forceinline__ int adj_gbits(const AdjBrickSmem<ND>& s, int j) {
  const unsigned long long b = 1ull << (j & 63);
  return ((s.gmask[j >> 6][0] & b) ? 1 : 0) | ((s.gmask[j >> 6][1] & b) ? 2 : 0);
}

template <int ND>
__device__ __forceinline__ int lds_stride(int d) {
  using G = Geo<ND, HOM3G_CNS, false>;
  return d == 0 ? 1 : (d == 1 ? G::E : G::E * G::E);
}

// tangential index of brick position y across direction d (the face-halo order)
template <int ND>
__device__ __forceinline__ int tang_of(const int* y, int d) {
  constexpr int B = BrickCfg<ND>::B;
  if constexpr (ND == 3) return d == 0 ? y[1] + B * y[2] : (d == 1 ? y[0] + B * y[2] : y[0] + B * y[1]);
  else return d == 0 ? y[1] : y[0];
}

// the variable of face-slot tangential entry j for tangential direction sd
// of a face along D (brick_face_rev's order)
template <int ND>
__device__ __forceinline__ constexpr int tan_var(int j, int D, int sd) {
  return j == 0 ? D : (j == 1 ? sd : ND);
}

// reverse of the face between LDS positions iR - stride(D) and iR (the
// forward's brick_face); hook as there (bit 0: L ghost, bit 1: R ghost,
// bit 2: wall hook)
template <int ND, int D>
__device__ __forceinline__ void adj_brick_face(const AdjBrickSmem<ND>& s, const PhysConst& pc, int iR, int hook, const double* Fb,
                               double inv_h, double* qLb, double* qRb, double (*Gtb)[ND - 1]) {
  constexpr int NV = ND + 2;
  const int iL = iR - lds_stride<ND>(D);
  double qL[NV], qR[NV], Gt[3][ND - 1];
#pragma unroll
  for (int v = 0; v < NV; ++v) {
    qL[v] = s.q[v][iL];
    qR[v] = s.q[v][iR];
  }
  // tangential difference sums; a ghost face takes twice its boundary
  // cell's difference (zero for wall momenta)
#pragma unroll
  for (int t = 0; t < ND - 1; ++t) {
    const int sd = t < D ? t : t + 1;
    const int st = lds_stride<ND>(sd);
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      const int v = tan_var<ND>(j, D, sd);
      const double dR = s.q[v][iR + st] - s.q[v][iR - st];
      const double dL = s.q[v][iL + st] - s.q[v][iL - st];
      double g = dR + dL;
      if (hook) {
        const double d = (hook & 1) ? dR : dL;
        g = ((hook & 4) && j < 2) ? 0.0 : d + d;
      }
      Gt[j][t] = g;
    }
  }
  brick_face_rev<ND, D>(pc, qL, qR, s.p[iL], s.p[iR], s.ia[iL], s.ia[iR], Gt, inv_h, Fb, qLb, qRb, Gtb);
}

// Morton bits of a brick coordinate v (0 <= v < B): bit i of v to bit ND i
template <int ND>
__device__ __forceinline__ unsigned spread_small(unsigned v) {
  if constexpr (ND == 3) return (v & 1u) | ((v & 2u) << 2) | ((v & 4u) << 4);
  else return (v & 1u) | ((v & 2u) << 1) | ((v & 4u) << 2) | ((v & 8u) << 3);
}

// ghost bits of the halo side of the face below brick position y along d:
// its L in a main pass (y_d = 0), its R in the tail pass
template <int ND, bool TAIL, int d>
__device__ __forceinline__ int adj_face_gbits(const AdjBrickSmem<ND>& s, const int* y) {
  constexpr int TANG = Geo<ND, HOM3G_CNS, false>::TANG;
  if (TAIL) return adj_gbits<ND>(s, (2 * d + 1) * TANG + tang_of<ND>(y, d));
  return y[d] == 0 ? adj_gbits<ND>(s, 2 * d * TANG + tang_of<ND>(y, d)) : 0;
}

// the weight of a stencil cell of the R side (side 0) or L side (side 1) in
// the tangential difference sum of a face with ghost bits gb: 1; at a ghost
// face 2 on the boundary cell's side (0 for wall momenta) and 0 on the ghost's
template <int ND, bool TAIL>
__device__ __forceinline__ double adj_tan_coef(int gb, int side, int v) {
  if (!(gb & 1)) return 1.0;
  const int ghost_side = TAIL ? 0 : 1;
  if (side == ghost_side) return 0.0;
  return ((gb & 2) && v < ND) ? 0.0 : 2.0;
}

// a[v] += what target x (brick coordinates, possibly in the halo) receives
// from the faces along d in the slots: the face below x (x is R), the face
// above x (x is L), and the faces whose R or L has x as a tangential
// neighbour.  Slots: in a main pass the faces below the brick's cells
// (slot = Morton index), in the tail pass the faces below the positions
// just above the brick (slot = d TANG + tangential index).  A face is
// addressed by the offsets o[k] - 1 of the position above it from x.
template <int ND, bool TAIL, int d>
__device__ __forceinline__ void adj_gather(const AdjBrickSmem<ND>& s, const int* x, double* a) {
  constexpr int NV = ND + 2;
  constexpr int b = TAIL ? ND % 2 : d % 2;  // the pass's slot set
  constexpr int B = BrickCfg<ND>::B;
  constexpr int TANG = Geo<ND, HOM3G_CNS, false>::TANG;
  // per coordinate k and offset o - 1: the Morton bits and whether the
  // coordinate lies in the brick
  unsigned P[3][3];
  bool in[3][3];
#pragma unroll
  for (int k = 0; k < 3; ++k)
#pragma unroll
    for (int o = 0; o < 3; ++o) {
      const int v = (k < ND ? x[k] : 0) + o - 1;
      in[k][o] = k >= ND || (v >= 0 && v < B);
      P[k][o] = k < ND ? spread_small<ND>((unsigned)v & (unsigned)(B - 1)) << k : 0u;
    }
  // slot of the face below y = x + o - 1, or -1
  auto slot = [&](const int* o, const int* y) -> int {
    bool ok = true;
    if constexpr (TAIL) {
#pragma unroll
      for (int k = 0; k < ND; ++k) ok = ok && (k == d ? y[k] == B : in[k][o[k]]);
      return ok ? d * TANG + tang_of<ND>(y, d) : -1;
    } else {
      unsigned m = 0u;
#pragma unroll
      for (int k = 0; k < ND; ++k) {
        ok = ok && in[k][o[k]];
        m |= P[k][o[k]];
      }
      return ok ? (int)m : -1;
    }
  };
  {
    int o[3] = {1, 1, 1};
    int y[3] = {x[0], x[1], ND == 3 ? x[2] : 0};
    // the face below x: in a main pass x is a brick cell whose thread took
    // that face's R adjoint already; in the tail pass x may be its halo R
    int f = TAIL ? slot(o, y) : -1;
    if (f >= 0) {
#pragma unroll
      for (int v = 0; v < NV; ++v) a[v] += s.f.fr[v][f];
    }
    o[d] = 2;
    y[d] += 1;
    f = slot(o, y);
    if (f >= 0) {
#pragma unroll
      for (int v = 0; v < NV; ++v) a[v] += s.f.fl[b][v][f];
    }
  }
#pragma unroll
  for (int sd = 0; sd < ND; ++sd) {
    if (sd == d) continue;
    const int ti = sd < d ? sd : sd - 1;
#pragma unroll
    for (int sg = 0; sg < 2; ++sg) {
      const int sgn = sg == 0 ? 1 : -1;
      // x = R + sgn e_sd: the face below R = x - sgn e_sd
      int o[3] = {1, 1, 1};
      int z[3] = {x[0], x[1], ND == 3 ? x[2] : 0};
      o[sd] = 1 - sgn;
      z[sd] -= sgn;
      int f = slot(o, z);
      if (f >= 0) {
        const int gb = adj_face_gbits<ND, TAIL, d>(s, z);
#pragma unroll
        for (int j = 0; j < 3; ++j) {
          const int v = tan_var<ND>(j, d, sd);
          a[v] += sgn * adj_tan_coef<ND, TAIL>(gb, 0, v) * s.f.fg[b][j][ti][f];
        }
      }
      // x = L + sgn e_sd: the face above L, below L + e_d
      o[d] = 2;
      z[d] += 1;
      f = slot(o, z);
      if (f >= 0) {
        const int gb = adj_face_gbits<ND, TAIL, d>(s, z);
#pragma unroll
        for (int j = 0; j < 3; ++j) {
          const int v = tan_var<ND>(j, d, sd);
          a[v] += sgn * adj_tan_coef<ND, TAIL>(gb, 1, v) * s.f.fg[b][j][ti][f];
        }
      }
    }
  }
}

// slot f of set b (WITH_R: a tail face, whose R adjoint is gathered too)
template <int ND, bool WITH_R>
__device__ __forceinline__ void adj_store_face(AdjBrickSmem<ND>& s, int b, int f, const double* qLb,
                                               const double* qRb, const double (*Gtb)[ND - 1]) {
#pragma unroll
  for (int v = 0; v < ND + 2; ++v) {
    s.f.fl[b][v][f] = qLb[v];
    if (WITH_R) s.f.fr[v][f] = qRb[v];
  }
#pragma unroll
  for (int j = 0; j < 3; ++j)
#pragma unroll
    for (int t = 0; t < ND - 1; ++t) s.f.fg[b][j][t][f] = Gtb[j][t];
}

// one main pass: the faces below the brick's cells along D.  Fbar from
// obar of the cell (ob) and of its lower neighbour (a brick cell, or the
// face-halo entry's cell: 0 for a ghost)
template <int ND, int D>
__device__ __forceinline__ void adj_brick_pass(AdjBrickSmem<ND>& s, const AdjBrickArgs& A, const PhysConst& pc, long long brick,
                               const int* cell, const double* ob, bool has_h, const int* hx, double* acc,
                               double* acch) {
  constexpr int NV = ND + 2;
  using G = Geo<ND, HOM3G_CNS, false>;
  using HL = HaloList<ND, HOM3G_CNS, false>;
  const int t = threadIdx.x;
  {
    // a face on the brick's lower side carries both cells' adjoints (its L,
    // a cell of the brick below, is that brick's upper-boundary face, which
    // the brick below leaves to this one; a ghost L carries none)
    int hook = 0;
    double Fb[NV];
    if (cell[D] == 0) {
      const int j = 2 * D * G::TANG + tang_of<ND>(cell, D);
      const int gb = adj_gbits<ND>(s, j);
      hook = (gb & 1) ? (1 | ((gb & 2) ? 4 : 0)) : 0;
      const long long iL = (long long)(A.halo[brick * HL::N + j] & 0x3fffffffu);
#pragma unroll
      for (int v = 0; v < NV; ++v) {
        const double oL = iL < A.n_int ? A.obar[v * A.ld + iL] : 0.0;
        Fb[v] = A.fbs * ob[v] - A.fbs * oL;
      }
    } else {
      int y[3] = {cell[0], cell[1], cell[2]};
      y[D] -= 1;
      const long long iL = brick * (long long)BrickCfg<ND>::NT + morton<ND>(y);
#pragma unroll
      for (int v = 0; v < NV; ++v) Fb[v] = A.fbs * ob[v] - A.fbs * A.obar[v * A.ld + iL];
    }
    const double inv_h = A.inv_h;
    double qLb[NV], qRb[NV], Gtb[3][ND - 1];
    bool any = false;
#pragma unroll
    for (int v = 0; v < NV; ++v) any = any || Fb[v] != 0.0;
    if (any) {
      adj_brick_face<ND, D>(s, pc, G::lin(cell), hook, Fb, inv_h, qLb, qRb, Gtb);
    } else {
#pragma unroll
      for (int v = 0; v < NV; ++v) qLb[v] = qRb[v] = 0.0;
#pragma unroll
      for (int j = 0; j < 3; ++j)
#pragma unroll
        for (int t = 0; t < ND - 1; ++t) Gtb[j][t] = 0.0;
    }
    // R is this thread's cell: its adjoint is the first term the cell
    // receives from this pass, as in the gather order
    adj_store_face<ND, false>(s, D % 2, t, qLb, qRb, Gtb);
#pragma unroll
    for (int v = 0; v < NV; ++v) acc[v] += qRb[v];
  }
  __syncthreads();
  // no barrier after the gather: the next pass writes the other slot set,
  // and the one after it comes after the next pass's barrier
  adj_gather<ND, false, D>(s, cell, acc);
  if (has_h) adj_gather<ND, false, D>(s, hx, acch);
}

template <int ND, int D>
__device__ __forceinline__ void adj_tail_face(AdjBrickSmem<ND>& s, const AdjBrickArgs& A, const PhysConst& pc,
                                              long long brick, int tg) {
  constexpr int NV = ND + 2;
  using G = Geo<ND, HOM3G_CNS, false>;
  constexpr int B = G::B;
  const double inv_h = A.inv_h;
  int y[3] = {0, 0, 0};
  if constexpr (ND == 3) {
    const int ta = tg % B, tb = tg / B;
    if (D == 0) { y[0] = B; y[1] = ta; y[2] = tb; }
    else if (D == 1) { y[0] = ta; y[1] = B; y[2] = tb; }
    else { y[0] = ta; y[1] = tb; y[2] = B; }
  } else {
    if (D == 0) { y[0] = B; y[1] = tg; } else { y[0] = tg; y[1] = B; }
  }
  const int gb = adj_gbits<ND>(s, (2 * D + 1) * G::TANG + tg);
  const int hook = (gb & 1) ? (2 | ((gb & 2) ? 4 : 0)) : 0;
  // only a face whose R is a ghost is reversed here (the adjoint of L, the
  // brick's top cell below it); a face towards a cell of the next brick (or
  // its periodic image) is that brick's lower-side face
  int yl[3] = {y[0], y[1], y[2]};
  yl[D] -= 1;
  const long long iL = brick * (long long)BrickCfg<ND>::NT + morton<ND>(yl);
  double Fb[NV];
#pragma unroll
  for (int v = 0; v < NV; ++v) Fb[v] = (gb & 1) ? 0.0 - A.fbs * A.obar[v * A.ld + iL] : 0.0;
  double qLb[NV], qRb[NV], Gtb[3][ND - 1];
  bool any = false;
#pragma unroll
  for (int v = 0; v < NV; ++v) any = any || Fb[v] != 0.0;
  if (any) {
    adj_brick_face<ND, D>(s, pc, G::lin(y), hook, Fb, inv_h, qLb, qRb, Gtb);
  } else {
#pragma unroll
    for (int v = 0; v < NV; ++v) qLb[v] = qRb[v] = 0.0;
#pragma unroll
    for (int j = 0; j < 3; ++j)
#pragma unroll
      for (int t = 0; t < ND - 1; ++t) Gtb[j][t] = 0.0;
  }
  adj_store_face<ND, true>(s, ND % 2, D * G::TANG + tg, qLb, qRb, Gtb);
}

template <int ND>
__global__ void __launch_bounds__(BrickCfg<ND>::NT) k_adj_brick(AdjBrickArgs A, PhysConst pc) {
  using G = Geo<ND, HOM3G_CNS, false>;
  using HL = HaloList<ND, HOM3G_CNS, false>;
  constexpr int NV = G::NV, NT = BrickCfg<ND>::NT, NH = HL::N, TANG = G::TANG;
  __shared__ AdjBrickSmem<ND> s;
  const int t = threadIdx.x;
  const long long nb = (long long)gridDim.x;
  long long brick = blockIdx.x;
  if ((nb & 7) == 0) brick = (long long)(blockIdx.x & 7) * (nb >> 3) + (blockIdx.x >> 3);
  int cell[3] = {0, 0, 0};
  demorton<ND>((unsigned)t, cell);
  const long long gc = brick * (long long)NT + t;
  const int ic = G::lin(cell);
  // stage input and output adjoint of the brick and its halo
  const bool has_h = t < NH;
  const int jh = has_h ? t : NH - 1;
  int hx[3] = {0, 0, 0};
  halo_closed_form<ND, HOM3G_CNS, false>(jh, hx);
  const int ih = G::lin(hx);
  const unsigned e = A.halo[brick * NH + jh];
  const int hp = has_h ? A.hpos[brick * NH + jh] : -1;
  const long long gh = (long long)(e & 0x3fffffffu);
  const unsigned code = e >> 30;
  {
    const bool face = t < 2 * ND * TANG;
    const unsigned long long g0 = __ballot(face && (code & 1u) != 0);
    const unsigned long long g1 = __ballot(face && code == 3u);
    if ((t & 63) == 0 && (t >> 6) < (2 * ND * TANG + 63) / 64) {
      s.gmask[t >> 6][0] = g0;
      s.gmask[t >> 6][1] = g1;
    }
  }
  double ob[NV];
  {
    double qc[NV], qh[NV];
#pragma unroll
    for (int v = 0; v < NV; ++v) {
      qc[v] = A.Q[v * A.ld + gc];
      ob[v] = A.obar[v * A.ld + gc];
      qh[v] = has_h ? A.Q[v * A.ld + gh] : 1.0;
    }
    double pv, iav;
#pragma unroll
    for (int v = 0; v < NV; ++v) s.q[v][ic] = qc[v];
    brick_prims<ND, double>(qc, pc.gamma, pc.gm1, pv, iav);
    s.p[ic] = pv;
    s.ia[ic] = iav;
    if (has_h) {
#pragma unroll
      for (int v = 0; v < NV; ++v) s.q[v][ih] = qh[v];
      brick_prims<ND, double>(qh, pc.gamma, pc.gm1, pv, iav);
      s.p[ih] = pv;
      s.ia[ih] = iav;
    }
  }
  __syncthreads();
  double acc[NV], acch[NV];
#pragma unroll
  for (int v = 0; v < NV; ++v) acc[v] = acch[v] = 0.0;
  adj_brick_pass<ND, 0>(s, A, pc, brick, cell, ob, has_h, hx, acc, acch);
  adj_brick_pass<ND, 1>(s, A, pc, brick, cell, ob, has_h, hx, acc, acch);
  if constexpr (ND == 3) adj_brick_pass<ND, 2>(s, A, pc, brick, cell, ob, has_h, hx, acc, acch);
  // the faces on the brick's upper sides whose R is a ghost, one wave per
  // direction (a pass only bricks at the domain's upper boundaries take)
  bool ghost_tail = false;
#pragma unroll
  for (int d = 0; d < ND; ++d)
    for (int k = 0; k < TANG; k += 64) {
      const int j0 = (2 * d + 1) * TANG + k;
      const int nb = TANG - k < 64 ? TANG - k : 64;
      const unsigned long long m = nb == 64 ? ~0ull : ((1ull << nb) - 1ull);
      ghost_tail = ghost_tail || (((s.gmask[j0 >> 6][0] >> (j0 & 63)) & m) != 0ull);
    }
  if (ghost_tail) {
    const int wv = t >> 6, tg = t & 63;
    if (tg < TANG) {
      if (wv == 0) adj_tail_face<ND, 0>(s, A, pc, brick, tg);
      else if (wv == 1) adj_tail_face<ND, 1>(s, A, pc, brick, tg);
      else if (ND == 3 && wv == 2) adj_tail_face<ND, (ND == 3 ? 2 : 1)>(s, A, pc, brick, tg);
    }
    __syncthreads();
    adj_gather<ND, true, 0>(s, cell, acc);
    if (has_h) adj_gather<ND, true, 0>(s, hx, acch);
    adj_gather<ND, true, 1>(s, cell, acc);
    if (has_h) adj_gather<ND, true, 1>(s, hx, acch);
    if constexpr (ND == 3) {
      adj_gather<ND, true, 2>(s, cell, acc);
      if (has_h) adj_gather<ND, true, 2>(s, hx, acch);
    }
  }
#pragma unroll
  for (int v = 0; v < NV; ++v) {
    double a = acc[v];
    if (A.Ubar) {
      // the stage's direct terms (k_adj_cell), fused
      const double o = ob[v];
      a += A.w * o;
      if (A.ubar_assign) A.Ubar[v * A.ld + gc] = A.wb * o;
      else if (A.wb != 0.0) A.Ubar[v * A.ld + gc] += A.wb * o;
    }
    if (A.assign) A.Ibar[v * A.ld + gc] = a;
    else A.Ibar[v * A.ld + gc] += a;
    if (hp >= 0) A.HS[(long long)hp * NV + v] = acch[v];
  }
}

// ---- adjoint kernels ---------------------------------------------------------

// Face slots of the reverse sweep, indexed by (d, R): the face between
// L = nghbr(R, 2d) and R.  FL/FR: the adjoint reaching L and R through the
// face flux (nv x nd x ld each); FG: half the adjoint of the face's
// tangential surface slope along sd (nv x nd x nd x ld), which reaches the
// slope stencils of both R and L.
struct FaceSlots {
  double* FL;
  double* FR;
  double* FG;
};

// One face (L, R, d) per thread: faces (m, c) of every internal cell c and
// faces (c, p) towards a ghost p.  Fbar = w dt/len_R obar_R [R internal]
// - w dt/len_L obar_L [L internal]; its pull-back through the face flux
// (flux_face_rev) goes to the face's slots (zeros when Fbar vanishes, so
// every existing face writes all of its slots).
template <int ND, int PHYS>
__global__ void __launch_bounds__(128) __attribute__((amdgpu_waves_per_eu(2))) k_adj_faces(AdjTopo t, PhysConst pc, const double* Q, const double* obar, double w, double dt,
                            FaceSlots fs) {
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
  const long long f = (long long)d * t.ld + R;
  const long long vs = (long long)ND * t.ld;   // stride between variables
  double Fb[NV];
  bool any = false;
  for (int v = 0; v < NV; ++v) {
    double fb = 0;
    if (R < t.n_int) fb += w * (dt / t.len[R]) * obar[v * t.ld + R];
    if (L < t.n_int) fb -= w * (dt / t.len[L]) * obar[v * t.ld + L];
    Fb[v] = fb;
    any = any || fb != 0.0;
  }
  if (!any) {
    for (int v = 0; v < NV; ++v) {
      fs.FL[v * vs + f] = 0.0;
      fs.FR[v * vs + f] = 0.0;
      if constexpr (PHYS == HOM3G_CNS)
        for (int sd = 0; sd < ND; ++sd)
          if (sd != d) fs.FG[(v * ND + sd) * vs + f] = 0.0;
    }
    return;
  }
  double qL[NV], qR[NV], Gt[NV][ND];
  for (int v = 0; v < NV; ++v) {
    qL[v] = Q[v * t.ld + L];
    qR[v] = Q[v * t.ld + R];
  }
  // slope walks per tangential direction: the stencil does not depend on
  // the variable; wall hooks flip the sign of the momentum slopes only.
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
    fs.FL[v * vs + f] = qLb[v];
    fs.FR[v * vs + f] = qRb[v];
    // G = 0.5 (slope(R) + slope(L)): half of the slope adjoint to each side
    if constexpr (PHYS == HOM3G_CNS)
      for (int sd = 0; sd < ND; ++sd)
        if (sd != d) fs.FG[(v * ND + sd) * vs + f] = 0.5 * Gtb[v][sd];
  }
}

// Adjoint of every cell's slopes: Sb[v][sd][c] = sum over the faces of c in
// the directions d != sd (the face below, then the face above) of their
// FG[v][sd] slot.
template <int ND>
__global__ void __launch_bounds__(256) k_adj_slopes(AdjTopo t, const double* FG, double* Sb) {
  constexpr int NV = ND + 2;
  const long long c = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (c >= t.n_int + t.ng) return;
  const long long vs = (long long)ND * t.ld;
  int m[ND], p[ND];
  for (int d = 0; d < ND; ++d) {
    m[d] = t.nghbr[(2 * d) * t.ld + c];
    p[d] = t.nghbr[(2 * d + 1) * t.ld + c];
  }
  for (int v = 0; v < NV; ++v)
    for (int sd = 0; sd < ND; ++sd) {
      const double* G = FG + (v * ND + sd) * vs;
      double a = 0.0;
      for (int d = 0; d < ND; ++d) {
        if (d == sd) continue;
        if (m[d] >= 0) a += G[(long long)d * t.ld + c];
        if (p[d] >= 0) a += G[(long long)d * t.ld + p[d]];
      }
      Sb[(v * ND + sd) * t.ld + c] = a;
    }
}

// Slope stencils reversed: slope(c, sd) = sign (Q[np] - Q[nn]) / (2 len)
// after its walk.  For every target cell the list holds the (c, sd) whose
// walk ends with the target as np or nn, in increasing c.
struct WalkList {
  const int32_t* rp;     // ncells + 1
  const int32_t* src;    // source cell c
  const uint8_t* meta;   // bit 0: target is nn (minus); bit 1: wall flip; bits 2-3: sd
  const double* coef;    // 1 / (2 len) of the stencil's centre
};

// Ibar[t] += the face adjoints of the faces of t (below: t is R; above: t
// is L) + the slope-stencil adjoints reaching t, in that order.
template <int ND, int PHYS>
__global__ void __launch_bounds__(256) k_adj_gather(AdjTopo t, FaceSlots fs, const double* Sb, WalkList wl,
                                                    double* Ibar) {
  constexpr int NV = ND + 2;
  const long long c = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (c >= t.n_int + t.ng) return;
  const long long vs = (long long)ND * t.ld;
  int m[ND], p[ND];
  for (int d = 0; d < ND; ++d) {
    m[d] = t.nghbr[(2 * d) * t.ld + c];
    p[d] = t.nghbr[(2 * d + 1) * t.ld + c];
  }
  int j0 = 0, j1 = 0;
  if constexpr (PHYS == HOM3G_CNS) {
    j0 = wl.rp[c];
    j1 = wl.rp[c + 1];
  }
  for (int v = 0; v < NV; ++v) {
    double a = 0.0;
    for (int d = 0; d < ND; ++d) {
      if (m[d] >= 0) a += fs.FR[v * vs + (long long)d * t.ld + c];
      if (p[d] >= 0) a += fs.FL[v * vs + (long long)d * t.ld + p[d]];
    }
    if constexpr (PHYS == HOM3G_CNS) {
      for (int j = j0; j < j1; ++j) {
        const int mt = wl.meta[j];
        const int sd = mt >> 2;
        double cf = wl.coef[j];
        if (mt & 1) cf = -cf;
        if ((mt & 2) && v < ND) cf = -cf;
        a += cf * Sb[(v * ND + sd) * t.ld + wl.src[j]];
      }
    }
    Ibar[v * t.ld + c] += a;
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

// The heat stage's reverse in one pass per cell c (internal or ghost).  The
// residual is linear, acc = sum_d (1/dx)(T_m - T_c) - (1/dx)(T_c - T_p), so
// with k(x) = (1/dx) w (dt/dx) obar(x) for an internal cell x with both
// neighbours along d, c receives -2 k(c) per such direction and +k of the
// internal cells it is the minus or plus neighbour of; then, for internal
// cells, the direct terms of out = w_base base + w (in + nf + dt S), added
// in that order.  fresh: Ibar holds no earlier terms (written, not read);
// ubar_fresh: the same for Ubar, whose ghost rows are cleared.
__device__ __forceinline__ double heat_k(const AdjTopo& t, const double* obar, double w, double dt, long long c) {
  const double dx = t.len[c];
  const double ab = w * (dt / dx) * obar[c];
  return (1 / dx) * ab;
}
__device__ __forceinline__ void heat_stage_cell(const AdjTopo& t, const double* obar, double w, double wb, double dt,
                                                double* Ibar, double* Ubar, double* Sbar, int fresh, int ubar_fresh,
                                                long long c) {
  if (c >= t.n_int + t.ng) return;
  double a = 0.0;
  for (int d = 0; d < t.nd; ++d) {
    const int m = t.nghbr[(2 * d) * t.ld + c];
    const int p = t.nghbr[(2 * d + 1) * t.ld + c];
    if (c < t.n_int && m >= 0 && p >= 0) a += -2.0 * heat_k(t, obar, w, dt, c);
    if (p >= 0 && p < t.n_int && t.nghbr[(2 * d) * t.ld + p] >= 0 && t.nghbr[(2 * d + 1) * t.ld + p] >= 0)
      a += heat_k(t, obar, w, dt, p);
    if (m >= 0 && m < t.n_int && t.nghbr[(2 * d) * t.ld + m] >= 0 && t.nghbr[(2 * d + 1) * t.ld + m] >= 0)
      a += heat_k(t, obar, w, dt, m);
  }
  double x = fresh ? a : Ibar[c] + a;
  if (c < t.n_int) {
    const double o = obar[c];
    x += w * o;
    if (ubar_fresh) Ubar[c] = wb * o;
    else if (wb != 0.0) Ubar[c] += wb * o;
    if (Sbar) Sbar[c] += w * dt * o;
  } else if (ubar_fresh) {
    Ubar[c] = 0.0;  // a ghost row, added to by the last stage
  }
  Ibar[c] = x;
}

// the adjoint a ghost row of the heat stage receives (the terms of
// heat_stage_cell for c >= n_int, in the same order)
__device__ __forceinline__ double heat_ghost_adj(const AdjTopo& t, const double* obar, double w, double dt,
                                                 long long c) {
  double a = 0.0;
  for (int d = 0; d < t.nd; ++d) {
    const int m = t.nghbr[(2 * d) * t.ld + c];
    const int p = t.nghbr[(2 * d + 1) * t.ld + c];
    if (p >= 0 && p < t.n_int && t.nghbr[(2 * d) * t.ld + p] >= 0 && t.nghbr[(2 * d + 1) * t.ld + p] >= 0)
      a += heat_k(t, obar, w, dt, p);
    if (m >= 0 && m < t.n_int && t.nghbr[(2 * d) * t.ld + m] >= 0 && t.nghbr[(2 * d + 1) * t.ld + m] >= 0)
      a += heat_k(t, obar, w, dt, m);
  }
  return a;
}


// The heat stage's output adjoint, weight and dt: with obar set, the
// ghost pull-back forms a ghost's adjoint with heat_ghost_adj instead of
// reading its Ibar row.
struct HeatStageAdj {
  const double* obar;
  double w, dt;
};

// pull-back of the ghost fill: the adjoint of ghost g (its Ibar row, or on
// tiled bricks the sum of the brick halo slots that read it: slots
// hrow[n_int + g] .. hrow[n_int + g + 1] of HS) through its condition, one
// input k per thread (k < nv the buffer's boundary cell,
// then this solver's lhs at the boundary cell (CHT), then the coupled
// solver's partner cell), into the slot GS[k * ng + g].  Periodic ghosts
// pass their adjoint on unchanged (slots k < nv, gathered at the image).
// Every (g, k) slot of a launched input k (ghost_inputs) is written.
template <int ND>
__device__ __forceinline__ void adj_ghost_slot(const AdjTopo& t, long long g, int k, const double* bufv,
                                               const double* ownv, const double* othv, const double* Ibar, double* GS,
                                               const double* HS, const int32_t* hrow, const HeatStageAdj& hs) {
  double* slot = GS + (long long)k * t.ng + g;
  const int bc = t.g_apply[g];
  if (bc < 0) { *slot = 0.0; return; }
  const long long ld = t.ld, gc = t.n_int + g;
  const int nv = t.nvars;
  const int kind = t.kind[bc];
  const bool periodic = kind == HOM3G_BC_PERIODIC;
  const bool cht = uses_own(kind);
  const int onv = (kind == HOM3G_BC_CHT_FLUID) ? 1 : ND + 2;
  const int nin = periodic ? nv : nv + (cht ? nv + onv : 0);
  if (k >= nin) { *slot = 0.0; return; }
  // the ghost's adjoint (tiled: its halo slots summed in entry order)
  double gb[ND + 2];
  if (HS) {
#pragma unroll
    for (int v = 0; v < ND + 2; ++v) gb[v] = 0.0;
    // four slots per round, their loads issued together
    const int h0 = hrow[gc], h1 = hrow[gc + 1];
    for (int h = h0; h < h1; h += 4) {
      double x[4][ND + 2];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const double* e = HS + (long long)(h + u < h1 ? h + u : h0) * nv;
#pragma unroll
        for (int v = 0; v < ND + 2; ++v) x[u][v] = (v < nv) ? e[v] : 0.0;
      }
#pragma unroll
      for (int u = 0; u < 4; ++u)
        if (h + u < h1) {
#pragma unroll
          for (int v = 0; v < ND + 2; ++v) gb[v] += x[u][v];
        }
    }
  } else if (hs.obar) {
#pragma unroll
    for (int v = 0; v < ND + 2; ++v) gb[v] = 0.0;
    gb[0] = heat_ghost_adj(t, hs.obar, hs.w, hs.dt, gc);
  } else {
#pragma unroll
    for (int v = 0; v < ND + 2; ++v) gb[v] = (v < nv) ? Ibar[v * ld + gc] : 0.0;
  }
  if (periodic) {
    double x = 0.0;
#pragma unroll
    for (int v = 0; v < ND + 2; ++v)
      if (v == k) x = gb[v];
    *slot = x;
    return;
  }
  bool any = false;
#pragma unroll
  for (int v = 0; v < ND + 2; ++v) any = any || gb[v] != 0.0;
  if (!any) { *slot = 0.0; return; }
  const int b = t.g_bndry[g];
  const int pc_ = cht ? t.g_partner[g] : 0;
  double cvb0[ND + 2], own0[ND + 2], oth0[ND + 2];
  for (int v = 0; v < ND + 2; ++v) {
    cvb0[v] = (v < nv) ? bufv[v * ld + b] : 0.0;
    own0[v] = (v < nv) ? ownv[v * ld + b] : 0.0;
  }
  for (int v = 0; v < ND + 2; ++v) oth0[v] = (cht && v < onv) ? othv[v * t.other_ld + pc_] : 0.0;
  // the seeds as bit tests of per-array masks (input k sets one bit): the
  // arrays stay in registers instead of taking a store at index k
  const unsigned mc = k < nv ? 1u << k : 0u;
  const unsigned mo = (k >= nv && k < 2 * nv) ? 1u << (k - nv) : 0u;
  const unsigned mp = (k >= 2 * nv && k < 2 * nv + onv) ? 1u << (k - 2 * nv) : 0u;
  Dual cvb[ND + 2], own[ND + 2], oth[ND + 2], out[ND + 2];
#pragma unroll
  for (int v = 0; v < ND + 2; ++v) {
    cvb[v] = Dual(cvb0[v], ((mc >> v) & 1u) ? 1.0 : 0.0);
    own[v] = Dual(own0[v], ((mo >> v) & 1u) ? 1.0 : 0.0);
    oth[v] = Dual(oth0[v], ((mp >> v) & 1u) ? 1.0 : 0.0);
  }
  ghost_value<ND, Dual>(t, bc, (int)g, cvb, own, oth, out);
  double xb = 0;
  for (int v = 0; v < ND + 2; ++v)
    if (v < nv) xb += gb[v] * out[v].d;
  *slot = xb;
}

template <int ND>
__global__ void __launch_bounds__(256) k_adj_ghosts(AdjTopo t, const double* bufv, const double* ownv, const double* othv,
                                                    const double* Ibar, double* GS, const double* HS,
                                                    const int32_t* hrow) {
  const long long g = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (g >= t.ng) return;
  adj_ghost_slot<ND>(t, g, (int)blockIdx.y, bufv, ownv, othv, Ibar, GS, HS, hrow, HeatStageAdj{nullptr, 0.0, 0.0});
}

// the heat stage's reverse and its ghost pull-back in one launch: blocks
// [0, nhb) run heat_stage_cell, the rest the (ghost, input) slots with the
// ghosts' adjoints formed from obar (they do not read Ibar's ghost rows,
// which the first blocks write)
template <int ND>
__global__ void __launch_bounds__(256) k_adj_heat_stage_ghosts(AdjTopo t, const double* obar, double w, double wb,
                                                               double dt, double* Ibar, double* Ubar, double* Sbar,
                                                               int fresh, int ubar_fresh, long long nhb, int nk,
                                                               const double* bufv, const double* ownv,
                                                               const double* othv, double* GS) {
  if (blockIdx.x < nhb) {
    heat_stage_cell(t, obar, w, wb, dt, Ibar, Ubar, Sbar, fresh, ubar_fresh,
                    blockIdx.x * (long long)blockDim.x + threadIdx.x);
    return;
  }
  const long long i = (blockIdx.x - nhb) * (long long)blockDim.x + threadIdx.x;
  if (i >= (long long)nk * t.ng) return;
  const int k = (int)(i / t.ng);
  const long long g = i - (long long)k * t.ng;
  adj_ghost_slot<ND>(t, g, k, bufv, ownv, othv, nullptr, GS, nullptr, nullptr, HeatStageAdj{obar, w, dt});
}

// Ghost lists: for every target cell, the ghosts whose slots reach it, in
// increasing ghost index; in the lists by internal cell the entries of
// coupled (CHT) ghosts carry bit 31 (kCoupledGhost).
constexpr uint32_t kCoupledGhost = 0x80000000u;
struct GhostList {
  const int32_t* rp;  // target tl[i] has the entries g[rp[i] .. rp[i + 1])
  const int32_t* g;
  const int32_t* tl;  // the targets with a non-empty list, increasing
  long long nt;       // their number
};

// One gather of ghost slots: out1[v][c] += sum slots[k0 + v] over the ghosts
// listed for c and out2[v][c] += sum slots[k1 + v] over the coupled ones
// (only they have inputs past nv: the other ghosts' slots there are zero;
// out2 may alias out1: the two sums are added in that order by one thread);
// out2 == nullptr skips the second sum, and so does a target without
// coupled ghosts.
struct GatherJob {
  GhostList gl;
  int nvo, k0, k1;
  double* out1;
  double* out2;
  long long ld;
};

// two independent gathers in one launch: threads [0, a.gl.nt) run a, the
// next b.gl.nt run b
// the ghost-slot sums of one target: sa[v] = sum slots[k0 + v] over the
// entries g[j0 .. j1), sb[v] = the same of slots[k1 + v] over the coupled
// ones (want2), in entry order; U entries per round, all their slot loads
// issued together
template <int U>
__device__ __forceinline__ void sum_slots(const int32_t* g, int j0, int j1, int nvo, int k0, int k1, bool want2,
                                          const double* slots, long long ng, double* sa, double* sb, bool& coupled) {
#pragma unroll
  for (int v = 0; v < 5; ++v) sa[v] = sb[v] = 0.0;
  coupled = false;
  for (int j = j0; j < j1; j += U) {
    uint32_t e[U];
#pragma unroll
    for (int u = 0; u < U; ++u) e[u] = (uint32_t)g[j + u < j1 ? j + u : j0];
    double x[U][5], y[U][5];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const long long gj = (long long)(e[u] & ~kCoupledGhost);
      const bool cu = want2 && (e[u] & kCoupledGhost);
#pragma unroll
      for (int v = 0; v < 5; ++v) {
        x[u][v] = (v < nvo) ? slots[(long long)(k0 + v) * ng + gj] : 0.0;
        y[u][v] = (v < nvo && cu) ? slots[(long long)(k1 + v) * ng + gj] : 0.0;
      }
    }
#pragma unroll
    for (int u = 0; u < U; ++u)
      if (j + u < j1) {
        const bool cu = (e[u] & kCoupledGhost) != 0u;
        coupled = coupled || cu;
#pragma unroll
        for (int v = 0; v < 5; ++v) {
          sa[v] += x[u][v];
          if (want2 && cu) sb[v] += y[u][v];
        }
      }
  }
}

// target i of gather J: out1 += sa, then out2 += sb where coupled
__device__ __forceinline__ void gather_target(const int32_t* rp, const int32_t* g, const int32_t* tl, int nvo, int k0,
                                              int k1, double* out1, double* out2, long long ld, long long i,
                                              const double* slots, long long ng) {
  const int c = tl[i];
  double sa[5], sb[5];
  bool coupled;
  sum_slots<4>(g, rp[i], rp[i + 1], nvo, k0, k1, out2 != nullptr, slots, ng, sa, sb, coupled);
#pragma unroll
  for (int v = 0; v < 5; ++v)
    if (v < nvo) {
      out1[v * ld + c] += sa[v];
      if (out2 && coupled) out2[v * ld + c] += sb[v];
    }
}

__global__ void __launch_bounds__(256) k_gather_ghost_slots(GatherJob a, GatherJob b, const double* slots,
                                                            long long ng) {
  long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  const bool first = i < a.gl.nt;
  if (!first) i -= a.gl.nt;
  // the job's fields selected one by one (a reference to either argument
  // would copy both to memory)
  if (i >= (first ? a.gl.nt : b.gl.nt)) return;
  gather_target(first ? a.gl.rp : b.gl.rp, first ? a.gl.g : b.gl.g, first ? a.gl.tl : b.gl.tl,
                first ? a.nvo : b.nvo, first ? a.k0 : b.k0, first ? a.k1 : b.k1, first ? a.out1 : b.out1,
                first ? a.out2 : b.out2, first ? a.ld : b.ld, i, slots, ng);
}

// Tiled bricks: one gather per stage into the internal cells.  Target i
// (cell tl[i]; rng[i] = its halo-slot range h0, h1 in HS and its ghost-slot
// range g0, g1 in the own ghost list g) takes Ibar += the halo slots of the
// bricks that read it (entry order), then += its ghosts' slots k < nv, and
// Ownbar += the coupled ghosts' slots nv .. 2 nv - 1; the threads past nt
// run the gather `part` into the coupled solver's cells.  Each sum is the
// one the separate halo and ghost gathers formed, added in the same order.
__global__ void __launch_bounds__(256) k_adj_tiled_gather(long long nt, const int32_t* tl, const int4* rng, int nv,
                                                          long long ld, const double* HS, const int32_t* g,
                                                          const double* slots, long long ng, double* Ibar,
                                                          double* Ownbar, GatherJob part) {
  long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i >= nt) {
    i -= nt;
    if (i < part.gl.nt)
      gather_target(part.gl.rp, part.gl.g, part.gl.tl, part.nvo, part.k0, part.k1, part.out1, part.out2, part.ld, i,
                    slots, ng);
    return;
  }
  const int c = tl[i];
  const int4 r = rng[i];
  double ib[5];
#pragma unroll
  for (int v = 0; v < 5; ++v) ib[v] = (v < nv) ? Ibar[v * ld + c] : 0.0;
  double a[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
  for (int j = r.x; j < r.y; j += 4) {
    double h[4][5];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const double* e = HS + (long long)(j + u < r.y ? j + u : r.x) * nv;
#pragma unroll
      for (int v = 0; v < 5; ++v) h[u][v] = (v < nv) ? e[v] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < 4; ++u)
      if (j + u < r.y) {
#pragma unroll
        for (int v = 0; v < 5; ++v) a[v] += h[u][v];
      }
  }
  double sa[5], sb[5];
  bool coupled;
  sum_slots<2>(g, r.z, r.w, nv, 0, nv, Ownbar != nullptr, slots, ng, sa, sb, coupled);
#pragma unroll
  for (int v = 0; v < 5; ++v)
    if (v < nv) {
      double x = ib[v];
      if (r.x < r.y) x += a[v];
      if (r.z < r.w) x += sa[v];
      Ibar[v * ld + c] = x;
      if (Ownbar && coupled) Ownbar[v * ld + c] += sb[v];
    }
}

// Ordered sum of n values into acc[0] (one block; each thread sums a
// contiguous chunk, then a fixed tree): the same result in every run.
__global__ void __launch_bounds__(256) k_sum_ordered(long long n, const double* x, long long stride, int m,
                                                     double* acc) {
  __shared__ double red[256];
  for (int k = 0; k < m; ++k) {
    const long long per = (n + 255) / 256;
    const long long i0 = threadIdx.x * per, i1 = i0 + per < n ? i0 + per : n;
    double a = 0.0;
    for (long long i = i0; i < i1; ++i) a += x[k * stride + i];
    red[threadIdx.x] = a;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
      if ((int)threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
      __syncthreads();
    }
    if (threadIdx.x == 0) acc[k] += red[0];
    __syncthreads();
  }
}

// wall heat flux Q of the solid (its CHT ghosts) times the heat dt: one
// term per ghost into term[g] (with tangents: its derivative into
// term[ng + g]); k_sum_ordered adds them up
template <int ND>
__global__ void k_objective(AdjTopo th, const double* Hv, const double* Ht, const double* Cv, const double* Ct,
                            const DevScalars* hscal, double* term) {
  const long long g = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (g >= th.ng) return;
  const int bc = th.g_apply[g];
  if (bc < 0 || th.kind[bc] != HOM3G_BC_CHT_SOLID) {
    term[g] = 0.0;
    if (Ht) term[th.ng + g] = 0.0;
    return;
  }
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
  term[g] = dt * q.v;
  if (Ht) term[th.ng + g] = dt * q.d;
}

// adjoint of one step's objective term: slot k = 0 for the solid boundary
// cell, k = 1 + v for the fluid partner's variables (OS[k * ng + g])
template <int ND>
__global__ void k_objective_adj(AdjTopo th, const double* Hv, const double* Cv, double scale, double* OS) {
  const long long g = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (g >= th.ng) return;
  const int bc = th.g_apply[g];
  if (bc < 0 || th.kind[bc] != HOM3G_BC_CHT_SOLID) {
    for (int k = 0; k < ND + 3; ++k) OS[(long long)k * th.ng + g] = 0.0;
    return;
  }
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
    OS[(long long)k * th.ng + g] = scale * q.d;
  }
}

// final-state misfit of the solid: per-block partial sums of
// 1/2 (T_c - obs_c)^2 (and, with a tangent, of (T_c - obs_c) dT_c) into
// part[blockIdx] and part[gridDim + blockIdx]; k_sum_ordered adds them up
__global__ void __launch_bounds__(256) k_misfit(long long n, const double* T, const double* Tt, const double* obs,
                                                double* part) {
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
    part[blockIdx.x] = (red[0][0] + red[0][1]) + (red[0][2] + red[0][3]);
    part[gridDim.x + blockIdx.x] = (red[1][0] + red[1][1]) + (red[1][2] + red[1][3]);
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
  // the primal steps run on the table-driven kernels or, for the fluid, on
  // tiled bricks (the same fluxes to a few ULP; the reverse sweep
  // differentiates the table-driven form)
  H3_REQUIRE((cns.kernel == HOM3G_KERNEL_GENERIC || cns.brick_table) && heat.kernel == HOM3G_KERNEL_GENERIC,
             "the adjoint runs on the table-driven kernels or tiled bricks");
  H3_REQUIRE(cns.P.precision == HOM3G_FP64 && heat.P.precision == HOM3G_FP64, "the adjoint runs on the FP64 state");
  for (SolverDev* S : {&cns, &heat}) {
    H3_REQUIRE(S->P.ti == HOM3G_TI_EULER_FWD || S->P.ti == HOM3G_TI_SSP_RK3,
               "the adjoint supports forward Euler and SSP-RK3");
    H3_REQUIRE(S->P.recon == HOM3G_RECON_NONE && S->P.flux != HOM3G_FLUX_LLF, "the adjoint supports AUSM without reconstruction");
    H3_REQUIRE(!S->nonuniform, "the adjoint supports uniform grids");
    for (const BCDesc& b : S->bcs) {
      H3_REQUIRE(b.kind != HOM3G_BC_ISENTROPIC_VORTEX, "the adjoint does not support the isentropic vortex boundary");
      H3_REQUIRE(b.kind != HOM3G_BC_HOST_CALLBACK, "the adjoint does not support host-callback conditions");
    }
  }
  H3_REQUIRE(cns.P.ti == heat.P.ti, "both solvers must use the same time integration");
}

double scaling_of(double Tr0, const SolverDev& cns) { return Tr0 * cns.pc.Pr0 * cns.pc.Re0; }

template <class F>
void by_nd(int nd, F f) {
  if (nd == 2) f(std::integral_constant<int, 2>());
  else f(std::integral_constant<int, 3>());
}

// acc[0] += J term of the current step (with tangents also acc[1] += its
// derivative), ordered; term: scratch of 2 ng values
void objective_launch(Ctx& C, SolverDev& heat, SolverDev& cns, const AdjTopo& th, const double* Ht,
                      const double* Ct, DevBuf<double>& term, double* acc) {
  if (heat.n_ghost == 0) return;
  term.alloc(2 * (size_t)heat.n_ghost);
  by_nd(heat.nd, [&](auto ndc) {
    constexpr int ND = decltype(ndc)::value;
    hipLaunchKernelGGL(k_objective<ND>, dim3(nblk(heat.n_ghost)), dim3(256), 0, C.stream, th,
                       (const double*)heat.lhs.p, Ht, (const double*)cns.lhs.p, Ct, (const DevScalars*)heat.scal.p,
                       term.p);
  });
  H3_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_sum_ordered, dim3(1), dim3(256), 0, C.stream, (long long)heat.n_ghost,
                     (const double*)term.p, (long long)heat.n_ghost, Ht ? 2 : 1, acc);
  H3_LAUNCH_CHECK();
}

void misfit_launch(Ctx& C, SolverDev& heat, const double* obs, const double* Ht, DevBuf<double>& part, double* acc) {
  const long long n = heat.n_int;
  const unsigned grid = (unsigned)std::min<long long>(nblk(n), 1024);
  part.alloc(2 * (size_t)grid);
  hipLaunchKernelGGL(k_misfit, dim3(grid), dim3(256), 0, C.stream, n, (const double*)heat.lhs.p, Ht, obs, part.p);
  H3_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_sum_ordered, dim3(1), dim3(256), 0, C.stream, (long long)grid, (const double*)part.p,
                     (long long)grid, Ht ? 2 : 1, acc);
  H3_LAUNCH_CHECK();
}

// ---- index lists of the reverse sweep (built once per run from the
// topology, on the host) ------------------------------------------------------
struct RevMaps {
  DevBuf<int32_t> wrp, wsrc, orp, og, otl, prp, pg, ptl;
  long long ont = 0, pnt = 0;
  DevBuf<uint8_t> wmeta;
  DevBuf<double> wcoef;
  WalkList walk() const { return WalkList{wrp.p, wsrc.p, wmeta.p, wcoef.p}; }
  GhostList own() const { return GhostList{orp.p, og.p, otl.p, ont}; }
  GhostList partner() const { return GhostList{prp.p, pg.p, ptl.p, pnt}; }
};

// the rows of a CSR map that are not empty
std::vector<int32_t> nonempty_rows(const std::vector<int32_t>& rp) {
  std::vector<int32_t> tl;
  for (size_t c = 0; c + 1 < rp.size(); ++c)
    if (rp[c + 1] > rp[c]) tl.push_back((int32_t)c);
  return tl;
}

// the entry ranges of the listed rows tl (the non-empty rows of rp, in
// order) by list position: row tl[i] spans [out[i], out[i + 1]) (the empty
// rows between listed ones add nothing)
std::vector<int32_t> ranges_by_position(const std::vector<int32_t>& rp, const std::vector<int32_t>& tl) {
  std::vector<int32_t> out;
  out.reserve(tl.size() + 1);
  for (int32_t c : tl) out.push_back(rp[c]);
  out.push_back(rp.back());
  return out;
}

GatherJob gather_job(const GhostList& gl, int nvo, int k0, double* out1, int k1, double* out2, long long ld) {
  return GatherJob{gl, nvo, k0, k1, out1, out2, ld};
}
GatherJob no_gather() { return GatherJob{GhostList{nullptr, nullptr, nullptr, 0}, 0, 0, 0, nullptr, nullptr, 0}; }

// the gathers a and b of ghost slots (see k_gather_ghost_slots), one launch
void gather_ghost_slots(Ctx& C, const GatherJob& a, const GatherJob& b, const double* slots, long long ng) {
  const long long n = a.gl.nt + b.gl.nt;
  if (n == 0) return;
  hipLaunchKernelGGL(k_gather_ghost_slots, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, C.stream, a, b, slots,
                     ng);
  H3_LAUNCH_CHECK();
}

// counting sort of (target, value) pairs into CSR rows, values kept in
// insertion order within a row
template <class Fill>
void csr(int64_t ntarget, size_t nent, Fill fill, std::vector<int32_t>& rp, std::vector<int32_t>& order) {
  std::vector<int64_t> tgt(nent);
  fill(tgt);
  rp.assign(ntarget + 1, 0);
  for (int64_t t : tgt) rp[t + 1]++;
  for (int64_t i = 0; i < ntarget; ++i) rp[i + 1] += rp[i];
  std::vector<int32_t> pos(rp.begin(), rp.end() - 1);
  order.assign(nent, 0);
  for (size_t e = 0; e < nent; ++e) order[pos[tgt[e]]++] = (int32_t)e;
}

void build_rev_maps(Ctx& C, SolverDev& S, const SolverDev* other, RevMaps& M) {
  const int64_t ld = S.cap, nc = S.n_int + S.n_ghost, ng = S.n_ghost;
  const int nd = S.nd;
  std::vector<int32_t> nb((size_t)S.npos * ld), gb(ng), gim(ng), ghook(ng), gapp(ng), gpart(ng);
  std::vector<double> len(ld);
  S.nghbr.download(nb.data(), nb.size(), C.stream);
  S.length.download(len.data(), ld, C.stream);
  if (ng) {
    S.g_bndry.download(gb.data(), ng, C.stream);
    S.g_image.download(gim.data(), ng, C.stream);
    S.g_hookbc.download(ghook.data(), ng, C.stream);
    S.g_applybc.download(gapp.data(), ng, C.stream);
    S.g_partner.download(gpart.data(), ng, C.stream);
  }
  H3_HIP(hipStreamSynchronize(C.stream));
  auto kind_of = [&](int bc) { return bc >= 0 && bc < (int)S.bcs.size() ? S.bcs[bc].kind : -1; };
  auto wall = [](int k) {
    return k == HOM3G_BC_ADIABATIC_NOSLIP || k == HOM3G_BC_ISOTHERMAL_NOSLIP || k == HOM3G_BC_CHT_FLUID;
  };
  std::vector<int32_t> rp, order;
  if (S.P.physics == HOM3G_CNS) {
    // the walk of slope (solver.hpp:596-623) for every cell and direction,
    // as slope_walk does it on the device
    struct E { int64_t tgt; int32_t src; uint8_t meta; double coef; };
    std::vector<E> ent;
    ent.reserve((size_t)nc * nd * 2);
    for (int64_t c0 = 0; c0 < nc; ++c0)
      for (int sd = 0; sd < nd; ++sd) {
        int64_t c = c0;
        bool flip = false, ok = false;
        int np = -1, nn = -1;
        for (int it = 0; it < 16; ++it) {
          nn = nb[(2 * sd) * ld + c];
          np = nb[(2 * sd + 1) * ld + c];
          if (nn >= 0 && np >= 0) { ok = true; break; }
          if (nn >= 0) { c = nn; continue; }
          if (np >= 0) { c = np; continue; }
          const int64_t g = c - S.n_int;
          if (g < 0 || g >= ng) break;
          const int k = kind_of(ghook[g]);
          if (k == HOM3G_BC_PERIODIC) { c = gim[g]; continue; }
          if (wall(k)) flip = !flip;
          c = gb[g];
        }
        if (!ok) continue;
        const double cf = 1.0 / (2. * len[c]);
        const uint8_t m = (uint8_t)((sd << 2) | (flip ? 2 : 0));
        ent.push_back(E{np, (int32_t)c0, m, cf});
        ent.push_back(E{nn, (int32_t)c0, (uint8_t)(m | 1), cf});
      }
    csr(nc, ent.size(), [&](std::vector<int64_t>& t) { for (size_t e = 0; e < ent.size(); ++e) t[e] = ent[e].tgt; },
        rp, order);
    std::vector<int32_t> src(ent.size());
    std::vector<uint8_t> meta(ent.size());
    std::vector<double> coef(ent.size());
    for (size_t j = 0; j < order.size(); ++j) {
      const E& e = ent[order[j]];
      src[j] = e.src;
      meta[j] = e.meta;
      coef[j] = e.coef;
    }
    M.wrp.alloc(rp.size());
    M.wrp.upload(rp.data(), rp.size(), C.stream);
    M.wsrc.alloc(std::max<size_t>(1, src.size()));
    M.wmeta.alloc(std::max<size_t>(1, meta.size()));
    M.wcoef.alloc(std::max<size_t>(1, coef.size()));
    if (!src.empty()) {
      M.wsrc.upload(src.data(), src.size(), C.stream);
      M.wmeta.upload(meta.data(), meta.size(), C.stream);
      M.wcoef.upload(coef.data(), coef.size(), C.stream);
    }
  }
  // ghosts by the internal cell their slots reach: the image (periodic) or
  // the boundary cell
  csr(S.n_int, ng, [&](std::vector<int64_t>& t) {
        for (int64_t g = 0; g < ng; ++g) t[g] = kind_of(gapp[g]) == HOM3G_BC_PERIODIC ? gim[g] : gb[g];
      }, rp, order);
  {
    std::vector<int32_t> flagged(order.size());
    for (size_t j = 0; j < order.size(); ++j) {
      const int k = kind_of(gapp[order[j]]);
      const bool coupled = k == HOM3G_BC_CHT_FLUID || k == HOM3G_BC_CHT_SOLID;
      flagged[j] = (int32_t)((uint32_t)order[j] | (coupled ? kCoupledGhost : 0u));
    }
    M.og.alloc(std::max<size_t>(1, flagged.size()));
    if (!flagged.empty()) M.og.upload(flagged.data(), flagged.size(), C.stream);
  }
  std::vector<int32_t> orl, prl;  // alive until the closing synchronize
  {
    const std::vector<int32_t> tl = nonempty_rows(rp);
    M.ont = (long long)tl.size();
    M.otl.alloc(std::max<size_t>(1, tl.size()));
    if (!tl.empty()) M.otl.upload(tl.data(), tl.size(), C.stream);
    orl = ranges_by_position(rp, tl);
    M.orp.alloc(orl.size());
    M.orp.upload(orl.data(), orl.size(), C.stream);
  }
  // coupled ghosts by their partner cell in the other solver
  if (other) {
    std::vector<int64_t> tg;
    std::vector<int32_t> who;
    for (int64_t g = 0; g < ng; ++g) {
      const int k = kind_of(gapp[g]);
      if (k == HOM3G_BC_CHT_FLUID || k == HOM3G_BC_CHT_SOLID) { tg.push_back(gpart[g]); who.push_back((int32_t)g); }
    }
    csr(other->n_int, tg.size(), [&](std::vector<int64_t>& t) { t = tg; }, rp, order);
    std::vector<int32_t> lst(order.size());
    for (size_t j = 0; j < order.size(); ++j) lst[j] = who[order[j]];
    M.pg.alloc(std::max<size_t>(1, lst.size()));
    if (!lst.empty()) M.pg.upload(lst.data(), lst.size(), C.stream);
    const std::vector<int32_t> tl = nonempty_rows(rp);
    M.pnt = (long long)tl.size();
    M.ptl.alloc(std::max<size_t>(1, tl.size()));
    if (!tl.empty()) M.ptl.upload(tl.data(), tl.size(), C.stream);
    prl = ranges_by_position(rp, tl);
    M.prp.alloc(prl.size());
    M.prp.upload(prl.data(), prl.size(), C.stream);
  }
  H3_HIP(hipStreamSynchronize(C.stream));
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
  double* Ucur;      // adjoint of U^n (zeroed by the caller unless ucur_fresh)
  double* Othbar;    // adjoint of the coupled solver's state
  double* Sbar;      // adjoint of the source (heat) or null
  double* I;         // work: stage input adjoint
  double* O;         // work: stage output adjoint
  double dt;
  bool ucur_fresh = false;  // Ucur is written before it is added to (writes_ucur_fresh)
};

// the reverse of a CNS solver on tiled bricks (k_adj_brick and
// k_adj_tiled_gather)
bool tiled_rev(const SolverDev& S) { return S.brick_table && S.P.physics == HOM3G_CNS; }

// work buffers and index lists of one solver's reverse sweep
struct RevWork {
  RevMaps maps;
  DevBuf<double> fl, fr, fg, sb, gs, hs;
  // tiled bricks: the slot of each halo entry in hs, and the internal cells
  // the stage gather adds to with their halo-slot and ghost-slot ranges
  DevBuf<int32_t> hpos, utl;
  DevBuf<int4> urng;
  long long nu = 0;
  void alloc(Ctx& C, SolverDev& S, const SolverDev* other) {
    build_rev_maps(C, S, other, maps);
    const size_t ld = S.cap, nd = S.nd, nv = S.nvars;
    if (S.P.physics == HOM3G_HEAT) {
      // the heat reverse needs no work buffer
    } else if (tiled_rev(S)) {
      alloc_tiled(C, S);
    } else {
      fl.alloc(nv * nd * ld);
      fr.alloc(nv * nd * ld);
      if (S.P.physics == HOM3G_CNS) {
        fg.alloc(nv * nd * nd * ld);
        sb.alloc(nv * nd * ld);
      }
    }
    gs.alloc(std::max<size_t>(1, 3 * (nd + 2) * (size_t)S.n_ghost));
  }
  // one slot per read halo entry and variable, the entries in the order of
  // the cells they read (the brick reverse map's order); the targets are
  // the internal cells some halo entry or ghost reaches
  void alloc_tiled(Ctx& C, SolverDev& S) {
    const int64_t nhe = S.nd == 3 ? (int64_t)HaloList<3, HOM3G_CNS, false>::N : (int64_t)HaloList<2, HOM3G_CNS, false>::N;
    const int64_t ncell = S.nd == 3 ? (int64_t)BrickCfg<3>::NT : (int64_t)BrickCfg<2>::NT;
    const int64_t nent = (S.n_int / ncell) * nhe;
    std::vector<int32_t> rp(S.brick_rev_rp.n), src(S.brick_rev_src.n), otl((size_t)maps.ont), orp((size_t)maps.ont + 1);
    S.brick_rev_rp.download(rp.data(), rp.size(), C.stream);
    S.brick_rev_src.download(src.data(), src.size(), C.stream);
    if (maps.ont > 0) {
      maps.otl.download(otl.data(), otl.size(), C.stream);
      maps.orp.download(orp.data(), orp.size(), C.stream);
    }
    H3_HIP(hipStreamSynchronize(C.stream));
    const int32_t total = rp.back();
    std::vector<int32_t> pos((size_t)nent, -1);
    for (int32_t k = 0; k < total; ++k) pos[(size_t)src[k]] = k;
    std::vector<int32_t> tl;
    std::vector<int4> rng;
    size_t p = 0;  // next own-list target
    for (int64_t c = 0; c < S.n_int; ++c) {
      const bool h = rp[c + 1] > rp[c];
      const bool g = p < otl.size() && otl[p] == c;
      if (!h && !g) continue;
      tl.push_back((int32_t)c);
      rng.push_back(make_int4(rp[c], rp[c + 1], g ? orp[p] : 0, g ? orp[p + 1] : 0));
      if (g) ++p;
    }
    H3_REQUIRE(p == otl.size(), "ghost targets beyond the internal cells");
    hs.alloc(std::max<size_t>(1, (size_t)total * (size_t)S.nvars));
    hpos.alloc(pos.size());
    hpos.upload(pos.data(), pos.size(), C.stream);
    nu = (long long)tl.size();
    utl.alloc(std::max<size_t>(1, tl.size()));
    urng.alloc(std::max<size_t>(1, rng.size()));
    if (!tl.empty()) {
      utl.upload(tl.data(), tl.size(), C.stream);
      urng.upload(rng.data(), rng.size(), C.stream);
    }
    H3_HIP(hipStreamSynchronize(C.stream));
  }
};

// Ibar += adjoint of the stage's face fluxes (obar: adjoint of the stage
// output, w: its weight)
void faces_rev(Ctx& C, SolverDev& S, const AdjTopo& t, RevWork& W, const double* Q, const double* obar, double w,
               double dt, double* Ibar, double wb = 0.0, double* Ubar = nullptr, bool fresh = false,
               bool ubar_fresh = false) {
  const long long nc = S.n_int + S.n_ghost;
  by_nd(S.nd, [&](auto ndc) {
    constexpr int ND = decltype(ndc)::value;
    if (tiled_rev(S)) {
      AdjBrickArgs A;
      A.Q = Q;
      A.obar = obar;
      A.Ibar = Ibar;
      A.HS = W.hs.p;
      A.hpos = W.hpos.p;
      A.halo = S.brick_halo.p;
      A.ld = S.cap;
      A.n_int = S.n_int;
      A.fbs = w * (dt / S.h_uniform);
      A.inv_h = 1.0 / S.h_uniform;
      A.w = Ubar ? w : 0.0;  // the direct terms, when fused here
      A.wb = Ubar ? wb : 0.0;
      A.Ubar = Ubar;
      A.assign = fresh ? 1 : 0;
      A.ubar_assign = (Ubar && ubar_fresh) ? 1 : 0;
      const long long nb = S.n_int / BrickCfg<ND>::NT;
      hipLaunchKernelGGL(k_adj_brick<ND>, dim3((unsigned)nb), dim3(BrickCfg<ND>::NT), 0, C.stream, A, S.pc);
      H3_LAUNCH_CHECK();
    } else {
      const FaceSlots fs{W.fl.p, W.fr.p, W.fg.p};
      const dim3 grid(nblk(S.n_int, 128), 2 * ND);
      if (S.P.physics == HOM3G_CNS) {
        hipLaunchKernelGGL((k_adj_faces<ND, HOM3G_CNS>), grid, dim3(128), 0, C.stream, t, S.pc, Q, obar, w, dt, fs);
        H3_LAUNCH_CHECK();
        hipLaunchKernelGGL(k_adj_slopes<ND>, dim3(nblk(nc)), dim3(256), 0, C.stream, t, (const double*)W.fg.p, W.sb.p);
        H3_LAUNCH_CHECK();
        hipLaunchKernelGGL((k_adj_gather<ND, HOM3G_CNS>), dim3(nblk(nc)), dim3(256), 0, C.stream, t, fs,
                           (const double*)W.sb.p, W.maps.walk(), Ibar);
      } else {
        hipLaunchKernelGGL((k_adj_faces<ND, HOM3G_EULER>), grid, dim3(128), 0, C.stream, t, S.pc, Q, obar, w, dt,
                           fs);
        H3_LAUNCH_CHECK();
        hipLaunchKernelGGL((k_adj_gather<ND, HOM3G_EULER>), dim3(nblk(nc)), dim3(256), 0, C.stream, t, fs,
                           (const double*)nullptr, W.maps.walk(), Ibar);
      }
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

// pull-back of the ghost fill of `buf`: into Ibar (boundary cells, periodic
// images), Ownbar (this solver's lhs, CHT) and Othbar (the coupled solver)
// the ghost-fill inputs whose slots the gathers read: the boundary cell's
// nv values, the solver's own lhs there (CHT) and the coupled solver's
// partner cell (1 value for the fluid's interface ghosts, nd + 2 for the
// solid's); the slots of larger inputs are never formed
int ghost_inputs(const SolverDev& S) {
  const int k = 2 * S.nvars + (S.P.physics == HOM3G_CNS ? 1 : S.nd + 2);
  return std::min(k, 3 * (S.nd + 2));
}

// (tiled bricks: the ghosts' adjoints are their halo slots, which the same
// launch's gather also adds to the internal cells they read)
void ghosts_rev(Ctx& C, SolverDev& S, const AdjTopo& t, RevWork& W, const double* buf, const double* own,
                const double* oth, double* Ibar, double* Ownbar, double* Othbar, long long other_nint) {
  const bool tiled = tiled_rev(S);
  if (S.n_ghost == 0 && !tiled) return;
  const long long ng = S.n_ghost;
  if (ng > 0) {
    by_nd(S.nd, [&](auto ndc) {
      constexpr int ND = decltype(ndc)::value;
      hipLaunchKernelGGL(k_adj_ghosts<ND>, dim3(nblk(ng), ghost_inputs(S)), dim3(256), 0, C.stream, t, buf, own, oth,
                         (const double*)Ibar, W.gs.p, tiled ? (const double*)W.hs.p : nullptr,
                         tiled ? (const int32_t*)S.brick_rev_rp.p : nullptr);
    });
    H3_LAUNCH_CHECK();
  }
  const int nv = S.nvars;
  // this solver's cells and the coupled solver's partner cells: disjoint
  // outputs, one launch
  GatherJob part = no_gather();
  if (ng > 0 && Othbar && other_nint > 0 && W.maps.prp.p) {
    const int onv = (S.P.physics == HOM3G_CNS) ? 1 : S.nd + 2;
    part = gather_job(W.maps.partner(), onv, 2 * nv, Othbar, 0, nullptr, t.other_ld);
  }
  if (tiled) {
    const long long n = W.nu + part.gl.nt;
    if (n == 0) return;
    hipLaunchKernelGGL(k_adj_tiled_gather, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, C.stream, W.nu,
                       (const int32_t*)W.utl.p, (const int4*)W.urng.p, nv, (long long)S.cap, (const double*)W.hs.p,
                       (const int32_t*)W.maps.og.p, (const double*)W.gs.p, ng, Ibar, Ownbar, part);
    H3_LAUNCH_CHECK();
    return;
  }
  const GatherJob mine = gather_job(W.maps.own(), nv, 0, Ibar, nv, Ownbar, (long long)S.cap);
  gather_ghost_slots(C, mine, part, (const double*)W.gs.p, ng);
}

// the face reverse and the stage's direct terms (out = wb base + w (in +
// nf [+ dt S])): fused into one kernel on tiled bricks
// (fresh / ubar_fresh: Ibar / Ubar hold no earlier terms: written, not
// added to)
void stage_rev(Ctx& C, SolverDev& S, const AdjTopo& t, RevWork& W, const double* Q, const double* obar, double w,
               double wb, double dt, double* Ibar, double* Ubar, double* Sbar, bool fresh = false,
               bool ubar_fresh = false) {
  H3_REQUIRE(S.P.physics != HOM3G_HEAT, "the heat stage reverses with its ghost fill (heat_stage_ghosts_rev)");
  if (tiled_rev(S)) {
    // the brick reverse writes (fresh) or adds to every internal cell; the
    // halo slots go to the stage gather (ghosts_rev), and no ghost row of
    // Ibar or Ubar is read
    H3_REQUIRE(Sbar == nullptr, "no source adjoint on tiled CNS bricks");
    faces_rev(C, S, t, W, Q, obar, w, dt, Ibar, wb, Ubar, fresh, ubar_fresh);
    return;
  }
  if (fresh) zero(C, Ibar, (long long)S.nvars * S.cap);
  if (ubar_fresh) zero(C, Ubar, (long long)S.nvars * S.cap);
  faces_rev(C, S, t, W, Q, obar, w, dt, Ibar);
  cell_rev(C, S, obar, w, wb, dt, Ibar, Ubar, Sbar);
}

// whether a solver's reverse step writes the adjoint of its stage-1 input
// (Ucur) before adding to it, so the caller need not clear it: SSP-RK3 on
// the heat stage (whose last stage's reverse writes the direct term at the
// internal cells and clears the ghost rows) or on tiled CNS bricks (which
// write the internal cells and read no ghost row)
bool writes_ucur_fresh(const SolverDev& S, bool has_sbar) {
  if (S.P.ti != HOM3G_TI_SSP_RK3) return false;
  return S.P.physics == HOM3G_HEAT || (tiled_rev(S) && !has_sbar);
}

// the reverse of one heat stage with its ghost fill: the stage and the
// ghost slots in one launch (k_adj_heat_stage_ghosts), then the gather
void heat_stage_ghosts_rev(Ctx& C, SolverDev& S, const AdjTopo& t, RevWork& W, const double* buf, const double* own,
                           const double* oth, const double* obar, double w, double wb, double dt, double* Ibar,
                           double* Ubar, double* Sbar, bool fresh, bool ubar_fresh, double* Othbar,
                           long long other_nint) {
  const long long ng = S.n_ghost;
  const long long nhb = nblk(S.n_int + ng);
  const int nk = ghost_inputs(S);
  const long long nkg = ng > 0 ? nblk((long long)nk * ng) : 0;
  by_nd(S.nd, [&](auto ndc) {
    constexpr int ND = decltype(ndc)::value;
    hipLaunchKernelGGL(k_adj_heat_stage_ghosts<ND>, dim3((unsigned)(nhb + nkg)), dim3(256), 0, C.stream, t, obar, w, wb,
                       dt, Ibar, Ubar, Sbar, fresh ? 1 : 0, ubar_fresh ? 1 : 0, nhb, nk, buf, own, oth, W.gs.p);
  });
  H3_LAUNCH_CHECK();
  if (ng == 0) return;
  const int nv = S.nvars;
  const GatherJob mine = gather_job(W.maps.own(), nv, 0, Ibar, nv, Ubar, (long long)S.cap);
  GatherJob part = no_gather();
  if (Othbar && other_nint > 0 && W.maps.prp.p) part = gather_job(W.maps.partner(), S.nd + 2, 2 * nv, Othbar, 0, nullptr, t.other_ld);
  gather_ghost_slots(C, mine, part, (const double*)W.gs.p, ng);
}

void rev_solver(Ctx& C, SolverDev& S, const AdjTopo& t, RevWork& W, const RevArgs& r, long long other_nint) {
  // one stage's reverse (Q: its input, also the buffer its ghost fill read)
  // and the reverse of that ghost fill
  auto stage = [&](const double* Q, const double* obar, double w, double wb, double* Ibar, bool fresh,
                   bool ubar_fresh) {
    if (S.P.physics == HOM3G_HEAT) {
      heat_stage_ghosts_rev(C, S, t, W, Q, r.B0, r.oth, obar, w, wb, r.dt, Ibar, r.Ucur, r.Sbar, fresh, ubar_fresh,
                            r.Othbar, other_nint);
      return;
    }
    stage_rev(C, S, t, W, Q, obar, w, wb, r.dt, Ibar, r.Ucur, r.Sbar, fresh, ubar_fresh);
    ghosts_rev(C, S, t, W, Q, r.B0, r.oth, Ibar, r.Ucur, r.Othbar, other_nint);
  };
  if (S.P.ti == HOM3G_TI_SSP_RK3) {
    // I: the stage input's adjoint being formed; O: the stage output's
    // (the previous I: the buffers swap instead of copying)
    double* I = r.I;
    double* O = r.O;
    stage(r.B2, r.Unext, 2. / 3., 1. / 3., I, true, r.ucur_fresh);
    std::swap(I, O);
    stage(r.B1, O, 0.25, 0.75, I, true, false);
    std::swap(I, O);
    stage(r.B0, O, 1.0, 0.0, r.Ucur, false, false);
  } else {
    stage(r.B0, r.Unext, 1.0, 0.0, r.Ucur, false, false);
  }
}

// the dt each solver took in the step just run (set_dt may clamp the forced
// dt at the end time), kept for the reverse sweep
__global__ void k_record_dt(const DevScalars* cs, const DevScalars* hs, double* out) {
  out[0] = cs->dt;
  out[1] = hs->dt;
}

}  // namespace

// ---- drivers -----------------------------------------------------------------

void cht_objective_run(Ctx& C, SolverDev& cns, SolverDev& heat, double Tr0, int nsteps, const double* dts,
                       double* dts_out, ChtRunStats& st, const double* obs) {
  check_pair(cns, heat);
  const double scaling = scaling_of(Tr0, cns);
  const AdjTopo th = make_adj_topo(heat, &cns);
  const Obs ob(C, heat, obs);
  DevBuf<double> acc, term;
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
    if (!ob.p) objective_launch(C, heat, cns, th, nullptr, nullptr, term, acc.p);
    if (dts_out) {
      DevScalars s;
      H3_HIP(hipMemcpyAsync(&s, cns.scal.p, sizeof(DevScalars), hipMemcpyDeviceToHost, C.stream));
      H3_HIP(hipStreamSynchronize(C.stream));
      dts_out[n] = s.dt;
    }
  }
  if (ob.p) misfit_launch(C, heat, ob.p, nullptr, term, acc.p);
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
  DevBuf<double> ds, zsrc, acc, term;
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
    if (!ob.p) objective_launch(C, heat, cns, th, Th.L.p, Tc.L.p, term, acc.p);
  }
  if (ob.p) misfit_launch(C, heat, ob.p, Th.L.p, term, acc.p);
  cns.dtmin_current = heat.dtmin_current = false;
  double h[2];
  acc.download(h, 2, C.stream);
  H3_HIP(hipStreamSynchronize(C.stream));
  *J = h[0];
  *dJ = h[1];
}

// Binomial checkpointing (the schedule of Griewank's revolve): at most
// `slots` saved states of both solvers besides the initial one, and one
// step's stage buffers.  A first forward sweep accumulates J and records the
// dts the solvers take; the reverse sweep then reverses [0, N) recursively:
// for a range [a, b) whose start state is saved, with s free slots, advance
// to the split point m, save it, reverse [m, b) with s - 1 slots, restore a
// and reverse [a, m) with s slots.  m is chosen from the binomial counts
// beta(s, t) = C(s + t, s): for the smallest t with beta(s, t) >= b - a,
// m - a = max(1, (b - a) - beta(s - 1, t)), so that each part fits the
// repetitions left to it.  With no free slot the range is reversed by
// recomputing from its start for every step.  Every step is recomputed from
// saved states by the same kernels, so the gradient equals the one with
// every step stored, bit for bit.
static double binom_beta(int s, int t) {
  double r = 1.0;
  for (int i = 1; i <= s; ++i) r = r * (double)(t + i) / (double)i;
  return r;
}

static int binomial_split(int n, int s) {
  int t = 1;
  while (binom_beta(s, t) < (double)n) ++t;
  const double right = s >= 1 ? binom_beta(s - 1, t) : 1.0;
  const double m = (double)n - right;
  return (int)std::max(1.0, std::min((double)(n - 1), m));
}

void cht_adjoint_binomial(Ctx& C, SolverDev& cns, SolverDev& heat, double Tr0, int nsteps, const double* dts,
                          int slots, double* grad_source, ChtRunStats& st, const double* obs) {
  H3_REQUIRE(slots >= 1 && slots <= 4096, "binomial checkpointing needs 1 to 4096 snapshot slots");
  const double scaling = scaling_of(Tr0, cns);
  const AdjTopo tc = make_adj_topo(cns, &heat), th = make_adj_topo(heat, &cns);
  const bool rk3 = cns.P.ti == HOM3G_TI_SSP_RK3;
  const long long nc = (long long)cns.nvars * cns.cap, nh = (long long)heat.nvars * heat.cap;
  const int nslot = std::min(slots, std::max(nsteps - 1, 1));
  {
    size_t free_b = 0, total_b = 0;
    H3_HIP(hipMemGetInfo(&free_b, &total_b));
    const double need = (double)(nslot + 1 + (rk3 ? 3 : 1) + 10) * (double)(nc + nh) * sizeof(double);
    if (need > 0.9 * (double)free_b) throw Error(HOM3G_ERR_CAPACITY, "binomial checkpoints do not fit in device memory");
  }
  // stage buffers of the step being reversed
  DevBuf<double> C0, C1, C2, H0, H1, H2;
  C0.alloc(nc);
  H0.alloc(nh);
  if (rk3) { C1.alloc(nc); C2.alloc(nc); H1.alloc(nh); H2.alloc(nh); }
  // snapshot 0: the initial state; 1..nslot: the checkpoints
  std::vector<DevBuf<double>> SC(nslot + 1), SH(nslot + 1);
  for (int k = 0; k <= nslot; ++k) { SC[k].alloc(nc); SH[k].alloc(nh); }
  DevBuf<DevScalars> SS;
  SS.alloc(2 * (size_t)(nslot + 1));
  auto save = [&](int k) {
    copyd(C, cns.lhs.p, SC[k].p, nc);
    copyd(C, heat.lhs.p, SH[k].p, nh);
    H3_HIP(hipMemcpyAsync(SS.p + 2 * k, cns.scal.p, sizeof(DevScalars), hipMemcpyDeviceToDevice, C.stream));
    H3_HIP(hipMemcpyAsync(SS.p + 2 * k + 1, heat.scal.p, sizeof(DevScalars), hipMemcpyDeviceToDevice, C.stream));
  };
  auto restore = [&](int k) {
    copyd(C, SC[k].p, cns.lhs.p, nc);
    copyd(C, SH[k].p, heat.lhs.p, nh);
    H3_HIP(hipMemcpyAsync(cns.scal.p, SS.p + 2 * k, sizeof(DevScalars), hipMemcpyDeviceToDevice, C.stream));
    H3_HIP(hipMemcpyAsync(heat.scal.p, SS.p + 2 * k + 1, sizeof(DevScalars), hipMemcpyDeviceToDevice, C.stream));
    cns.dtmin_current = heat.dtmin_current = false;
  };
  auto advance = [&](int n) {
    hipLaunchKernelGGL(k_force_pair, dim3(1), dim3(1), 0, C.stream, cns.scal.p, heat.scal.p, dts[n], scaling);
    H3_LAUNCH_CHECK();
    cns.dtmin_current = heat.dtmin_current = false;
    solver_step(C, cns, true);
    solver_step(C, heat, true);
  };
  // forward sweep: J and the dts taken
  DevBuf<double> dtrec, acc, term;
  dtrec.alloc(2 * (size_t)std::max(nsteps, 1));
  acc.alloc(1);
  zero(C, acc.p, 1);
  const Obs ob(C, heat, obs);
  save(0);
  H3_HIP(hipEventRecord(C.ev0, C.stream));
  for (int n = 0; n < nsteps; ++n) {
    advance(n);
    hipLaunchKernelGGL(k_record_dt, dim3(1), dim3(1), 0, C.stream, (const DevScalars*)cns.scal.p,
                       (const DevScalars*)heat.scal.p, dtrec.p + 2 * n);
    H3_LAUNCH_CHECK();
    if (!ob.p) objective_launch(C, heat, cns, th, nullptr, nullptr, term, acc.p);
  }
  if (ob.p) misfit_launch(C, heat, ob.p, nullptr, term, acc.p);
  H3_HIP(hipEventRecord(C.ev1, C.stream));
  H3_HIP(hipEventSynchronize(C.ev1));
  H3_HIP(hipEventElapsedTime(&st.ms_forward, C.ev0, C.ev1));
  st.ms_primal_step = nsteps ? st.ms_forward / nsteps : 0.f;
  acc.download(&st.J, 1, C.stream);
  std::vector<double> dt_used(2 * (size_t)std::max(nsteps, 1));
  dtrec.download(dt_used.data(), dt_used.size(), C.stream);
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
  RevWork Wc, Wh;
  Wc.alloc(C, cns, &heat);
  Wh.alloc(C, heat, &cns);
  DevBuf<double> Cn1, Cn, Hn1, Hn, Ic, Oc, Ih, Oh, Sbar, Cend, Hend, OS;
  for (DevBuf<double>* b : {&Cn1, &Cn, &Ic, &Oc, &Cend}) b->alloc(nc);
  for (DevBuf<double>* b : {&Hn1, &Hn, &Ih, &Oh, &Sbar, &Hend}) b->alloc(nh);
  OS.alloc(std::max<size_t>(1, (size_t)(heat.nd + 3) * heat.n_ghost));
  zero(C, Cn1.p, nc);
  zero(C, Hn1.p, nh);
  zero(C, Sbar.p, nh);
  // Cn / Hn: cleared per step, or (fresh_c / fresh_h) written by the
  // reverse step before it adds to them; its rows past the ghosts, which no
  // step writes, stay as cleared here
  const bool fresh_c = writes_ucur_fresh(cns, false), fresh_h = writes_ucur_fresh(heat, true);
  zero(C, Cn.p, nc);
  zero(C, Hn.p, nh);
  if (ob.p) {
    hipLaunchKernelGGL(k_misfit_seed, dim3(nblk(heat.n_int)), dim3(256), 0, C.stream, (long long)heat.n_int,
                       (const double*)finH.p, ob.p, Hn1.p);
    H3_LAUNCH_CHECK();
  }
  // step n from the state at n held by the solvers: record its stage inputs,
  // keep the state at n + 1, reverse it
  auto reverse_one = [&](int n) {
    hipLaunchKernelGGL(k_force_pair, dim3(1), dim3(1), 0, C.stream, cns.scal.p, heat.scal.p, dts[n], scaling);
    H3_LAUNCH_CHECK();
    cns.dtmin_current = heat.dtmin_current = false;
    solver_fill_ghosts(C, cns, cns.lhs.p);
    copyd(C, cns.lhs.p, C0.p, nc);
    solver_step(C, cns, true);
    if (rk3) { copyd(C, cns.rhs.p, C1.p, nc); copyd(C, cns.stg.p, C2.p, nc); }
    solver_fill_ghosts(C, heat, heat.lhs.p);
    copyd(C, heat.lhs.p, H0.p, nh);
    solver_step(C, heat, true);
    if (rk3) { copyd(C, heat.rhs.p, H1.p, nh); copyd(C, heat.stg.p, H2.p, nh); }
    copyd(C, cns.lhs.p, Cend.p, nc);
    copyd(C, heat.lhs.p, Hend.p, nh);
    const double dtc = dt_used[2 * n], dth = dt_used[2 * n + 1];
    if (heat.n_ghost > 0 && !ob.p) {
      by_nd(heat.nd, [&](auto ndc) {
        constexpr int ND = decltype(ndc)::value;
        hipLaunchKernelGGL(k_objective_adj<ND>, dim3(nblk(heat.n_ghost)), dim3(256), 0, C.stream, th,
                           (const double*)Hend.p, (const double*)Cend.p, dth, OS.p);
      });
      H3_LAUNCH_CHECK();
      gather_ghost_slots(C, gather_job(Wh.maps.own(), 1, 0, Hn1.p, 0, nullptr, (long long)heat.cap),
                         gather_job(Wh.maps.partner(), cns.nvars, 1, Cn1.p, 0, nullptr, (long long)cns.cap),
                         (const double*)OS.p, (long long)heat.n_ghost);
    }
    if (!fresh_h) zero(C, Hn.p, nh);
    RevArgs rh{H0.p, rk3 ? H1.p : nullptr, rk3 ? H2.p : nullptr, Cend.p, Hn1.p, Hn.p, Cn1.p, Sbar.p,
               Ih.p, Oh.p, dth, fresh_h};
    rev_solver(C, heat, th, Wh, rh, cns.n_int);
    if (!fresh_c) zero(C, Cn.p, nc);
    RevArgs rc{C0.p, rk3 ? C1.p : nullptr, rk3 ? C2.p : nullptr, H0.p, Cn1.p, Cn.p, Hn.p, nullptr,
               Ic.p, Oc.p, dtc, fresh_c};
    rev_solver(C, cns, tc, Wc, rc, heat.n_int);
    std::swap(Cn1.p, Cn.p);
    std::swap(Hn1.p, Hn.p);
  };
  // [a, b) with the state at a saved in slot `base` and held by the solvers;
  // free slots base + 1 .. nslot
  std::function<void(int, int, int)> rev = [&](int a, int b, int base) {
    const int n = b - a, s = nslot - base;
    if (n == 1) {
      reverse_one(a);
      return;
    }
    if (s == 0) {
      for (int k = b - 1; k >= a; --k) {
        if (k != b - 1) restore(base);
        for (int i = a; i < k; ++i) advance(i);
        reverse_one(k);
      }
      return;
    }
    const int m = a + binomial_split(n, s);
    for (int i = a; i < m; ++i) advance(i);
    save(base + 1);
    rev(m, b, base + 1);
    restore(base);
    rev(a, m, base);
  };
  H3_HIP(hipEventRecord(C.ev0, C.stream));
  if (nsteps > 0) {
    restore(0);
    rev(0, nsteps, 0);
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

void cht_adjoint_run(Ctx& C, SolverDev& cns, SolverDev& heat, double Tr0, int nsteps, const double* dts,
                     int ckpt_every, double* grad_source, ChtRunStats& st, const double* obs) {
  check_pair(cns, heat);
  H3_REQUIRE(dts != nullptr, "the adjoint needs a dt schedule");
  if (ckpt_every < 0) {
    cht_adjoint_binomial(C, cns, heat, Tr0, nsteps, dts, -ckpt_every, grad_source, st, obs);
    return;
  }
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
  // the dts both solvers took (set_dt clamps the forced dt at the end time)
  DevBuf<double> dtrec;
  dtrec.alloc(2 * (size_t)std::max(nsteps, 1));
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

  // forward sweep: checkpoints (state and scalars) every K steps (in device
  // memory, or in pinned host memory when asked or when device memory is
  // short), or every stage input
  bool host_ck = false;
  if (!store_all) {
    size_t free_b = 0, total_b = 0;
    H3_HIP(hipMemGetInfo(&free_b, &total_b));
    const double need = (double)nseg * (double)(nc + nh) * sizeof(double);
    host_ck = C.ckpt_memory == HOM3G_CKPT_HOST || (C.ckpt_memory == HOM3G_CKPT_AUTO && need > 0.8 * (double)free_b);
  }
  std::vector<DevBuf<double>> ckC(store_all || host_ck ? 0 : nseg), ckH(store_all || host_ck ? 0 : nseg);
  std::vector<PinnedBuf<double>> hkC(host_ck ? nseg : 0), hkH(host_ck ? nseg : 0);
  DevBuf<DevScalars> ckS;
  ckS.alloc(2 * (size_t)std::max(nseg, 1));
  DevBuf<double> acc, term;
  acc.alloc(1);
  zero(C, acc.p, 1);
  const Obs ob(C, heat, obs);
  H3_HIP(hipEventRecord(C.ev0, C.stream));
  for (int n = 0; n < nsteps; ++n) {
    if (store_all) {
      record_step(n, n);
    } else {
      if (n % K == 0) {
        H3_HIP(hipMemcpyAsync(ckS.p + 2 * (n / K), cns.scal.p, sizeof(DevScalars), hipMemcpyDeviceToDevice, C.stream));
        H3_HIP(hipMemcpyAsync(ckS.p + 2 * (n / K) + 1, heat.scal.p, sizeof(DevScalars), hipMemcpyDeviceToDevice,
                              C.stream));
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
    hipLaunchKernelGGL(k_record_dt, dim3(1), dim3(1), 0, C.stream, (const DevScalars*)cns.scal.p,
                       (const DevScalars*)heat.scal.p, dtrec.p + 2 * n);
    H3_LAUNCH_CHECK();
    if (!ob.p) objective_launch(C, heat, cns, th, nullptr, nullptr, term, acc.p);
  }
  if (ob.p) misfit_launch(C, heat, ob.p, nullptr, term, acc.p);
  H3_HIP(hipEventRecord(C.ev1, C.stream));
  H3_HIP(hipEventSynchronize(C.ev1));
  H3_HIP(hipEventElapsedTime(&st.ms_forward, C.ev0, C.ev1));
  st.ms_primal_step = nsteps ? st.ms_forward / nsteps : 0.f;
  acc.download(&st.J, 1, C.stream);
  std::vector<double> dt_used(2 * (size_t)std::max(nsteps, 1));
  dtrec.download(dt_used.data(), dt_used.size(), C.stream);
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
  RevWork Wc, Wh;
  Wc.alloc(C, cns, &heat);
  Wh.alloc(C, heat, &cns);
  DevBuf<double> Cn1, Cn, Hn1, Hn, Ic, Oc, Ih, Oh, Sbar, Cend, Hend, OS;
  for (DevBuf<double>* b : {&Cn1, &Cn, &Ic, &Oc}) b->alloc(nc);
  for (DevBuf<double>* b : {&Hn1, &Hn, &Ih, &Oh, &Sbar}) b->alloc(nh);
  OS.alloc(std::max<size_t>(1, (size_t)(heat.nd + 3) * heat.n_ghost));
  if (!store_all) {
    Cend.alloc(nc);
    Hend.alloc(nh);
  }
  const double* CendP = store_all ? finC.p : Cend.p;
  const double* HendP = store_all ? finH.p : Hend.p;
  zero(C, Cn1.p, nc);
  zero(C, Hn1.p, nh);
  zero(C, Sbar.p, nh);
  // Cn / Hn: cleared per step, or (fresh_c / fresh_h) written by the
  // reverse step before it adds to them; its rows past the ghosts, which no
  // step writes, stay as cleared here
  const bool fresh_c = writes_ucur_fresh(cns, false), fresh_h = writes_ucur_fresh(heat, true);
  zero(C, Cn.p, nc);
  zero(C, Hn.p, nh);
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
      H3_HIP(hipMemcpyAsync(cns.scal.p, ckS.p + 2 * sg, sizeof(DevScalars), hipMemcpyDeviceToDevice, C.stream));
      H3_HIP(hipMemcpyAsync(heat.scal.p, ckS.p + 2 * sg + 1, sizeof(DevScalars), hipMemcpyDeviceToDevice, C.stream));
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
      const double dtc = dt_used[2 * n], dth = dt_used[2 * n + 1];
      const double* Cnext = (j + 1 < len) ? C0[j + 1].p : CendP;
      const double* Hnext = (j + 1 < len) ? H0[j + 1].p : HendP;
      if (heat.n_ghost > 0 && !ob.p) {
        by_nd(heat.nd, [&](auto ndc) {
          constexpr int ND = decltype(ndc)::value;
          hipLaunchKernelGGL(k_objective_adj<ND>, dim3(nblk(heat.n_ghost)), dim3(256), 0, C.stream, th, Hnext, Cnext,
                             dth, OS.p);
        });
        H3_LAUNCH_CHECK();
        gather_ghost_slots(C, gather_job(Wh.maps.own(), 1, 0, Hn1.p, 0, nullptr, (long long)heat.cap),
                           gather_job(Wh.maps.partner(), cns.nvars, 1, Cn1.p, 0, nullptr, (long long)cns.cap),
                           (const double*)OS.p, (long long)heat.n_ghost);
      }
      if (!fresh_h) zero(C, Hn.p, nh);
      RevArgs rh{H0[j].p, rk3 ? H1[j].p : nullptr, rk3 ? H2[j].p : nullptr, Cnext, Hn1.p, Hn.p, Cn1.p, Sbar.p,
                 Ih.p, Oh.p, dth, fresh_h};
      rev_solver(C, heat, th, Wh, rh, cns.n_int);
      if (!fresh_c) zero(C, Cn.p, nc);
      RevArgs rc{C0[j].p, rk3 ? C1[j].p : nullptr, rk3 ? C2[j].p : nullptr, H0[j].p, Cn1.p, Cn.p, Hn.p, nullptr,
                 Ic.p, Oc.p, dtc, fresh_c};
      rev_solver(C, cns, tc, Wc, rc, heat.n_int);
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
