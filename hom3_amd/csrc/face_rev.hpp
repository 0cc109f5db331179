// Reverse mode of the tiled brick's CNS face flux.
//
// The forward is brick.hip brick_face for CNS without MUSCL: per LDS
// position p = (g-1)(rho E - |rho u|^2 / (2 rho)) and 1/(rho a) =
// (g p rho)^(-1/2) (brick_prims), AUSM with those (ausm_flux_fast), and the
// viscous flux from unscaled differences (viscous_flux_fast): the normal
// difference q_R - q_L and, per tangential direction, the sum of the two
// cells' central differences.  The reverse below differentiates exactly
// that arrangement, one forward recomputation of the intermediates and then
// their adjoints in reverse order; piecewise formulas take the branch of the
// forward values.  The reference's algorithm (AUSM, euler/physics.hpp:276-325;
// chain-rule viscous flux with Sutherland viscosity, cns/physics.hpp:221-338)
// is the same function; only the arithmetic arrangement differs.
#pragma once
#include "physics.hpp"

namespace hom3g {

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

// qb += the adjoint of (p, ia)(q) for the adjoints pb, iab.  With
// ia = (g p rho)^(-1/2): 1/rho = g p ia^2 and 1/p = g rho ia^2, so no
// division is needed; d ia/dp = -ia/(2p), d ia/drho = -ia/(2 rho).
template <int ND>
__device__ __forceinline__ void brick_prims_rev(const double* q, double p, double ia, double gamma, double gm1,
                                                double pb, double iab, double* qb) {
  const double ia2 = ia * ia;
  const double ir = gamma * p * ia2;
  const double ip = gamma * q[ND] * ia2;
  pb -= 0.5 * ia * ip * iab;
  double rhob = -0.5 * ia * ir * iab;
  // p = (g-1)(rho E - |m|^2 / (2 rho)), u = m / rho
  qb[ND + 1] += gm1 * pb;
  double um2 = 0.0;
#pragma unroll
  for (int k = 0; k < ND; ++k) {
    const double u = q[k] * ir;
    qb[k] -= gm1 * u * pb;
    um2 += u * u;
  }
  rhob += 0.5 * gm1 * um2 * pb;
  qb[ND] += rhob;
}

// Reverse of the face flux along D between L and R.  In: the two states,
// their p and 1/(rho a), Gt[j][t] the tangential difference sums of the
// variables the flux reads (t over the directions other than D in
// increasing order; j = 0 the momentum along D, 1 the momentum along the
// tangential direction, 2 rho), 1/h and the flux adjoint Fb.  Out (set):
// qLb, qRb (including the paths through p and 1/(rho a)) and Gtb[j][t].
template <int ND, int D>
__device__ __forceinline__ void brick_face_rev(const PhysConst& pc, const double* qL, const double* qR, double pL,
                                               double pR, double iaL, double iaR, const double (*Gt)[ND - 1],
                                               double inv_h, const double* Fb, double* qLb, double* qRb,
                                               double (*Gtb)[ND - 1]) {
  constexpr int NV = ND + 2, RHO = ND, E = ND + 1;
  const double gamma = pc.gamma, gm1 = pc.gm1;
#pragma unroll
  for (int v = 0; v < NV; ++v) qLb[v] = qRb[v] = 0.0;
  double pLb = 0.0, pRb = 0.0, iaLb = 0.0, iaRb = 0.0;

  // ---- AUSM: f[k] = q_up[k] a_up m_i (k <= ND), f[E] = (q_up[E] + p_up) a_up m_i, f[D] += p_i
  const double aL = gamma * pL * iaL, aR = gamma * pR * iaR;
  const double ML = qL[D] * iaL, MR = qR[D] * iaR;
  const double m_i = m_int<+1>(ML) + m_int<-1>(MR);
  const bool left = m_i >= 0.0;
  const double a_up = left ? aL : aR;
  const double p_up = left ? pL : pR;
  const double am = a_up * m_i;
  double amb = 0.0;
#pragma unroll
  for (int k = 0; k < ND + 1; ++k) {
    amb += Fb[k] * (left ? qL[k] : qR[k]);
    const double x = Fb[k] * am;
    qLb[k] += left ? x : 0.0;
    qRb[k] += left ? 0.0 : x;
  }
  {
    amb += Fb[E] * ((left ? qL[E] : qR[E]) + p_up);
    const double x = Fb[E] * am;
    qLb[E] += left ? x : 0.0;
    qRb[E] += left ? 0.0 : x;
    pLb += left ? x : 0.0;
    pRb += left ? 0.0 : x;
  }
  const double a_upb = amb * m_i, m_ib = amb * a_up;
  // a = g p ia
  pLb += left ? gamma * iaL * a_upb : 0.0;
  iaLb += left ? gamma * pL * a_upb : 0.0;
  pRb += left ? 0.0 : gamma * iaR * a_upb;
  iaRb += left ? 0.0 : gamma * pR * a_upb;
  const double pib = Fb[D];
  pLb += pib * p_int_fast<+1>(ML);
  pRb += pib * p_int_fast<-1>(MR);
  const double MLb = pib * pL * dp_int<+1>(ML) + m_ib * dm_int<+1>(ML);
  const double MRb = pib * pR * dp_int<-1>(MR) + m_ib * dm_int<-1>(MR);
  qLb[D] += MLb * iaL;
  iaLb += MLb * qL[D];
  qRb[D] += MRb * iaR;
  iaRb += MRb * qR[D];

  // ---- viscous flux: forward intermediates (viscous_flux_fast)
  const double S = qL[RHO] + qR[RHO];
  const double iS = dev_rcp(S);
  const double rho_s = 0.5 * S, irs = 2.0 * iS, irs4 = 0.5 * iS;
  double u_s[ND], umag2 = 0.0;
#pragma unroll
  for (int k = 0; k < ND; ++k) {
    u_s[k] = (qL[k] + qR[k]) * iS;
    umag2 += u_s[k] * u_s[k];
  }
  const double p_s = 0.5 * gm1 * ((qL[E] + qR[E]) - rho_s * umag2);
  double dn[NV];
#pragma unroll
  for (int v = 0; v < NV; ++v) dn[v] = qR[v] - qL[v];
  double gcol[ND], grow[ND], diag[ND];
#pragma unroll
  for (int k = 0; k < ND; ++k) gcol[k] = (dn[k] - dn[RHO] * u_s[k]) * irs;
#pragma unroll
  for (int k = 0; k < ND; ++k) {
    if (k == D) {
      grow[k] = diag[k] = gcol[D];
    } else {
      const int t = k < D ? k : k - 1;
      grow[k] = (Gt[0][t] - Gt[2][t] * u_s[D]) * irs4;
      diag[k] = (Gt[1][t] - Gt[2][t] * u_s[k]) * irs4;
    }
  }
  double gru = dn[RHO] * umag2;
#pragma unroll
  for (int k = 0; k < ND; ++k) gru += S * u_s[k] * gcol[k];
  const double gradP = gm1 * (dn[E] - 0.5 * gru);
  const double Tm = gamma * (pL + pR) * iS;
  const double sq = fast_sqrt(Tm);
  const double iTs = dev_rcp(Tm + pc.sOverTref);
  const double mu = Tm * sq * pc.sOverTrefP1 * iTs;
  const double mh = mu * inv_h;
  const double Y = gradP - p_s * irs * dn[RHO];
  const double Kq = -mh * pc.q_coef * irs;  // heat flux = Kq Y
  double div = 0.0;
#pragma unroll
  for (int k = 0; k < ND; ++k) div += diag[k];
  div = div * (2.0 / 3.0);

  // ---- viscous flux: reverse
  const double iRe = pc.inv_Re0;
  const double fEb = iRe * Fb[E];
  double u_sb[ND], gcolb[ND], growb[ND], diagb[ND];
#pragma unroll
  for (int k = 0; k < ND; ++k) u_sb[k] = gcolb[k] = growb[k] = diagb[k] = 0.0;
  double mhb = 0.0, divb = 0.0;
#pragma unroll
  for (int i = 0; i < ND; ++i) {
    const double Si = grow[i] + gcol[i] - (i == D ? div : 0.0);
    const double tb = iRe * Fb[i] + fEb * u_s[i];
    u_sb[i] -= fEb * mh * Si;
    mhb -= tb * Si;
    const double Sb = -mh * tb;
    growb[i] += Sb;
    gcolb[i] += Sb;
    if (i == D) divb -= Sb;
  }
#pragma unroll
  for (int k = 0; k < ND; ++k) diagb[k] += (2.0 / 3.0) * divb;
  const double Yb = fEb * Kq;
  mhb -= fEb * Y * pc.q_coef * irs;
  double irsb = -fEb * Y * mh * pc.q_coef;
  const double gradPb = Yb;
  double p_sb = -Yb * irs * dn[RHO];
  irsb -= Yb * p_s * dn[RHO];
  double dnb[NV];
#pragma unroll
  for (int v = 0; v < NV; ++v) dnb[v] = 0.0;
  dnb[RHO] -= Yb * p_s * irs;
  // mu = Tm^1.5 c1 / (Tm + c0): d mu / d Tm = c1 sqrt(Tm) / (Tm + c0) (1.5 - Tm / (Tm + c0))
  const double Tmb = mhb * inv_h * pc.sOverTrefP1 * sq * iTs * (1.5 - Tm * iTs);
  pLb += Tmb * gamma * iS;
  pRb += Tmb * gamma * iS;
  double iSb = Tmb * gamma * (pL + pR);
  dnb[E] += gm1 * gradPb;
  const double grub = -0.5 * gm1 * gradPb;
  dnb[RHO] += grub * umag2;
  double umag2b = grub * dn[RHO];
  double Sb = 0.0;
#pragma unroll
  for (int k = 0; k < ND; ++k) {
    Sb += grub * u_s[k] * gcol[k];
    u_sb[k] += grub * S * gcol[k];
    gcolb[k] += grub * S * u_s[k];
  }
  gcolb[D] += growb[D] + diagb[D];
  double irs4b = 0.0;
#pragma unroll
  for (int k = 0; k < ND; ++k) {
    if (k == D) continue;
    const int t = k < D ? k : k - 1;
    const double gw = growb[k] * irs4, dg = diagb[k] * irs4;
    Gtb[0][t] = gw;
    Gtb[1][t] = dg;
    Gtb[2][t] = -gw * u_s[D] - dg * u_s[k];
    u_sb[D] -= gw * Gt[2][t];
    u_sb[k] -= dg * Gt[2][t];
    irs4b += growb[k] * (Gt[0][t] - Gt[2][t] * u_s[D]) + diagb[k] * (Gt[1][t] - Gt[2][t] * u_s[k]);
  }
#pragma unroll
  for (int k = 0; k < ND; ++k) {
    const double g = gcolb[k] * irs;
    dnb[k] += g;
    dnb[RHO] -= g * u_s[k];
    u_sb[k] -= g * dn[RHO];
    irsb += gcolb[k] * (dn[k] - dn[RHO] * u_s[k]);
  }
  const double EEb = 0.5 * gm1 * p_sb;
  const double rho_sb = -0.5 * gm1 * umag2 * p_sb;
  umag2b -= 0.5 * gm1 * rho_s * p_sb;
#pragma unroll
  for (int k = 0; k < ND; ++k) {
    u_sb[k] += 2.0 * u_s[k] * umag2b;
    const double x = u_sb[k] * iS;
    qLb[k] += x;
    qRb[k] += x;
    iSb += u_sb[k] * (qL[k] + qR[k]);
  }
  iSb += 2.0 * irsb + 0.5 * irs4b;
  Sb += 0.5 * rho_sb - iSb * iS * iS;
  qLb[RHO] += Sb;
  qRb[RHO] += Sb;
  qLb[E] += EEb;
  qRb[E] += EEb;
#pragma unroll
  for (int v = 0; v < NV; ++v) {
    qRb[v] += dnb[v];
    qLb[v] -= dnb[v];
  }
  // ---- p and 1/(rho a) of both sides
  brick_prims_rev<ND>(qL, pL, iaL, gamma, gm1, pLb, iaLb, qLb);
  brick_prims_rev<ND>(qR, pR, iaR, gamma, gm1, pRb, iaRb, qRb);
}

}  // namespace hom3g
