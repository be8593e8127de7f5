// hlld_fast.hpp -- HLLD flux for an axis-aligned face, in the face frame.
//
// Same Riemann solver as chlld's n_state = 3 branch with Davis wave speeds
// (SET_Fortran_Cuda_MHD/Solvers.f90:1316-1319, 1598-1855): the same star and
// double-star states, the same |tt| <= 1e-9 and |Bn| <= 1e-6 switches, the
// same region tests with the later test winning on ties, and the same
// normal-B flux -max(|SL|,|SR|)(BnR - BnL) (:1828-1839).  It differs from the
// reference in arithmetic only:
//   * the face normal is a grid axis, so the tangential basis is the other
//     two axes instead of the B-aligned basis built at :1235-1273 (the flux is
//     equivariant under rotations about the normal, so only rounding differs);
//   * each repeated divisor is inverted once (frcp) and multiplied;
//   * fast speeds arrive precomputed per cell and per axis (cf_cell below);
//   * B is carried in Heaviside-Lorentz units b = B/sqrt(4 pi), as chlld uses
//     internally, so no /sqrt(4 pi) and *sqrt(4 pi) per face.
// Templated on R = double (FP64 form) or float (mixed form's star states).
#pragma once
#include "common.hpp"

namespace setmhd {

// Face-frame state of one side: normal n, tangents t1, t2.
template <class R>
struct Side {
  R rho, vn, vt1, vt2, bn, bt1, bt2, pt, e, cf;
};

// Fast magnetosonic speed along an axis with normal field component bd:
// cf^2 = (s + sqrt(s^2 - 4 c^2 bd^2/rho)) / 2, s = c^2 + |b|^2/rho -- the
// closed form of chlld's (qp + qm)/2 (:1282-1288).
template <class R>
__device__ __forceinline__ R cf_cell(R c2, R s, R bd2_over_rho) {
  R disc = s * s - R(4) * c2 * bd2_over_rho;
  disc = disc > R(0) ? disc : R(0);
  return fsqrt(R(0.5) * (s + fsqrt(disc)));
}

// Physical flux of one side in the face frame (:1377-1393).
template <class R>
__device__ __forceinline__ void phys_flux(const Side<R>& s, R F[8]) {
  R mv = s.rho * s.vn;
  R vb = s.vn * s.bn + s.vt1 * s.bt1 + s.vt2 * s.bt2;
  F[0] = mv;
  F[1] = mv * s.vn + s.pt - s.bn * s.bn;
  F[2] = mv * s.vt1 - s.bn * s.bt1;
  F[3] = mv * s.vt2 - s.bn * s.bt2;
  F[4] = (s.e + s.pt) * s.vn - s.bn * vb;
  F[5] = R(0);
  F[6] = s.vn * s.bt1 - s.vt1 * s.bn;
  F[7] = s.vn * s.bt2 - s.vt2 * s.bn;
}

// HLLD flux F (face frame) and max(|SL|,|SR|) for the CFL estimate.
// Returns false where no wave region matches (NaN input; the reference
// STOPs there, :1816-1825).
template <class R>
__device__ __forceinline__ bool hlld(const Side<R>& L, const Side<R>& Rr, R F[8], R& smax) {
  const R SL = fmin(L.vn - L.cf, Rr.vn - Rr.cf);  // :1316-1319
  const R SR = fmax(L.vn + L.cf, Rr.vn + Rr.cf);
  const R aSL = fabs(SL), aSR = fabs(SR);
  smax = aSL > aSR ? aSL : aSR;
  const R fbn = -(smax * (Rr.bn - L.bn));  // :1828-1839 (w = 0)

  if (SL > R(0) || SR < R(0)) {  // regions 1 and 6: upwind
    const bool up = SL > R(0);
    Side<R> U;
    U.rho = up ? L.rho : Rr.rho;
    U.vn = up ? L.vn : Rr.vn;
    U.vt1 = up ? L.vt1 : Rr.vt1;
    U.vt2 = up ? L.vt2 : Rr.vt2;
    U.bn = up ? L.bn : Rr.bn;
    U.bt1 = up ? L.bt1 : Rr.bt1;
    U.bt2 = up ? L.bt2 : Rr.bt2;
    U.pt = up ? L.pt : Rr.pt;
    U.e = up ? L.e : Rr.e;
    phys_flux(U, F);
    F[5] = fbn;
    return true;
  }
  const R suR = SR - Rr.vn, suL = SL - L.vn;
  const R mR = suR * Rr.rho, mL = suL * L.rho;
  const R idn = frcp(mR - mL);
  const R SM = (mR * Rr.vn - Rr.pt + L.pt - mL * L.vn) * idn;            // :1343
  const R ptz = (mR * L.pt - mL * Rr.pt + mR * mL * (Rr.vn - L.vn)) * idn;  // :1600
  const R bn = (SR * Rr.bn - SL * L.bn) * frcp(SR - SL);                  // UZ(6)
  const R bn2 = bn * bn;
  const R idR = frcp(SR - SM), idL = frcp(SL - SM);
  const R rzR = mR * idR, rzL = mL * idL;  // r*su/(S-SM)

  // single-star states (:1624-1656)
  R tvL = R(0), tbL = R(0), tvR = R(0), tbR = R(0);
  {
    R tt = mL * (SL - SM) - bn2;
    if (fabs(tt) > R(1e-9)) {
      R it = frcp(tt);
      tvL = (SM - L.vn) * it;
      tbL = (mL * suL - bn2) * it;
    }
    tt = mR * (SR - SM) - bn2;
    if (fabs(tt) > R(1e-9)) {
      R it = frcp(tt);
      tvR = (SM - Rr.vn) * it;
      tbR = (mR * suR - bn2) * it;
    }
  }
  const R vL1 = L.vt1 - bn * L.bt1 * tvL, vL2 = L.vt2 - bn * L.bt2 * tvL;
  const R vR1 = Rr.vt1 - bn * Rr.bt1 * tvR, vR2 = Rr.vt2 - bn * Rr.bt2 * tvR;
  const R bL1 = L.bt1 * tbL, bL2 = L.bt2 * tbL;
  const R bR1 = Rr.bt1 * tbR, bR2 = Rr.bt2 * tbR;
  const R sbvL = bn * SM + bL1 * vL1 + bL2 * vL2;
  const R sbvR = bn * SM + bR1 * vR1 + bR2 * vR2;
  const R vbL = L.vn * L.bn + L.vt1 * L.bt1 + L.vt2 * L.bt2;
  const R vbR = Rr.vn * Rr.bn + Rr.vt1 * Rr.bt1 + Rr.vt2 * Rr.bt2;
  const R ezL = L.e * (suL * idL) + (ptz * SM - L.pt * L.vn + bn * (vbL - sbvL)) * idL;
  const R ezR = Rr.e * (suR * idR) + (ptz * SM - Rr.pt * Rr.vn + bn * (vbR - sbvR)) * idR;

  // double-star states (:1658-1698)
  const R abn = fabs(bn);
  const bool ibn = abn > R(1e-6);
  const R irL = frsq(rzL), irR = frsq(rzR);  // 1/sqrt(rho*)
  const R sqL = rzL * irL, sqR = rzR * irR;
  const R SZL = ibn ? SM - abn * irL : SM;
  const R SZR = ibn ? SM + abn * irR : SM;

  // region, sequential tests of :1722-1814 (later wins)
  int j = 0;
  if (SZL >= R(0)) j = 2;                   // SL <= 0 holds here
  if (ibn && SZL <= R(0) && SM >= R(0)) j = 3;
  if (ibn && SM <= R(0) && SZR >= R(0)) j = 4;
  if (SZR <= R(0)) j = 5;                   // SR >= 0 holds here
  const bool left = j <= 3;

  // star state of the chosen side and its HLL jump
  const R S = left ? SL : SR;
  const R rz = left ? rzL : rzR;
  const R vz1 = left ? vL1 : vR1, vz2 = left ? vL2 : vR2;
  const R bz1 = left ? bL1 : bR1, bz2 = left ? bL2 : bR2;
  const R ez = left ? ezL : ezR;
  Side<R> Q;  // the side whose fan holds the face
  Q.rho = left ? L.rho : Rr.rho;
  Q.vn = left ? L.vn : Rr.vn;
  Q.vt1 = left ? L.vt1 : Rr.vt1;
  Q.vt2 = left ? L.vt2 : Rr.vt2;
  Q.e = left ? L.e : Rr.e;
  Q.bn = left ? L.bn : Rr.bn;
  Q.bt1 = left ? L.bt1 : Rr.bt1;
  Q.bt2 = left ? L.bt2 : Rr.bt2;
  Q.pt = left ? L.pt : Rr.pt;
  phys_flux(Q, F);  // F_L or F_R
  // F* = F_Q + S (U* - U_Q)
  F[0] += S * (rz - Q.rho);
  F[1] += S * (rz * SM - Q.rho * Q.vn);
  F[2] += S * (rz * vz1 - Q.rho * Q.vt1);
  F[3] += S * (rz * vz2 - Q.rho * Q.vt2);
  F[4] += S * (ez - Q.e);
  F[6] += S * (bz1 - Q.bt1);
  F[7] += S * (bz2 - Q.bt2);

  if (j == 3 || j == 4) {  // U** - U*, weighted by SZL / SZR
    const R sbn = bn > R(0) ? R(1) : R(-1);
    const R iss = frcp(sqL + sqR);
    const R w1 = (sqL * vL1 + sqR * vR1 + sbn * (bR1 - bL1)) * iss;
    const R w2 = (sqL * vL2 + sqR * vR2 + sbn * (bR2 - bL2)) * iss;
    const R c1 = (sqL * bR1 + sqR * bL1 + sbn * sqL * sqR * (vR1 - vL1)) * iss;
    const R c2 = (sqL * bR2 + sqR * bL2 + sbn * sqL * sqR * (vR2 - vL2)) * iss;
    const R sbzz = bn * SM + c1 * w1 + c2 * w2;
    const R sq = left ? sqL : sqR;
    const R sbvz = left ? sbvL : sbvR;
    const R sgn = left ? R(-1) : R(1);
    const R SZ = left ? SZL : SZR;
    // e** = e* -/+ sqrt(rho*) sgn(bn) (b*.v* - b**.v**)
    const R dezz = sgn * sq * sbn * (sbvz - sbzz);
    F[2] += SZ * (rz * (w1 - vz1));
    F[3] += SZ * (rz * (w2 - vz2));
    F[4] += SZ * dezz;
    F[6] += SZ * (c1 - bz1);
    F[7] += SZ * (c2 - bz2);
  }
  F[5] = fbn;
  return j != 0;
}

}  // namespace setmhd
