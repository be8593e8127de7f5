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
//   * each repeated divisor is inverted once (frcp) and multiplied, and in
//     FP64 divisors known at the same point share one reciprocal: the pair
//     (mR - mL, SR - SL), the triple (S - SM, So - SM, tt) of the side whose
//     fan holds the face (S), the other side (So) and the single-star switch
//     denominator, and in the double-star regions the other side's tt with
//     sqrt(rhoL*) + sqrt(rhoR*) (frcp2, frcp3);
//   * 1/sqrt(rho*) is formed for the chosen side only, and for the other side
//     inside the double-star branch, the one place it is used;
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

// The default side accessor of hlld: pick(true) is the left state,
// pick(false) the right one, chosen per lane by selects.  Callers whose
// states sit in LDS pass an accessor that re-reads the chosen state instead
// (a select costs one VALU instruction per 32 bits).
template <class R>
struct PickSelect {
  const Side<R>& L;
  const Side<R>& Rr;
  __device__ __forceinline__ Side<R> operator()(bool left) const {
    Side<R> s;
    s.rho = left ? L.rho : Rr.rho;
    s.vn = left ? L.vn : Rr.vn;
    s.vt1 = left ? L.vt1 : Rr.vt1;
    s.vt2 = left ? L.vt2 : Rr.vt2;
    s.bn = left ? L.bn : Rr.bn;
    s.bt1 = left ? L.bt1 : Rr.bt1;
    s.bt2 = left ? L.bt2 : Rr.bt2;
    s.pt = left ? L.pt : Rr.pt;
    s.e = left ? L.e : Rr.e;
    s.cf = left ? L.cf : Rr.cf;
    return s;
  }
};

// Tangential velocity and field of one side's single-star state
// (:1624-1656): X is the side, m = rho (S - vn), su = S - vn, S its outer
// wave speed, it = 1/tt for tt = m (S - SM) - bn^2, or 0 where
// |tt| <= 1e-9, which leaves the side's tangential state (tv = tb = 0).
template <class R>
__device__ __forceinline__ void star_tangential(const Side<R>& X, R m, R su, R SM, R bn, R bn2, R it,
                                                R& v1, R& v2, R& b1, R& b2) {
  const R tv = (SM - X.vn) * it;
  const R tb = (m * su - bn2) * it;
  v1 = X.vt1 - bn * X.bt1 * tv;
  v2 = X.vt2 - bn * X.bt2 * tv;
  b1 = X.bt1 * tb;
  b2 = X.bt2 * tb;
}

// HLLD flux F (face frame) and max(|SL|,|SR|) for the CFL estimate.
// Returns false where no wave region matches (NaN input; the reference
// STOPs there, :1816-1825).
//
// The regions are checked in the reference's order (:1722-1814, later test
// wins); for any non-NaN state the winner lies left of the contact (regions
// 2, 3) exactly when SM > 0, so the side whose fan holds the face is known
// as soon as SM is: only that side's star state is formed, and the other
// side's tangential star state only in the double-star regions 3 and 4.
template <class R, class Pick>
__device__ __forceinline__ bool hlld(const Side<R>& L, const Side<R>& Rr, R F[8], R& smax,
                                     const Pick& pick) {
  const R SL = fmin(L.vn - L.cf, Rr.vn - Rr.cf);  // :1316-1319
  const R SR = fmax(L.vn + L.cf, Rr.vn + Rr.cf);
  const R aSL = fabs(SL), aSR = fabs(SR);
  smax = aSL > aSR ? aSL : aSR;
  const R fbn = -(smax * (Rr.bn - L.bn));  // :1828-1839 (w = 0)

  if (SL > R(0) || SR < R(0)) {  // regions 1 and 6: upwind
    phys_flux(pick(SL > R(0)), F);
    F[5] = fbn;
    return true;
  }
  const R suR = SR - Rr.vn, suL = SL - L.vn;
  const R mR = suR * Rr.rho, mL = suL * L.rho;
  R idn, isrl;  // 1/(mR - mL), 1/(SR - SL)
  frcp2(mR - mL, SR - SL, idn, isrl);
  const R SM = (mR * Rr.vn - Rr.pt + L.pt - mL * L.vn) * idn;            // :1343
  const bool left = SM > R(0);
  const Side<R> Q = pick(left);  // the side whose fan holds the face
  const R ptz = (mR * L.pt - mL * Rr.pt + mR * mL * (Rr.vn - L.vn)) * idn;  // :1600
  const R bn = (SR * Rr.bn - SL * L.bn) * isrl;                          // UZ(6)
  const R bn2 = bn * bn;

  // single-star state of the chosen side (:1624-1656).  su and m repeat
  // suL/suR and mL/mR's operations on the chosen side's operands instead of
  // selecting them.
  const R S = left ? SL : SR, So = left ? SR : SL;
  const R su = S - Q.vn;
  const R m = su * Q.rho;
  const R tt = m * (S - SM) - bn2;
  const bool tbig = fabs(tt) > R(1e-9);
  R id, ido, it;  // 1/(S - SM), 1/(So - SM), 1/tt
  frcp3(S - SM, So - SM, tbig ? tt : R(1), id, ido, it);
  it = tbig ? it : R(0);
  const R rz = m * id;  // rho* = r su/(S - SM)

  // double-star speed of the chosen side (:1658-1698).  The region tests of
  // :1722-1814 run in order with the later test winning.  Left of the
  // contact (SM > 0) the right Alfven speed SZR >= SM > 0 rules out regions
  // 4 and 5, so region 2 (SZL >= 0) or 3 (|bn| > 1e-6, SZL <= 0) wins and
  // only SZL is needed; otherwise (SM <= 0) SZL <= SM <= 0 makes a match of
  // regions 2 or 3 imply one of 4 or 5 (SM = 0), which win, and only SZR is
  // needed.  NaN states match no region.
  const R abn = fabs(bn);
  const bool ibn = abn > R(1e-6);
  const R irz = frsq(rz);  // 1/sqrt(rho*)
  const R sab = left ? -abn : abn;
  const R SZ = ibn ? SM + sab * irz : SM;  // SZL or SZR
  bool found, dstar;
  if (left) {
    found = SZ >= R(0) || (ibn && SZ <= R(0));
    dstar = ibn && SZ <= R(0);
  } else {
    const bool t4 = ibn && SM <= R(0) && SZ >= R(0);
    const bool t5 = SZ <= R(0);
    found = t4 || t5;
    dstar = t4 && !t5;
  }

  R vz1, vz2, bz1, bz2;
  star_tangential(Q, m, su, SM, bn, bn2, it, vz1, vz2, bz1, bz2);
  const R sbvz = bn * SM + bz1 * vz1 + bz2 * vz2;
  const R vbQ = Q.vn * Q.bn + Q.vt1 * Q.bt1 + Q.vt2 * Q.bt2;
  const R ez = Q.e * (su * id) + (ptz * SM - Q.pt * Q.vn + bn * (vbQ - sbvz)) * id;
  phys_flux(Q, F);  // F_L or F_R
  // F* = F_Q + S (U* - U_Q)
  F[0] += S * (rz - Q.rho);
  F[1] += S * (rz * SM - Q.rho * Q.vn);
  F[2] += S * (rz * vz1 - Q.rho * Q.vt1);
  F[3] += S * (rz * vz2 - Q.rho * Q.vt2);
  F[4] += S * (ez - Q.e);
  F[6] += S * (bz1 - Q.bt1);
  F[7] += S * (bz2 - Q.bt2);

  if (dstar) {  // U** - U*, weighted by SZL / SZR
    // kept a branch (an empty volatile asm is not speculated): waves whose
    // lanes all sit outside regions 3 and 4 skip it
    asm volatile("");
    // the other side O, its density and tangential single-star state
    const Side<R> O = pick(!left);
    const R suO = So - O.vn;
    const R mO = suO * O.rho;
    const R rzO = mO * ido;
    const R sq = rz * irz, sqo = rzO * frsq(rzO);  // sqrt(rho*)
    const R ttO = mO * (So - SM) - bn2;
    const bool tbigO = fabs(ttO) > R(1e-9);
    R itO, iss;  // 1/ttO, 1/(sqrt(rhoL*) + sqrt(rhoR*))
    frcp2(tbigO ? ttO : R(1), sq + sqo, itO, iss);
    itO = tbigO ? itO : R(0);
    R vo1, vo2, bo1, bo2;
    star_tangential(O, mO, suO, SM, bn, bn2, itO, vo1, vo2, bo1, bo2);
    // sbn (b_R - b_L) = so (b_O - b_Q)
    const R sbn = bn > R(0) ? R(1) : R(-1);
    const R so = left ? sbn : -sbn;
    const R w1 = (sq * vz1 + sqo * vo1 + so * (bo1 - bz1)) * iss;
    const R w2 = (sq * vz2 + sqo * vo2 + so * (bo2 - bz2)) * iss;
    const R c1 = (sq * bo1 + sqo * bz1 + so * sq * sqo * (vo1 - vz1)) * iss;
    const R c2 = (sq * bo2 + sqo * bz2 + so * sq * sqo * (vo2 - vz2)) * iss;
    const R sbzz = bn * SM + c1 * w1 + c2 * w2;
    // e** = e* -/+ sqrt(rho*) sgn(bn) (b*.v* - b**.v**)
    const R dezz = so * sq * (sbzz - sbvz);
    F[2] += SZ * (rz * (w1 - vz1));
    F[3] += SZ * (rz * (w2 - vz2));
    F[4] += SZ * dezz;
    F[6] += SZ * (c1 - bz1);
    F[7] += SZ * (c2 - bz2);
  }
  F[5] = fbn;
  return found;
}

template <class R>
__device__ __forceinline__ bool hlld(const Side<R>& L, const Side<R>& Rr, R F[8], R& smax) {
  return hlld(L, Rr, F, smax, PickSelect<R>{L, Rr});
}

}  // namespace setmhd
