// chlld_dev.hpp -- device restatement of SET's point Riemann solver chlld
// (SET_Fortran_Cuda_MHD/Solvers.f90:1143-1859) and of signum/minmod/linear
// (Solvers.f90:108-150), for the unstructured drop-in kernels.
//
// Same evaluation order as the Fortran text; the including translation unit
// turns floating-point contraction off so every product and sum is rounded
// separately, and division and sqrt are the correctly rounded IEEE forms
// (quot, sqrt_rn of ieee_dev.hpp), a divisor used more than once being
// inverted once.  Returns 0, or -2 where the reference prints and STOPs
// (:1816-1825).
#pragma once
#include "ieee_dev.hpp"

namespace setmhd {

__device__ __forceinline__ int signum_ref(double x) {  // Solvers.f90:108-123
  if (x > 0) return 1;
  if (x < 0) return -1;
  return 0;
}

// Solvers.f90:125-137; ((s1+s2)/2.0) is integer/default-REAL: 0, +-0.5, +-1
__device__ __forceinline__ double minmod_ref(double x, double y) {
  int s = signum_ref(x) + signum_ref(y);
  if (s == 0) return 0.0;
  double h = (double)((float)s / 2.0f);
  double ax = fabs(x), ay = fabs(y);
  return h * (ax < ay ? ax : ay);
}

// Solvers.f90:140-150
__device__ __forceinline__ double linear_ref(double x1, double t1, double x2, double t2,
                                             double x3, double t3, double y) {
  double d = minmod_ref(quot(t1 - t2, x1 - x2), quot(t2 - t3, x2 - x3));
  return d * (y - x2) + t2;
}

// linear_ref with the two divisors inverted beforehand (the TVD calls it for
// five variables on the same three points): the same bits.
__device__ __forceinline__ double linear_inv(const Recip& i12, const Recip& i23, double x2, double t1, double t2,
                                             double t3, double y) {
  double d = minmod_ref(quot(t1 - t2, i12), quot(t2 - t3, i23));
  return d * (y - x2) + t2;
}

__device__ __forceinline__ double dmin1(double a, double b) { return a < b ? a : b; }
__device__ __forceinline__ double dmax1(double a, double b) { return a > b ? a : b; }
__device__ __forceinline__ double sq(double a) { return a * a; }

/* Solvers.f90:1143-1859.  Returns 0, or -2 where the
 * reference prints and STOPs (:1816-1825).  aco(i,j) is aco[i-1][j-1]. */
__device__ int chlld_ref(int n_state, double al, double be, double ge, double w,
              const double *qqq1, const double *qqq2, double *dsl, double *dsp,
              double *dsc, double *qqq, int n_disc) {
  double FR[8], FL[8], FW[8] = {0}, UL[8], UZ[8], UR[8];
  double UZL[8], UZR[8], UZZL[8], UZZR[8], dq[8];
  double vL[3], vR[3], bL[3], bR[3];
  double vzL[3], vzR[3], bzL[3], bzR[3];
  double vzzL[3], vzzR[3], bzzL[3], bzzR[3];
  double aco[3][3], qv[3], qb[3];
  const double x0 = 0., x1 = 1., x2 = 2., x3 = 3., x5 = 5.;
  int n_disco = n_disc; /* :1182-1186, present(n_disc) */
  int ik, i;

  double spi4 = sqrt(kCpi4); /* :1189-1191 */
  const Recip i_spi4 = recip_of(spi4);
  double eps = 1.e-12;          /* :1193 */
  double epsb = 1.e-06;         /* :1194 */
  double ga = x5 / x3;          /* :1199 */
  double g1 = ga - x1;
  double wv = w;

  double r1 = qqq1[0], u1 = qqq1[1], v1 = qqq1[2], w1 = qqq1[3], p1 = qqq1[4];
  double bx1 = quot(qqq1[5], i_spi4), by1 = quot(qqq1[6], i_spi4), bz1 = quot(qqq1[7], i_spi4);
  double r2 = qqq2[0], u2 = qqq2[1], v2 = qqq2[2], w2 = qqq2[3], p2 = qqq2[4];
  double bx2 = quot(qqq2[5], i_spi4), by2 = quot(qqq2[6], i_spi4), bz2 = quot(qqq2[7], i_spi4);

  double ro = (r2 + r1) / x2; /* :1225-1232 */
  double ap = (p2 + p1) / x2;
  double abx = (bx2 + bx1) / x2;
  double aby = (by2 + by1) / x2;
  double abz = (bz2 + bz1) / x2;

  double bk = abx * al + aby * be + abz * ge; /* :1235-1240 */
  double b2 = sq(abx) + sq(aby) + sq(abz);
  double d = b2 - sq(bk);
  aco[0][0] = al;
  aco[1][0] = be;
  aco[2][0] = ge;
  if (d > eps) { /* :1243-1251 */
    d = sqrt_rn(d);
    const Recip i_d = recip_of(d);
    aco[0][1] = quot(abx - bk * al, i_d);
    aco[1][1] = quot(aby - bk * be, i_d);
    aco[2][1] = quot(abz - bk * ge, i_d);
    aco[0][2] = quot(aby * ge - abz * be, i_d);
    aco[1][2] = quot(abz * al - abx * ge, i_d);
    aco[2][2] = quot(abx * be - aby * al, i_d);
  } else { /* :1252-1273 */
    double aix, aiy, aiz;
    if (fabs(al) < fabs(be) && fabs(al) < fabs(ge)) {
      aix = x1; aiy = x0; aiz = x0;
    } else if (fabs(be) < fabs(ge)) {
      aix = x0; aiy = x1; aiz = x0;
    } else {
      aix = x0; aiy = x0; aiz = x1;
    }
    double aik = aix * al + aiy * be + aiz * ge;
    d = sqrt_rn(x1 - sq(aik));
    const Recip i_d = recip_of(d);
    aco[0][1] = quot(aix - aik * al, i_d);
    aco[1][1] = quot(aiy - aik * be, i_d);
    aco[2][1] = quot(aiz - aik * ge, i_d);
    aco[0][2] = quot(aiy * ge - aiz * be, i_d);
    aco[1][2] = quot(aiz * al - aix * ge, i_d);
    aco[2][2] = quot(aix * be - aiy * al, i_d);
  }

  for (i = 0; i < 3; i++) { /* :1275-1280 */
    vL[i] = aco[0][i] * u1 + aco[1][i] * v1 + aco[2][i] * w1;
    vR[i] = aco[0][i] * u2 + aco[1][i] * v2 + aco[2][i] * w2;
    bL[i] = aco[0][i] * bx1 + aco[1][i] * by1 + aco[2][i] * bz1;
    bR[i] = aco[0][i] * bx2 + aco[1][i] * by2 + aco[2][i] * bz2;
  }

  const Recip i_r1 = recip_of(r1), i_r2 = recip_of(r2);
  double aaL = quot(bL[0], sqrt_rn(r1)); /* :1282-1289 */
  double b2L = sq(bL[0]) + sq(bL[1]) + sq(bL[2]);
  double b21 = quot(b2L, i_r1);
  double cL = sqrt_rn(quot(ga * p1, i_r1));
  double qp = sqrt_rn(b21 + cL * (cL + x2 * aaL));
  double qm = sqrt_rn(b21 + cL * (cL - x2 * aaL));
  double cfL = (qp + qm) / x2;
  double ptL = p1 + b2L / x2;

  double aaR = quot(bR[0], sqrt_rn(r2)); /* :1291-1298 */
  double b2R = sq(bR[0]) + sq(bR[1]) + sq(bR[2]);
  double b22 = quot(b2R, i_r2);
  double cR = sqrt_rn(quot(ga * p2, i_r2));
  qp = sqrt_rn(b22 + cR * (cR + x2 * aaR));
  qm = sqrt_rn(b22 + cR * (cR - x2 * aaR));
  double cfR = (qp + qm) / x2;
  double ptR = p2 + b2R / x2;

  double aC = (aaL + aaR) / x2; /* :1300-1306 */
  double b2o = (b22 + b21) / x2;
  double cC = sqrt_rn(quot(ga * ap, ro));
  qp = sqrt_rn(b2o + cC * (cC + x2 * aC));
  qm = sqrt_rn(b2o + cC * (cC - x2 * aC));
  double cfC = (qp + qm) / x2;
  double vC1 = (vL[0] + vR[0]) / x2;

  double SL = 0, SR = 0; /* :1308-1331 */
  if (n_disco == 1) {
    SL = dmin1((vL[0] - cfL), (vC1 - cfC));
    SR = dmax1((vR[0] + cfR), (vC1 + cfC));
  }
  if (n_disco == 0) {
    SL = dmin1((vL[0] - cfL), (vR[0] - cfR));
    SR = dmax1((vL[0] + cfL), (vR[0] + cfR));
  }
  if (n_disco == 2) {
    double SL_1 = dmin1((vL[0] - cfL), (vC1 - cfC));
    double SR_1 = dmax1((vR[0] + cfR), (vC1 + cfC));
    double SL_2 = dmin1((vL[0] - cfL), (vR[0] - cfR));
    double SR_2 = dmax1((vL[0] + cfL), (vR[0] + cfR));
    double oo = 0.75, oo1 = 1.0 - oo;
    SL = oo * SL_1 + oo1 * SL_2;
    SR = oo * SR_1 + oo1 * SR_2;
  }

  double suR = SR - vR[0]; /* :1341-1344 */
  double suL = SL - vL[0];
  const Recip i_dm = recip_of(suR * r2 - suL * r1);  // SM's divisor, and ptz's (HLLD)
  double SM = quot(suR * r2 * vR[0] - ptR + ptL - suL * r1 * vL[0], i_dm);

  *dsl = SL; /* :1346-1348: captured before the LAX override */
  *dsc = SM;
  *dsp = SR;

  if (n_state == 0) { /* :1360-1365 */
    double TR0 = fabs(vL[0] + vR[0]) / x2 + cfC;
    double TL0 = -TR0;
    SR = TR0;
    SL = TL0;
  }

  double upt1 = (sq(u1) + sq(v1) + sq(w1)) / x2; /* :1368-1375 */
  double sbv1 = u1 * bx1 + v1 * by1 + w1 * bz1;
  double upt2 = (sq(u2) + sq(v2) + sq(w2)) / x2;
  double sbv2 = u2 * bx2 + v2 * by2 + w2 * bz2;
  const Recip i_g1 = recip_of(g1);
  double e1 = quot(p1, i_g1) + r1 * upt1 + b2L / x2;
  double e2 = quot(p2, i_g1) + r2 * upt2 + b2R / x2;

  FL[0] = r1 * vL[0]; /* :1377-1393 */
  FL[1] = r1 * vL[0] * vL[0] + ptL - sq(bL[0]);
  FL[2] = r1 * vL[0] * vL[1] - bL[0] * bL[1];
  FL[3] = r1 * vL[0] * vL[2] - bL[0] * bL[2];
  FL[4] = (e1 + ptL) * vL[0] - bL[0] * sbv1;
  FL[5] = x0;
  FL[6] = vL[0] * bL[1] - vL[1] * bL[0];
  FL[7] = vL[0] * bL[2] - vL[2] * bL[0];

  FR[0] = r2 * vR[0];
  FR[1] = r2 * vR[0] * vR[0] + ptR - sq(bR[0]);
  FR[2] = r2 * vR[0] * vR[1] - bR[0] * bR[1];
  FR[3] = r2 * vR[0] * vR[2] - bR[0] * bR[2];
  FR[4] = (e2 + ptR) * vR[0] - bR[0] * sbv2;
  FR[5] = x0;
  FR[6] = vR[0] * bR[1] - vR[1] * bR[0];
  FR[7] = vR[0] * bR[2] - vR[2] * bR[0];

  UL[0] = r1; /* :1395-1405 */
  UL[4] = e1;
  UR[0] = r2;
  UR[4] = e2;
  for (ik = 0; ik < 3; ik++) {
    UL[ik + 1] = r1 * vL[ik];
    UL[ik + 5] = bL[ik];
    UR[ik + 1] = r2 * vR[ik];
    UR[ik + 5] = bR[ik];
  }

  const Recip i_srl = recip_of(SR - SL);
  for (ik = 0; ik < 8; ik++) /* :1407-1409 */
    UZ[ik] = quot(SR * UR[ik] - SL * UL[ik] + FL[ik] - FR[ik], i_srl);

  if (n_state <= 1) { /* LAX / HLL, :1415-1463 */
    for (ik = 0; ik < 8; ik++) dq[ik] = UR[ik] - UL[ik];
    double TL = SL, TR = SR;
    if (SL > wv) {
      TL = x0;
      for (ik = 0; ik < 8; ik++) FW[ik] = wv * UL[ik];
    }
    if (SL <= wv && wv <= SR)
      for (ik = 0; ik < 8; ik++) FW[ik] = wv * UZ[ik];
    if (SR < wv) {
      TR = x0;
      for (ik = 0; ik < 8; ik++) FW[ik] = wv * UR[ik];
    }
    double a = TR * TL;
    const Recip i_b = recip_of(TR - TL);
    qqq[0] = quot(TR * FL[0] - TL * FR[0] + a * dq[0], i_b) - FW[0];
    qqq[4] = quot(TR * FL[4] - TL * FR[4] + a * dq[4], i_b) - FW[4];
    for (ik = 1; ik < 4; ik++)
      qv[ik - 1] = quot(TR * FL[ik] - TL * FR[ik] + a * dq[ik], i_b) - FW[ik];
    for (ik = 5; ik < 8; ik++)
      qb[ik - 5] = quot(TR * FL[ik] - TL * FR[ik] + a * dq[ik], i_b) - FW[ik];
    for (i = 0; i < 3; i++) {
      qqq[i + 1] = aco[i][0] * qv[0] + aco[i][1] * qv[1] + aco[i][2] * qv[2];
      qqq[i + 5] = aco[i][0] * qb[0] + aco[i][1] * qb[1] + aco[i][2] * qb[2];
      qqq[i + 5] = spi4 * qqq[i + 5];
    }
    return 0;
  }

  if (n_state == 2) { /* HLLC, :1465-1596 */
    const Recip i_srm = recip_of(SR - SM), i_slm = recip_of(SL - SM);
    double suRm = quot(suR, i_srm);
    double suLm = quot(suL, i_slm);
    double rzR = r2 * suRm;
    double rzL = r1 * suLm;
    vzR[0] = SM;
    vzL[0] = SM;
    double ptzR = ptR + r2 * suR * (SM - vR[0]);
    double ptzL = ptL + r1 * suL * (SM - vL[0]);
    double ptz = (ptzR + ptzL) / x2;
    bzR[0] = UZ[5];
    bzL[0] = UZ[5];

    const Recip i_uz0 = recip_of(UZ[0]);
    vzR[1] = quot(UZ[2], i_uz0);
    vzR[2] = quot(UZ[3], i_uz0);
    vzL[1] = vzR[1];
    vzL[2] = vzR[2];

    const Recip i_sur = recip_of(suR), i_sul = recip_of(suL);
    vzR[1] = vR[1] + quot(quot(UZ[5] * (bR[1] - UZ[6]), i_sur), i_r2);
    vzR[2] = vR[2] + quot(quot(UZ[5] * (bR[2] - UZ[7]), i_sur), i_r2);
    vzL[1] = vL[1] + quot(quot(UZ[5] * (bL[1] - UZ[6]), i_sul), i_r1);
    vzL[2] = vL[2] + quot(quot(UZ[5] * (bL[2] - UZ[7]), i_sul), i_r1);

    bzR[1] = UZ[6];
    bzR[2] = UZ[7];
    bzL[1] = bzR[1];
    bzL[2] = bzR[2];

    double sbvz = quot(UZ[5] * UZ[1] + UZ[6] * UZ[2] + UZ[7] * UZ[3], i_uz0);

    double ezR = e2 * suRm + quot(ptz * SM - ptR * vR[0] + UZ[5] * (sbv2 - sbvz), i_srm);
    double ezL = e1 * suLm + quot(ptz * SM - ptL * vL[0] + UZ[5] * (sbv1 - sbvz), i_slm);

    if (fabs(UZ[5]) < epsb) { /* :1499-1508 */
      vzR[1] = vR[1];
      vzR[2] = vR[2];
      vzL[1] = vL[1];
      vzL[2] = vL[2];
      bzR[1] = bR[1] * suRm;
      bzR[2] = bR[2] * suRm;
      bzL[1] = bL[1] * suLm;
      bzL[2] = bL[2] * suLm;
    }

    UZL[0] = rzL;
    UZL[4] = ezL;
    UZR[0] = rzR;
    UZR[4] = ezR;
    for (ik = 0; ik < 3; ik++) {
      UZL[ik + 1] = vzL[ik] * rzL;
      UZL[ik + 5] = bzL[ik];
      UZR[ik + 1] = vzR[ik] * rzR;
      UZR[ik + 5] = bzR[ik];
    }

    if (SL > wv) { /* :1521-1531 */
      qqq[0] = FL[0] - wv * UL[0];
      qqq[4] = FL[4] - wv * UL[4];
      for (ik = 1; ik < 4; ik++) qv[ik - 1] = FL[ik] - wv * UL[ik];
      for (ik = 5; ik < 8; ik++) qb[ik - 5] = FL[ik] - wv * UL[ik];
    }
    if (SL <= wv && SM >= wv) { /* :1533-1543 */
      qqq[0] = FL[0] + SL * (rzL - r1) - wv * UZL[0];
      qqq[4] = FL[4] + SL * (ezL - e1) - wv * UZL[4];
      for (ik = 1; ik < 4; ik++)
        qv[ik - 1] = FL[ik] + SL * (UZL[ik] - UL[ik]) - wv * UZL[ik];
      for (ik = 5; ik < 8; ik++)
        qb[ik - 5] = FL[ik] + SL * (UZL[ik] - UL[ik]) - wv * UZL[ik];
    }
    if (SM <= wv && SR >= wv) { /* :1545-1556 */
      qqq[0] = FR[0] + SR * (rzR - r2) - wv * UZR[0];
      qqq[4] = FR[4] + SR * (ezR - e2) - wv * UZR[4];
      for (ik = 1; ik < 4; ik++)
        qv[ik - 1] = FR[ik] + SR * (UZR[ik] - UR[ik]) - wv * UZR[ik];
      for (ik = 5; ik < 8; ik++)
        qb[ik - 5] = FR[ik] + SR * (UZR[ik] - UR[ik]) - wv * UZR[ik];
    }
    if (SR < wv) { /* :1558-1567 */
      qqq[0] = FR[0] - wv * UR[0];
      qqq[4] = FR[4] - wv * UR[4];
      for (ik = 1; ik < 4; ik++) qv[ik - 1] = FR[ik] - wv * UR[ik];
      for (ik = 5; ik < 8; ik++) qb[ik - 5] = FR[ik] - wv * UR[ik];
    }

    double SN = dmax1(fabs(SL), fabs(SR)); /* :1569-1581 */
    double wbn;
    if (wv >= SR)
      wbn = wv * bR[0];
    else if (wv <= SL)
      wbn = wv * bL[0];
    else
      wbn = wv * (bL[0] + bR[0]) / x2;
    qb[0] = -(SN * (bR[0] - bL[0])) - wbn;

    for (i = 0; i < 3; i++) { /* :1585-1589 */
      qqq[i + 1] = aco[i][0] * qv[0] + aco[i][1] * qv[1] + aco[i][2] * qv[2];
      qqq[i + 5] = aco[i][0] * qb[0] + aco[i][1] * qb[1] + aco[i][2] * qb[2];
      qqq[i + 5] = spi4 * qqq[i + 5];
    }
    return 0;
  }

  if (n_state == 3) { /* HLLD, :1598-1855 */
    double ptz = quot(suR * r2 * ptL - suL * r1 * ptR + r1 * r2 * suR * suL * (vR[0] - vL[0]), i_dm);
    vzL[0] = SM;
    vzR[0] = SM;
    vzzL[0] = SM;
    vzzR[0] = SM;

    const Recip i_srm = recip_of(SR - SM), i_slm = recip_of(SL - SM);
    double suRm = quot(suR, i_srm);
    double suLm = quot(suL, i_slm);
    double rzR = r2 * suRm;
    double rzL = r1 * suLm;

    double bn = UZ[5];
    double bn2 = bn * bn;
    bzL[0] = bn;
    bzR[0] = bn;
    bzzL[0] = bn;
    bzzR[0] = bn;

    double ttR = r2 * suR * (SR - SM) - bn2, tvR, tbR; /* :1624-1640 */
    if (fabs(ttR) <= 1.e-09) {
      tvR = x0;
      tbR = x0;
    } else {
      const Recip i_tt = recip_of(ttR);
      tvR = quot(SM - vR[0], i_tt);
      tbR = quot(r2 * suR * suR - bn2, i_tt);
    }
    double ttL = r1 * suL * (SL - SM) - bn2, tvL, tbL;
    if (fabs(ttL) <= 1.e-09) {
      tvL = x0;
      tbL = x0;
    } else {
      const Recip i_tt = recip_of(ttL);
      tvL = quot(SM - vL[0], i_tt);
      tbL = quot(r1 * suL * suL - bn2, i_tt);
    }

    vzL[1] = vL[1] - bn * bL[1] * tvL; /* :1642-1650 */
    vzL[2] = vL[2] - bn * bL[2] * tvL;
    vzR[1] = vR[1] - bn * bR[1] * tvR;
    vzR[2] = vR[2] - bn * bR[2] * tvR;
    bzL[1] = bL[1] * tbL;
    bzL[2] = bL[2] * tbL;
    bzR[1] = bR[1] * tbR;
    bzR[2] = bR[2] * tbR;

    double sbvL = bzL[0] * vzL[0] + bzL[1] * vzL[1] + bzL[2] * vzL[2];
    double sbvR = bzR[0] * vzR[0] + bzR[1] * vzR[1] + bzR[2] * vzR[2];

    double ezR = e2 * suRm + quot(ptz * SM - ptR * vR[0] + bn * (sbv2 - sbvR), i_srm);
    double ezL = e1 * suLm + quot(ptz * SM - ptL * vL[0] + bn * (sbv1 - sbvL), i_slm);

    double rzzR = rzR; /* :1658-1666 */
    double rzzL = rzL;
    double rzRs = sqrt_rn(rzR);
    double rzLs = sqrt_rn(rzL);
    double rzss = rzRs + rzLs;
    double rzps = rzRs * rzLs;

    double SZL = SM - quot(fabs(bn), rzLs);
    double SZR = SM + quot(fabs(bn), rzRs);

    int ibn; /* :1668-1677 */
    double sbn;
    if (fabs(bn) > epsb) {
      sbn = quot(fabs(bn), bn);
      ibn = 1;
    } else {
      sbn = x0;
      ibn = 0;
      SZL = SM;
      SZR = SM;
    }

    const Recip i_rzss = recip_of(rzss);
    vzzL[1] = quot(rzLs * vzL[1] + rzRs * vzR[1] + sbn * (bzR[1] - bzL[1]), i_rzss);
    vzzL[2] = quot(rzLs * vzL[2] + rzRs * vzR[2] + sbn * (bzR[2] - bzL[2]), i_rzss);
    vzzR[1] = vzzL[1];
    vzzR[2] = vzzL[2];

    bzzL[1] = quot(rzLs * bzR[1] + rzRs * bzL[1] + sbn * rzps * (vzR[1] - vzL[1]), i_rzss);
    bzzL[2] = quot(rzLs * bzR[2] + rzRs * bzL[2] + sbn * rzps * (vzR[2] - vzL[2]), i_rzss);
    bzzR[1] = bzzL[1];
    bzzR[2] = bzzL[2];

    double sbzz = bzzL[0] * vzzL[0] + bzzL[1] * vzzL[1] + bzzL[2] * vzzL[2];

    double ezzR = ezR + rzRs * sbn * (sbvR - sbzz);
    double ezzL = ezL - rzLs * sbn * (sbvL - sbzz);

    UZL[0] = rzL; /* :1700-1720 */
    UZL[4] = ezL;
    UZR[0] = rzR;
    UZR[4] = ezR;
    for (ik = 0; ik < 3; ik++) {
      UZL[ik + 1] = vzL[ik] * rzL;
      UZL[ik + 5] = bzL[ik];
      UZR[ik + 1] = vzR[ik] * rzR;
      UZR[ik + 5] = bzR[ik];
    }
    UZZL[0] = rzzL;
    UZZL[4] = ezzL;
    UZZR[0] = rzzR;
    UZZR[4] = ezzR;
    for (ik = 0; ik < 3; ik++) {
      UZZL[ik + 1] = vzzL[ik] * rzzL;
      UZZL[ik + 5] = bzzL[ik];
      UZZR[ik + 1] = vzzR[ik] * rzzR;
      UZZR[ik + 5] = bzzR[ik];
    }

    int j_ccs = -1;
    if (SL > wv) { /* :1722-1732 */
      qqq[0] = FL[0] - wv * UL[0];
      qqq[4] = FL[4] - wv * UL[4];
      for (ik = 1; ik < 4; ik++) qv[ik - 1] = FL[ik] - wv * UL[ik];
      for (ik = 5; ik < 8; ik++) qb[ik - 5] = FL[ik] - wv * UL[ik];
      j_ccs = 1;
    }
    if (SL <= wv && SZL >= wv) { /* :1734-1746 */
      ik = 0;
      qqq[ik] = FL[ik] + SL * (UZL[ik] - UL[ik]) - wv * UZL[ik];
      ik = 4;
      qqq[ik] = FL[ik] + SL * (UZL[ik] - UL[ik]) - wv * UZL[ik];
      for (ik = 1; ik < 4; ik++)
        qv[ik - 1] = FL[ik] + SL * (UZL[ik] - UL[ik]) - wv * UZL[ik];
      for (ik = 5; ik < 8; ik++)
        qb[ik - 5] = FL[ik] + SL * (UZL[ik] - UL[ik]) - wv * UZL[ik];
      j_ccs = 2;
    }
    if (ibn == 1) { /* :1748-1787 */
      if (SZL <= wv && SM >= wv) {
        ik = 0;
        qqq[ik] = FL[ik] + SZL * (UZZL[ik] - UZL[ik]) + SL * (UZL[ik] - UL[ik]) - wv * UZZL[ik];
        ik = 4;
        qqq[ik] = FL[ik] + SZL * (UZZL[ik] - UZL[ik]) + SL * (UZL[ik] - UL[ik]) - wv * UZZL[ik];
        for (ik = 1; ik < 4; ik++)
          qv[ik - 1] = FL[ik] + SZL * (UZZL[ik] - UZL[ik]) + SL * (UZL[ik] - UL[ik]) - wv * UZZL[ik];
        for (ik = 5; ik < 8; ik++)
          qb[ik - 5] = FL[ik] + SZL * (UZZL[ik] - UZL[ik]) + SL * (UZL[ik] - UL[ik]) - wv * UZZL[ik];
        j_ccs = 3;
      }
      if (SM <= wv && SZR >= wv) {
        ik = 0;
        qqq[ik] = FR[ik] + SZR * (UZZR[ik] - UZR[ik]) + SR * (UZR[ik] - UR[ik]) - wv * UZZR[ik];
        ik = 4;
        qqq[ik] = FR[ik] + SZR * (UZZR[ik] - UZR[ik]) + SR * (UZR[ik] - UR[ik]) - wv * UZZR[ik];
        for (ik = 1; ik < 4; ik++)
          qv[ik - 1] = FR[ik] + SZR * (UZZR[ik] - UZR[ik]) + SR * (UZR[ik] - UR[ik]) - wv * UZZR[ik];
        for (ik = 5; ik < 8; ik++)
          qb[ik - 5] = FR[ik] + SZR * (UZZR[ik] - UZR[ik]) + SR * (UZR[ik] - UR[ik]) - wv * UZZR[ik];
        j_ccs = 4;
      }
    }
    if (SZR <= wv && SR >= wv) { /* :1789-1802 */
      ik = 0;
      qqq[ik] = FR[ik] + SR * (UZR[ik] - UR[ik]) - wv * UZR[ik];
      ik = 4;
      qqq[ik] = FR[ik] + SR * (UZR[ik] - UR[ik]) - wv * UZR[ik];
      for (ik = 1; ik < 4; ik++)
        qv[ik - 1] = FR[ik] + SR * (UZR[ik] - UR[ik]) - wv * UZR[ik];
      for (ik = 5; ik < 8; ik++)
        qb[ik - 5] = FR[ik] + SR * (UZR[ik] - UR[ik]) - wv * UZR[ik];
      j_ccs = 5;
    }
    if (SR < wv) { /* :1804-1814 */
      qqq[0] = FR[0] - wv * UR[0];
      qqq[4] = FR[4] - wv * UR[4];
      for (ik = 1; ik < 4; ik++) qv[ik - 1] = FR[ik] - wv * UR[ik];
      for (ik = 5; ik < 8; ik++) qb[ik - 5] = FR[ik] - wv * UR[ik];
      j_ccs = 6;
    }
    if (j_ccs == -1) { /* :1816-1825 print + STOP */
      for (ik = 0; ik < 8; ik++) qqq[ik] = __builtin_nan("");
      return -2;
    }

    double SN = dmax1(fabs(SL), fabs(SR)); /* :1828-1839 */
    double wbn;
    if (wv >= SR)
      wbn = wv * bR[0];
    else if (wv <= SL)
      wbn = wv * bL[0];
    else
      wbn = wv * (bL[0] + bR[0]) / x2;
    qb[0] = -(SN * (bR[0] - bL[0])) - wbn;

    for (i = 0; i < 3; i++) { /* :1843-1847 */
      qqq[i + 1] = aco[i][0] * qv[0] + aco[i][1] * qv[1] + aco[i][2] * qv[2];
      qqq[i + 5] = aco[i][0] * qb[0] + aco[i][1] * qb[1] + aco[i][2] * qb[2];
      qqq[i + 5] = spi4 * qqq[i + 5];
    }
    return 0;
  }
  return 0; /* :1857, other n_state: qqq untouched */
}

}  // namespace setmhd
