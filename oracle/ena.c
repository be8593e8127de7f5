/*
 * ena.c -- TEST INFRASTRUCTURE: FP64 restatement of SET_ENA's line-of-sight
 * ENA spectrum (SET_ENA/SET_ENA.cpp:544-956, the loop of Culc_ENA and
 * Culc_ENA_the) on a triangulation of the cell centres.
 *
 * The reference interpolates the 16 cell parameters barycentrically in the
 * CGAL Delaunay triangle that contains each sample (SET_ENA.cpp:242-339).
 * CGAL is not present here, so the triangulation is an input; the oracle
 * finds the containing triangle by brute force (the triangle CGAL's exact
 * locate would return, up to points on an edge, where both neighbours give
 * the same value to rounding).  The march stops at the first sample outside
 * the triangulation.
 */
#include <math.h>
#include <stdlib.h>

#include "oracle.h"

#define ENA_SQRTPI 1.77245385     /* SET_ENA.cpp:58,86 (truncated)          */
#define ENA_A2 0.1307345665       /* :61                                    */
#define ENA_KN 0.4326569          /* :63                                    */
#define ENA_AE 0.00313166         /* 1 AU, :791                             */
#define ENA_KMS 0.0963358         /* 1 km/s, :792                           */

static double kv(double x) { return x * x; }

/* sigma(x), :64 */
static double ena_sigma(double x) { return kv(1.0 - ENA_A2 * log(x)); }

/* Velosity_1, :341-351: mean relative speed of a particle at speed u in a
 * Maxwellian of thermal speed cp */
double orc_ena_velocity1(double u, double cp) {
  if (u < 0.00001)
    return 2.0 * cp / ENA_SQRTPI + 2.0 * u * u / (3.0 * cp * ENA_SQRTPI) -
           u * u * u * u / (15.0 * cp * cp * cp * ENA_SQRTPI);
  return exp(-u * u / kv(cp)) * cp / ENA_SQRTPI + (u + kv(cp) / (2.0 * u)) * erf(u / cp);
}

/* maxwell_distribution, :353-370 */
double orc_ena_maxwell(double rho, double cp, double u, double v, double w, double vx, double vy,
                          double vz) {
  const double pi = 3.14159265358979323846;
  double dvx = vx - u, dvy = vy - v, dvz = vz - w;
  double v_sq = dvx * dvx + dvy * dvy + dvz * dvz;
  double norm = rho / (pi * sqrt(pi) * cp * cp * cp);
  return norm * exp(-v_sq / (cp * cp));
}

/* int_1_f1, int_1_f2, int_1_f3 (:426-536): the reference's piecewise
 * sixth-degree fits, coefficients as the reference states them. */
double orc_ena_f(int which, double x) {
  if (which == 1) {
    if (x <= 1.0)
      return 6.283185155644284 + 0.000024846677279866114 * x + 2.0934329078277405 * x * x +
             0.008055998193903208 * x * x * x - 0.2355169235647438 * x * x * x * x +
             0.03820480582423355 * x * x * x * x * x + 0.006992274370591744 * x * x * x * x * x * x;
    if (x <= 3.0)
      return 6.437524091973454 - 0.6331520099380095 * x + 3.1348881317268997 * x * x -
             0.8454201478027856 * x * x * x + 0.1004702004260311 * x * x * x * x +
             0.0009895488638964746 * x * x * x * x * x - 0.000920750276197054 * x * x * x * x * x * x;
    if (x <= 5)
      return 4.4920780630505135 + 2.5133093267020654 * x + 1.1327223176567935 * x * x -
             0.24648691152318875 * x * x * x + 0.031326738629523766 * x * x * x * x -
             0.0021366031960331384 * x * x * x * x * x + 0.00005954097505746697 * x * x * x * x * x * x;
    if (x <= 7)
      return 1.9138683588136232 + 5.350374732905213 * x - 0.16380205801427633 * x * x +
             0.06765898334856263 * x * x * x - 0.011071118267864083 * x * x * x * x +
             0.0008673476933852199 * x * x * x * x * x - 0.00002691859374483661 * x * x * x * x * x * x;
    return 1.3138472469154294 + 5.336877156136497 * x + 0.020286308991329983 * x * x -
           0.9780973533544137 * pow(x / 10.0, 3) + 0.26354051936651874 * pow(x / 10.0, 4) -
           0.03711733070841712 * pow(x / 10.0, 5) + 0.002120935433043921 * pow(x / 10.0, 6);
  }
  if (which == 2) {
    if (x <= 1.0)
      return 1.328216167939543 - 0.000004545681954848391 * x + 2.537368073155103 * x * x -
             0.0020584991728545624 * x * x * x - 0.03742568018912792 * x * x * x * x -
             0.010312136385277346 * x * x * x * x * x + 0.002767736179209713 * x * x * x * x * x * x;
    if (x <= 3.0)
      return 1.2959616295629246 + 0.1533684067037866 * x + 2.2354849981206106 * x * x +
             0.3113395567715921 * x * x * x - 0.21656309882941488 * x * x * x * x +
             0.041957500887605075 * x * x * x * x * x - 0.0029978773724628604 * x * x * x * x * x * x;
    if (x <= 5.0)
      return 1.903643456971281 - 1.4801836911099535 * x + 3.973958664572268 * x * x -
             0.6482729779428982 * x * x * x + 0.07665007314658864 * x * x * x * x -
             0.005369758193338703 * x * x * x * x * x + 0.00016605531871992049 * x * x * x * x * x * x;
    if (x <= 7.0)
      return -4.484415105552316 + 5.3747429756690766 * x + 0.8892806582308143 * x * x +
             0.09767316152573671 * x * x * x - 0.025704749778475783 * x * x * x * x +
             0.0021937998296249206 * x * x * x * x * x - 0.00006928845984076111 * x * x * x * x * x * x;
    return 0.0;
  }
  if (x <= 1.0)
    return 1.2938345594193854 - 0.000031719847351174835 * x + 1.3183710041280094 * x * x -
           0.014150512069488197 * x * x * x + 0.4226114681928129 * x * x * x * x -
           0.06985750969880078 * x * x * x * x * x - 0.015347864048406958 * x * x * x * x * x * x;
  if (x <= 3.0)
    return 0.9667460440956788 + 1.336271810704016 * x - 0.8687355257991665 * x * x +
           1.7676868273627229 * x * x * x - 0.2731222764016417 * x * x * x * x +
           0.004801770033831665 * x * x * x * x * x + 0.001780776080720323 * x * x * x * x * x * x;
  if (x <= 5.0)
    return 4.760566734123174 - 5.048204299463048 * x + 3.332342585228025 * x * x +
           0.47584339615235993 * x * x * x - 0.12072786272726124 * x * x * x * x +
           0.011870955604980658 * x * x * x * x * x - 0.0004580199652402304 * x * x * x * x * x * x;
  if (x <= 7.0)
    return 9.370493362469261 - 10.848615619383615 * x + 6.423326878282571 * x * x -
           0.4148977656870439 * x * x * x + 0.025300923044176957 * x * x * x * x -
           0.0010108688120876522 * x * x * x * x * x + 0.00001864423130429156 * x * x * x * x * x * x;
  return 0.0;
}

/* int_1, :538-542: Maxwellian average of |u| sigma(|u|) */
double orc_ena_int1(double x, double cp) {
  double b = 1.0 - ENA_A2 * log(cp);
  return (cp / (ENA_SQRTPI * ENA_SQRTPI * ENA_SQRTPI)) *
         (b * b * orc_ena_f(1, x / cp) - 2.0 * ENA_A2 * b * orc_ena_f(2, x / cp) +
          kv(ENA_A2) * orc_ena_f(3, x / cp));
}

/* charge-exchange rate of a particle at relative speed u in a population of
 * density rho and thermal speed cp, :879-888 */
static double ena_rate(double u, double rho, double cp) {
  if (u / cp > 7.0) {
    double uz = orc_ena_velocity1(u, cp);
    return (rho * uz * ena_sigma(uz)) / ENA_KN;
  }
  return (rho * orc_ena_int1(u, cp)) / ENA_KN;
}

/* CGAL::area of (p, q, r) */
static double area2d(double px, double py, double qx, double qy, double rx, double ry) {
  return ((qx - px) * (ry - py) - (rx - px) * (qy - py)) / 2.0;
}

/* Barycentric interpolation of the 16 parameters at (x, y), :242-339;
 * returns 0 outside every triangle. */
static int ena_interpolate(const orc_ena_mesh *m, double x, double y, double *res) {
  for (int t = 0; t < m->n_tri; t++) {
    const int i0 = m->tri[3 * t], i1 = m->tri[3 * t + 1], i2 = m->tri[3 * t + 2];
    const double *p0 = &m->cell_center[3 * i0], *p1 = &m->cell_center[3 * i1], *p2 = &m->cell_center[3 * i2];
    double lo = p0[0], hi = p0[0];
    if (p1[0] < lo) lo = p1[0];
    if (p2[0] < lo) lo = p2[0];
    if (p1[0] > hi) hi = p1[0];
    if (p2[0] > hi) hi = p2[0];
    if (x < lo - 1e-9 || x > hi + 1e-9) continue;
    double area = area2d(p0[0], p0[1], p1[0], p1[1], p2[0], p2[1]);
    double c0 = area2d(x, y, p1[0], p1[1], p2[0], p2[1]) / area;
    double c1 = area2d(p0[0], p0[1], x, y, p2[0], p2[1]) / area;
    double c2 = area2d(p0[0], p0[1], p1[0], p1[1], x, y) / area;
    const double eps = 1e-12;
    if (c0 < -eps || c1 < -eps || c2 < -eps) continue;
    for (int p = 0; p < 16; p++)
      res[p] = c0 * m->cell_par[16 * i0 + p] + c1 * m->cell_par[16 * i1 + p] + c2 * m->cell_par[16 * i2 + p];
    return 1;
  }
  return 0;
}

/* One line of sight at angle `the` from the x-axis (:782-931): jena[i] of
 * the n_v velocities v_i = vL + (i + 1/2)(vR - vL)/n_v, vL = 1 km/s,
 * vR = 1000 km/s; returns the number of samples inside the mesh. */
int orc_ena_los(const orc_ena_mesh *m, double the, int n_v, double *jena) {
  const double ae = ENA_AE, kms = ENA_KMS;
  double LOS_x = cos(the), LOS_y = sin(the);
  double a = sqrt(kv(0.0) + kv(LOS_x) + kv(LOS_y));
  LOS_x /= a;
  LOS_y /= a;
  double x0 = 0.0 + 5.0 * LOS_x * ae, y0 = 0.0 + 5.0 * LOS_y * ae;
  const double vL = 1 * kms, vR = 1000 * kms;
  const double ds = ae / 30.0;
  x0 = x0 + LOS_x * ds / 2.0;
  y0 = y0 + LOS_y * ds / 2.0;
  const double step = (vR - vL) / n_v;
  double *S = (double *)malloc(sizeof(double) * n_v);
  for (int i = 0; i < n_v; i++) {
    S[i] = 1.0;
    jena[i] = 0.0;
  }
  int samples = 0;
  double r16[16];
  while (ena_interpolate(m, x0, y0, r16)) {
    samples++;
    double rho_p = r16[0], p_p = r16[4], u_p = r16[1], v_p = r16[2];
    double cp_p = sqrt(p_p / rho_p);
    double rho_H3 = r16[8] * 3.0, p_H3 = 0.5 * rho_H3 * r16[11], u_H3 = r16[9], v_H3 = r16[10];
    double cp_H3 = sqrt(p_H3 / rho_H3);
    double rho_H4 = r16[12] * 3.0, p_H4 = 0.5 * rho_H4 * r16[15], u_H4 = r16[13], v_H4 = r16[14];
    double cp_H4 = sqrt(p_H4 / rho_H4);
    double r = sqrt(x0 * x0 + y0 * y0);
    double nu_ph = 770.291 * kv(ae / r);
    for (int i = 0; i < n_v; i++) {
      double v = vL + (i + 0.5) * step;
      double u = sqrt(kv(-v * LOS_x - u_p) + kv(-v * LOS_y - v_p) + kv(0.0));
      double nu_ex = ena_rate(u, rho_p, cp_p);
      u = sqrt(kv(-v * LOS_x - u_H3) + kv(-v * LOS_y - v_H3) + kv(0.0));
      double nu_H3 = ena_rate(u, rho_H3, cp_H3);
      u = sqrt(kv(-v * LOS_x - u_H4) + kv(-v * LOS_y - v_H4) + kv(0.0));
      double nu_H4 = ena_rate(u, rho_H4, cp_H4);
      double nu_H = nu_H3 + nu_H4;
      S[i] = S[i] * exp(-(nu_ex + nu_ph) * ds / v);
      jena[i] = jena[i] + nu_H * orc_ena_maxwell(rho_p, cp_p, u_p, v_p, 0.0, -v * LOS_x, -v * LOS_y, 0.0) * v * S[i] * ds;
    }
    x0 = x0 + LOS_x * ds;
    y0 = y0 + LOS_y * ds;
  }
  free(S);
  return samples;
}
