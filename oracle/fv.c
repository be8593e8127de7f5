/*
 * fv.c -- TEST INFRASTRUCTURE: FP64 restatement of SET's face pass, cell
 * pass, charge-exchange source and step loop, on the reference's own
 * face-list layout.  See oracle.h for the file:line map.
 *
 * Semantic switches (orc_config) make the reference's heliosphere-specific
 * behaviour optional so the same code also states the ideal-MHD Cartesian
 * runs: with ORC_F_ZERO_BW cleared, w and B are advanced with the update
 * formulas that the reference keeps as comments (cuf_solvers.cuf:725-733,
 * sks = 0 because Gran_POTOK(9) = 0 at :433) and the total energy carries
 * the magnetic part |B|^2/(8 pi) = |b|^2/2 that chlld's energy flux uses.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "oracle.h"

#define NPAR 16

static inline double sq(double a) { return a * a; }

/* dev_norm2, cuf_kernel.cuf:196-203 */
static inline double norm2v(const double *x) {
  return sqrt(sq(x[0]) + sq(x[1]) + sq(x[2]));
}

/* Source_MK, cuf_kernel.cuf:215-288 */
void orc_source_mk(const double *cp, double *out) {
  const double pi = acos(-1.0);
  double rho = cp[0], u = cp[1], v = cp[2], p = cp[4];
  double rhoH3 = cp[8], uH3 = cp[9], vH3 = cp[10];
  double pH3 = 0.5 * cp[11] * rhoH3;
  if (rhoH3 < F4(0.0000001)) {
    rhoH3 = F4(0.0000001);
    pH3 = 0.0;
  }
  double rhoH4 = cp[12], uH4 = cp[13], vH4 = cp[14];
  double pH4 = 0.5 * cp[15] * rhoH4;
  if (rhoH4 < F4(0.0000001)) {
    rhoH4 = F4(0.0000001);
    pH4 = 0.0;
  }
  const double c64 = 64.0 / (9.0 * pi), c4 = 4.0 / (pi);
  const double ksig = F4(0.1307345665);
  const double kk = (double)(3.0f / 0.4326569f); /* REAL(4) constant expression */

  double U_M_H3 = sqrt(sq(u - uH3) + sq(v - vH3) + c64 * (p / rho + 2.0 * pH3 / rhoH3));
  double U_H3 = sqrt(sq(u - uH3) + sq(v - vH3) + c4 * (p / rho + 2.0 * pH3 / rhoH3));
  double sigma_H3 = sq(1.0 - ksig * log(U_M_H3));
  double nu_H3 = rho * rhoH3 * U_M_H3 * sigma_H3;

  double U_M_H4 = sqrt(sq(u - uH4) + sq(v - vH4) + c64 * (p / rho + 2.0 * pH4 / rhoH4));
  double U_H4 = sqrt(sq(u - uH4) + sq(v - vH4) + c4 * (p / rho + 2.0 * pH4 / rhoH4));
  double sigma_H4 = sq(1.0 - ksig * log(U_M_H4));
  double nu_H4 = rho * rhoH4 * U_M_H4 * sigma_H4;

  double sum1 = 0.0, sum2 = 0.0, sum3 = 0.0;
  sum1 = sum1 + kk * nu_H3 * (uH3 - u);
  sum2 = sum2 + kk * nu_H3 * (vH3 - v);
  sum3 = sum3 + kk * nu_H3 * ((sq(uH3) + sq(vH3) - sq(u) - sq(v)) / 2.0 +
                              (U_H3 / U_M_H3) * (2.0 * pH3 / rhoH3 - p / rho));
  sum1 = sum1 + kk * nu_H4 * (uH4 - u);
  sum2 = sum2 + kk * nu_H4 * (vH4 - v);
  sum3 = sum3 + kk * nu_H4 * ((sq(uH4) + sq(vH4) - sq(u) - sq(v)) / 2.0 +
                              (U_H4 / U_M_H4) * (2.0 * pH4 / rhoH4 - p / rho));
  out[0] = sum1;
  out[1] = sum2;
  out[2] = sum3;
}

/* The two states of face gr that CUF_MGD_grans_RK2 (cuf_solvers.cuf:
 * 300-420) hands to chlld: the cells' primitives, the boundary ghost of
 * types 4/5/wall/1/3 (:318-354) and the minmod TVD reconstruction with its
 * positivity fallback (:363-412).  Returns 1 if the face is skipped (the
 * inner-sphere test, :282-284), 0, or ORC_ERR_FACE_RHO (:414-420). */
int orc_face_states(const orc_config *cfg, const orc_mesh *m, int gr, int stage,
                    const double *W, const double *W2, double *q1, double *q2) {
  const double *S = (stage == 2) ? W2 : W;
  int type = m->gran_type[gr];
  int s1 = m->gran_neighbour[2 * gr + 0];
  int s2 = m->gran_neighbour[2 * gr + 1];
  const double *c1 = &m->cell_center[3 * (s1 - 1)];
  if (norm2v(c1) < cfg->r_skip_face) return 1; /* :282-284 */
  const double *gc = &m->gran_center[3 * gr];
  const double *nrm = &m->gran_normal[3 * gr];
  memcpy(q1, &S[NPAR * (s1 - 1)], 8 * sizeof(double)); /* :305-309 */
  const double *c2 = c1;
  int tvd = 1;
  if (type == 4) { /* :318-324 */
    memcpy(q2, q1, 8 * sizeof(double));
    if ((cfg->flags & ORC_F_HELIO_HACKS) && q2[1] > F4(-0.1)) q2[1] = F4(-0.1);
    tvd = 0;
  } else if (type == 5) { /* :325-329 */
    memcpy(q2, q1, 8 * sizeof(double));
    q2[2] = -q2[2];
    tvd = 0;
  } else if (type == 2 && s2 == -1) { /* :330-345 */
    double vel = nrm[0] * q1[1] + nrm[1] * q1[2] + nrm[2] * q1[3];
    double vel1[3], vel2[3];
    for (int k = 0; k < 3; k++) {
      vel1[k] = q1[1 + k] - vel * nrm[k];
      vel2[k] = vel * nrm[k];
    }
    memcpy(q2, q1, 8 * sizeof(double));
    for (int k = 0; k < 3; k++) q2[1 + k] = vel1[k] - vel2[k];
    tvd = 0;
  } else if (s2 < 1) {
    /* Types 1 and 3 with s2 = -1 reach Cell_par(:, s2) unguarded in the
     * reference (:346-354); stated here as an outflow copy ghost. */
    memcpy(q2, q1, 8 * sizeof(double));
    tvd = 0;
  } else { /* :346-354 */
    memcpy(q2, &S[NPAR * (s2 - 1)], 8 * sizeof(double));
    c2 = &m->cell_center[3 * (s2 - 1)];
  }
  int ss1 = m->gran_neighbour_tvd[2 * gr + 0];
  int ss2 = m->gran_neighbour_tvd[2 * gr + 1];
  if (cfg->recon == 1 && tvd && ss1 >= 1 && ss2 >= 1) { /* :363-412 */
    const double *css1 = &m->cell_center[3 * (ss1 - 1)];
    const double *css2 = &m->cell_center[3 * (ss2 - 1)];
    double rast[3];
    for (int k = 0; k < 3; k++) rast[k] = gc[k] - c1[k];
    double df1 = norm2v(rast);
    for (int k = 0; k < 3; k++) rast[k] = gc[k] - c2[k];
    double df2 = norm2v(rast);
    for (int k = 0; k < 3; k++) rast[k] = gc[k] - css1[k];
    double dff1 = norm2v(rast);
    for (int k = 0; k < 3; k++) rast[k] = gc[k] - css2[k];
    double dff2 = norm2v(rast);
    /* Only variables 1..5 of ss1/ss2 are read (the reference assigns the
     * 16-long column to an 8-long local, :379-384).  Variables 6..8 of
     * qqq1_TVD are never assigned in the reference; they are stated as
     * the first-order values. */
    const double *q11 = &S[NPAR * (ss1 - 1)];
    const double *q22 = &S[NPAR * (ss2 - 1)];
    double t1[8], t2[8];
    memcpy(t1, q1, sizeof t1);
    memcpy(t2, q2, sizeof t2);
    for (int i = 0; i < 5; i++) {
      t1[i] = orc_linear(-dff1, q11[i], -df1, q1[i], df2, q2[i], 0.0);
      t2[i] = orc_linear(-dff2, q22[i], -df2, q2[i], df1, q1[i], 0.0);
    }
    if (t1[0] < 0.0) t1[0] = q1[0];
    if (t1[4] < 0.0) t1[4] = q1[4];
    if (t2[0] < 0.0) t2[0] = q2[0];
    if (t2[4] < 0.0) t2[4] = q2[4];
    memcpy(q1, t1, sizeof t1);
    memcpy(q2, t2, sizeof t2);
  }
  if (q1[0] < 0.0 || q2[0] < 0.0) return ORC_ERR_FACE_RHO; /* :414-420 */
  return 0;
}

/* One face of CUF_MGD_grans_RK2 (cuf_solvers.cuf:243-440); the Euler
 * kernel CUF_MGD_grans (:62-241) is the same with stage 0 reading W.
 * Returns 1 if the face was skipped (inner-sphere test), 0 if computed,
 * <0 on a reference STOP condition. */
static int face_one(const orc_config *cfg, const orc_mesh *m, int gr,
                    int stage, const double *W, const double *W2,
                    double *potok, double *time_out) {
  double time = cfg->dt_cap; /* 'time = 100000.0' */
  double q1[8], q2[8];
  int st0 = orc_face_states(cfg, m, gr, stage, W, W2, q1, q2);
  if (st0 != 0) return st0;
  const double *c1 = &m->cell_center[3 * (m->gran_neighbour[2 * gr] - 1)];
  const double *gc = &m->gran_center[3 * gr];
  double distant[3] = {gc[0] - c1[0], gc[1] - c1[1], gc[2] - c1[2]};
  double dist = norm2v(distant); /* :291-293 */
  const double *nrm = &m->gran_normal[3 * gr];
  double square = m->gran_square[gr];

  int metod = cfg->riemann;
  if ((cfg->flags & ORC_F_HELIO_HACKS) && gc[0] > F4(0.22) && gc[1] < F4(0.1) &&
      fabs(nrm[0]) < F4(0.1))
    metod = 0; /* :296-298 */

  double wc = 0.0, dsl, dsp, dsc, POTOK[8];
  int st = orc_chlld(metod, nrm[0], nrm[1], nrm[2], wc, q1, q2, &dsl, &dsp,
                     &dsc, POTOK, 0);
  if (st != ORC_OK) return st;

  double smax = fabs(dsl) > fabs(dsp) ? fabs(dsl) : fabs(dsp);
  double cand = cfg->cfl * dist / (smax + fabs(wc)); /* :429 */
  time = time < cand ? time : cand;
  for (int k = 0; k < 8; k++) potok[9 * gr + k] = POTOK[k] * square;
  potok[9 * gr + 8] = 0.0;
  *time_out = time;
  return 0;
}

int orc_face_pass(const orc_config *cfg, const orc_mesh *m, int stage,
                  const double *W, const double *W2, double *potok,
                  double *dt_min, int nthreads) {
  int status = ORC_OK;
  double tmin = *dt_min;
  int reduce = (stage != 2); /* atomicmin: Euler always, RK2 stage 1 */
#pragma omp parallel for num_threads(nthreads) reduction(min : tmin) schedule(static)
  for (int gr = 0; gr < m->n_gran; gr++) {
    double t;
    int r = face_one(cfg, m, gr, stage, W, W2, potok, &t);
    if (r < 0) {
#pragma omp critical
      status = r;
    } else if (r == 0 && reduce) {
      tmin = tmin < t ? tmin : t;
    }
  }
  if (reduce) *dt_min = tmin;
  return status;
}

double orc_dt_scan(const orc_config *cfg, const orc_mesh *m, const double *W,
                   int nthreads) {
  double tmin = 1.0e15;
  double *scratch = (double *)malloc(sizeof(double) * 9 * (size_t)m->n_gran);
  orc_face_pass(cfg, m, 1, W, W, scratch, &tmin, nthreads);
  free(scratch);
  return tmin;
}

/* magnetic energy density of stored (Gaussian) B: |B|^2/(8 pi) */
static inline double emag(double bx, double by, double bz) {
  return 0.5 * (sq(bx) + sq(by) + sq(bz)) / ORC_CPI4;
}

/* One cell of CUF_MGD_cells (:443-558, stage 0) / CUF_MGD_cell_RK2
 * (:561-794, stage 1 and 2).  Writes Cell_par2 (stage 1) or Cell_par. */
static int cell_one(const orc_config *cfg, const orc_mesh *m, int gr,
                    int stage, double time, double *W, double *W2,
                    const double *potok) {
  const double ggg = ORC_GGG;
  const int mhd = !(cfg->flags & ORC_F_ZERO_BW);
  const int axi = (cfg->flags & ORC_F_AXISYM) != 0;
  const int smk = (cfg->flags & ORC_F_SOURCE_MK) != 0;
  double Volume = m->cell_volume[gr];
  const double *rr = &m->cell_center[3 * gr];
  double dist = norm2v(rr);
  const double *qqq = (stage == 2) ? &W2[NPAR * gr] : &W[NPAR * gr];
  const double *qn = &W[NPAR * gr]; /* qqq2 = Cell_par in stage 2 */
  if (dist < cfg->r_hold_cell) return 0; /* :611 / :475 */

  double source[3] = {0, 0, 0};
  if (smk) orc_source_mk(qqq, source); /* :684 */

  double POTOK[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  /* :691-703 sums slots 1..4 ('do i = 1, 4', also :491); SET's meshes hold
   * 0 in slots 5..6 (Setka.cpp:1344-1396), where summing all six is the same.
   * The six-slot sum is this build's extension for 3-D face-list meshes,
   * whose z- and z+ faces sit in slots 5 and 6. */
  for (int i = 0; i < 6; i++) {
    int j = m->cell_gran[6 * gr + i];
    if (j == 0) continue;
    const double *f = &potok[9 * (j - 1)];
    if (m->gran_neighbour[2 * (j - 1)] == gr + 1)
      for (int k = 0; k < 8; k++) POTOK[k] = POTOK[k] + f[k];
    else
      for (int k = 0; k < 8; k++) POTOK[k] = POTOK[k] - f[k];
  }

  double ro3, u3, v3, w3, p3, bx3, by3, bz3;
  double nrm = sqrt(sq(qqq[1]) + sq(qqq[2]) + sq(qqq[3]));
  double E1 = qqq[4] / (ggg - 1.0) + 0.5 * qqq[0] * sq(nrm);
  if (mhd) E1 = E1 + emag(qqq[5], qqq[6], qqq[7]);

  if (stage != 2) { /* Euler (:705-739) and Heun stage 1 (:705-739) */
    if (axi)
      ro3 = qqq[0] - time * (POTOK[0] / Volume + qqq[0] * qqq[2] / rr[1]);
    else
      ro3 = qqq[0] - time * (POTOK[0] / Volume);
    if (ro3 <= 0.0) return ORC_ERR_RHO; /* :708-711 STOP */
    double a;
    a = qqq[0] * qqq[1] - time * (POTOK[1]) / Volume;
    if (smk) a = a + time * source[0];
    if (axi) a = a - time * qqq[0] * qqq[1] * qqq[2] / rr[1];
    u3 = a / ro3;
    a = qqq[0] * qqq[2] - time * (POTOK[2]) / Volume;
    if (smk) a = a + time * source[1];
    if (axi) a = a - time * qqq[0] * qqq[2] * qqq[2] / rr[1];
    v3 = a / ro3;
    if (mhd) {
      w3 = (qqq[0] * qqq[3] - time * (POTOK[3]) / Volume) / ro3;
      bx3 = qqq[5] - time * (POTOK[5]) / Volume;
      by3 = qqq[6] - time * (POTOK[6]) / Volume;
      bz3 = qqq[7] - time * (POTOK[7]) / Volume;
    } else {
      w3 = 0.0;
      bx3 = 0.0;
      by3 = 0.0;
      bz3 = 0.0;
    }
    a = E1 - time * (POTOK[4]) / Volume;
    if (smk) a = a + time * source[2];
    if (axi)
      a = a - time * qqq[2] * (ggg * qqq[4] / (ggg - 1) + qqq[0] * (sq(qqq[1]) + sq(qqq[2])) * 0.5) / rr[1];
    a = a - 0.5 * ro3 * (sq(u3) + sq(v3) + sq(w3));
    if (mhd) a = a - emag(bx3, by3, bz3);
    p3 = a * (ggg - 1.0);
  } else { /* Heun stage 2 (:740-772) */
    double nrm2 = sqrt(sq(qn[1]) + sq(qn[2]) + sq(qn[3]));
    double E0 = qn[4] / (ggg - 1.0) + 0.5 * qn[0] * sq(nrm2);
    if (mhd) E0 = E0 + emag(qn[5], qn[6], qn[7]);
    if (axi)
      ro3 = 0.5 * (qn[0] + qqq[0] - time * (POTOK[0] / Volume + qqq[0] * qqq[2] / rr[1]));
    else
      ro3 = 0.5 * (qn[0] + qqq[0] - time * (POTOK[0] / Volume));
    if (ro3 <= 0.0) return ORC_ERR_RHO;
    double a;
    a = qn[0] * qn[1] + qqq[0] * qqq[1] - time * (POTOK[1]) / Volume;
    if (smk) a = a + time * source[0];
    if (axi) a = a - time * qqq[0] * qqq[1] * qqq[2] / rr[1];
    u3 = 0.5 * a / ro3;
    a = qn[0] * qn[2] + qqq[0] * qqq[2] - time * (POTOK[2]) / Volume;
    if (smk) a = a + time * source[1];
    if (axi) a = a - time * qqq[0] * qqq[2] * qqq[2] / rr[1];
    v3 = 0.5 * a / ro3;
    if (mhd) {
      w3 = 0.5 * (qn[0] * qn[3] + qqq[0] * qqq[3] - time * (POTOK[3]) / Volume) / ro3;
      bx3 = 0.5 * (qn[5] + qqq[5] - time * (POTOK[5]) / Volume);
      by3 = 0.5 * (qn[6] + qqq[6] - time * (POTOK[6]) / Volume);
      bz3 = 0.5 * (qn[7] + qqq[7] - time * (POTOK[7]) / Volume);
    } else {
      w3 = 0.0;
      bx3 = 0.0;
      by3 = 0.0;
      bz3 = 0.0;
    }
    a = E0 + E1 - time * (POTOK[4]) / Volume;
    if (smk) a = a + time * source[2];
    if (axi)
      a = a - time * qqq[2] * (ggg * qqq[4] / (ggg - 1) + qqq[0] * (sq(qqq[1]) + sq(qqq[2])) * 0.5) / rr[1];
    a = 0.5 * a - 0.5 * ro3 * (sq(u3) + sq(v3) + sq(w3));
    if (mhd) a = a - emag(bx3, by3, bz3);
    p3 = a * (ggg - 1.0);
  }
  if (p3 <= 0.0) p3 = F4(0.0000001); /* :777-779 */
  if ((cfg->flags & ORC_F_HELIO_HACKS) && rr[0] > F4(0.28) && rr[1] < F4(0.07) && v3 < 0.0)
    v3 = F4(0.01); /* :783-785 */

  double *out = (stage == 1) ? &W2[NPAR * gr] : &W[NPAR * gr];
  out[0] = ro3;
  out[1] = u3;
  out[2] = v3;
  out[3] = w3;
  out[4] = p3;
  out[5] = bx3;
  out[6] = by3;
  out[7] = bz3;
  return 0;
}

int orc_cell_pass(const orc_config *cfg, const orc_mesh *m, int stage,
                  double dt, double *W, double *W2, const double *potok,
                  int nthreads) {
  int status = ORC_OK;
#pragma omp parallel for num_threads(nthreads) schedule(static)
  for (int gr = 0; gr < m->n_cell; gr++) {
    int r = cell_one(cfg, m, gr, stage, dt, W, W2, potok);
    if (r < 0) {
#pragma omp critical
      status = r;
    }
  }
  return status;
}

int orc_run(const orc_config *cfg, const orc_mesh *m, double *W, double *W2,
            double *potok, double *t, double *dt_next, int nsteps,
            int nthreads) {
  for (int step = 0; step < nsteps; step++) { /* cuf_kernel.cuf:780-885 */
    double time_step = *dt_next;              /* :789 */
    *dt_next = 1.0e15;                        /* :790 */
    *t = *t + time_step;                      /* :792-794 */
    int st;
    if (cfg->time_scheme == 0) { /* :837-846 */
      st = orc_face_pass(cfg, m, 0, W, W2, potok, dt_next, nthreads);
      if (st) return st;
      st = orc_cell_pass(cfg, m, 0, time_step, W, W2, potok, nthreads);
      if (st) return st;
    } else { /* :847-877 */
      st = orc_face_pass(cfg, m, 1, W, W2, potok, dt_next, nthreads);
      if (st) return st;
      st = orc_cell_pass(cfg, m, 1, time_step, W, W2, potok, nthreads);
      if (st) return st;
      st = orc_face_pass(cfg, m, 2, W, W2, potok, dt_next, nthreads);
      if (st) return st;
      st = orc_cell_pass(cfg, m, 2, time_step, W, W2, potok, nthreads);
      if (st) return st;
    }
  }
  return ORC_OK;
}
