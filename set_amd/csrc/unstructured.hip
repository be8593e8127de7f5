// unstructured.hip -- drop-in twins of SET's face and cell kernels on the
// reference's face-list mesh (SET_Fortran_Cuda_MHD/cuf_solvers.cuf).
//
//   k_face_pass  = CUF_MGD_grans (Euler, :62-241) / CUF_MGD_grans_RK2 (:243-440)
//   k_cell_pass  = CUF_MGD_cells (Euler, :443-558) / CUF_MGD_cell_RK2 (:561-794)
//
// One thread per face / per cell, as in the reference, with the same
// arithmetic in the same order (contraction off, IEEE div/sqrt), so a run
// reproduces the CPU restatement of the reference bit for bit.  Device
// layout differs from the reference where that costs nothing in meaning:
// primitives are kept as 8 doubles per cell (one 64-byte gather), the
// neutral-H moments (Cell_par(9:16)) in their own array, face fluxes as 8
// doubles per face (Gran_POTOK(9) is identically 0, cuf_solvers.cuf:433).
// The global CFL minimum is a wave64 shuffle reduction plus one atomic per
// block instead of one atomicmin per thread (:436-438).
#pragma clang fp contract(off)

#include "chlld_dev.hpp"
#include "kernels.hpp"

namespace setmhd {

__device__ __forceinline__ double norm2_3(double a, double b, double c) {
  return sqrt(a * a + b * b + c * c);  // dev_norm2, cuf_kernel.cuf:196-203
}

// Source_MK, cuf_kernel.cuf:215-288
__device__ void source_mk(const double* q, const double* h, double out[3]) {
  const double pi = 3.141592653589793115997963468544185161590576171875;  // acos(-1)
  double rho = q[0], u = q[1], v = q[2], p = q[4];
  double rhoH3 = h[0], uH3 = h[1], vH3 = h[2];
  double pH3 = 0.5 * h[3] * rhoH3;
  if (rhoH3 < SETMHD_F4(0.0000001)) {
    rhoH3 = SETMHD_F4(0.0000001);
    pH3 = 0.0;
  }
  double rhoH4 = h[4], uH4 = h[5], vH4 = h[6];
  double pH4 = 0.5 * h[7] * rhoH4;
  if (rhoH4 < SETMHD_F4(0.0000001)) {
    rhoH4 = SETMHD_F4(0.0000001);
    pH4 = 0.0;
  }
  const double c64 = 64.0 / (9.0 * pi), c4 = 4.0 / (pi);
  const double ksig = SETMHD_F4(0.1307345665);
  const double kk = (double)(3.0f / 0.4326569f);
  double U_M_H3 = sqrt((u - uH3) * (u - uH3) + (v - vH3) * (v - vH3) +
                       c64 * (p / rho + 2.0 * pH3 / rhoH3));
  double U_H3 = sqrt((u - uH3) * (u - uH3) + (v - vH3) * (v - vH3) +
                     c4 * (p / rho + 2.0 * pH3 / rhoH3));
  double s3 = 1.0 - ksig * log(U_M_H3);
  double nu_H3 = rho * rhoH3 * U_M_H3 * (s3 * s3);
  double U_M_H4 = sqrt((u - uH4) * (u - uH4) + (v - vH4) * (v - vH4) +
                       c64 * (p / rho + 2.0 * pH4 / rhoH4));
  double U_H4 = sqrt((u - uH4) * (u - uH4) + (v - vH4) * (v - vH4) +
                     c4 * (p / rho + 2.0 * pH4 / rhoH4));
  double s4 = 1.0 - ksig * log(U_M_H4);
  double nu_H4 = rho * rhoH4 * U_M_H4 * (s4 * s4);
  double sum1 = 0.0, sum2 = 0.0, sum3 = 0.0;
  sum1 = sum1 + kk * nu_H3 * (uH3 - u);
  sum2 = sum2 + kk * nu_H3 * (vH3 - v);
  sum3 = sum3 + kk * nu_H3 * ((uH3 * uH3 + vH3 * vH3 - u * u - v * v) / 2.0 +
                              (U_H3 / U_M_H3) * (2.0 * pH3 / rhoH3 - p / rho));
  sum1 = sum1 + kk * nu_H4 * (uH4 - u);
  sum2 = sum2 + kk * nu_H4 * (vH4 - v);
  sum3 = sum3 + kk * nu_H4 * ((uH4 * uH4 + vH4 * vH4 - u * u - v * v) / 2.0 +
                              (U_H4 / U_M_H4) * (2.0 * pH4 / rhoH4 - p / rho));
  out[0] = sum1;
  out[1] = sum2;
  out[2] = sum3;
}

// One face.  Returns 1 if skipped, 0 if computed, <0 on a STOP condition.
__device__ int face_one(const UMesh& m, const UCfg& c, int gr, const double* __restrict__ S,
                        double* __restrict__ potok, double* time_out) {
  double time = c.dt_cap;  // 'time = 100000.0'
  int type = m.gran_type[gr];
  int s1 = m.gran_nb[2 * gr + 0];
  int s2 = m.gran_nb[2 * gr + 1];
  const double* c1 = &m.cell_center[3 * (s1 - 1)];
  if (norm2_3(c1[0], c1[1], c1[2]) < c.r_skip) return 1;  // :282-284
  const double* gc = &m.gran_center[3 * gr];
  double dist = norm2_3(gc[0] - c1[0], gc[1] - c1[1], gc[2] - c1[2]);  // :291-293
  const double* nr = &m.gran_normal[3 * gr];
  double n0 = nr[0], n1 = nr[1], n2 = nr[2];
  int metod = c.riemann;
  if ((c.flags & kHelioHacks) && gc[0] > SETMHD_F4(0.22) && gc[1] < SETMHD_F4(0.1) &&
      fabs(n0) < SETMHD_F4(0.1))
    metod = 0;  // :296-298

  double q1[8], q2[8];
#pragma unroll
  for (int k = 0; k < 8; k++) q1[k] = S[8 * (s1 - 1) + k];
  const double* c2 = c1;
  bool tvd = true;
  if (type == 4) {  // :318-324
#pragma unroll
    for (int k = 0; k < 8; k++) q2[k] = q1[k];
    if ((c.flags & kHelioHacks) && q2[1] > SETMHD_F4(-0.1)) q2[1] = SETMHD_F4(-0.1);
    tvd = false;
  } else if (type == 5) {  // :325-329
#pragma unroll
    for (int k = 0; k < 8; k++) q2[k] = q1[k];
    q2[2] = -q2[2];
    tvd = false;
  } else if (type == 2 && s2 == -1) {  // :330-345
    double vel = n0 * q1[1] + n1 * q1[2] + n2 * q1[3];
    double v10 = q1[1] - vel * n0, v11 = q1[2] - vel * n1, v12 = q1[3] - vel * n2;
    double v20 = vel * n0, v21 = vel * n1, v22 = vel * n2;
#pragma unroll
    for (int k = 0; k < 8; k++) q2[k] = q1[k];
    q2[1] = v10 - v20;
    q2[2] = v11 - v21;
    q2[3] = v12 - v22;
    tvd = false;
  } else if (s2 < 1) {  // unguarded in the reference (:346-354): copy ghost
#pragma unroll
    for (int k = 0; k < 8; k++) q2[k] = q1[k];
    tvd = false;
  } else {  // :346-354
#pragma unroll
    for (int k = 0; k < 8; k++) q2[k] = S[8 * (s2 - 1) + k];
    c2 = &m.cell_center[3 * (s2 - 1)];
  }
  double square = m.gran_square[gr];
  int ss1 = m.gran_tvd[2 * gr + 0];
  int ss2 = m.gran_tvd[2 * gr + 1];
  if (c.recon == 1 && tvd && ss1 >= 1 && ss2 >= 1) {  // :363-412
    const double* cs1 = &m.cell_center[3 * (ss1 - 1)];
    const double* cs2 = &m.cell_center[3 * (ss2 - 1)];
    double df1 = norm2_3(gc[0] - c1[0], gc[1] - c1[1], gc[2] - c1[2]);
    double df2 = norm2_3(gc[0] - c2[0], gc[1] - c2[1], gc[2] - c2[2]);
    double dff1 = norm2_3(gc[0] - cs1[0], gc[1] - cs1[1], gc[2] - cs1[2]);
    double dff2 = norm2_3(gc[0] - cs2[0], gc[1] - cs2[1], gc[2] - cs2[2]);
    const double* q11 = &S[8 * (ss1 - 1)];
    const double* q22 = &S[8 * (ss2 - 1)];
    double t1[5], t2[5];
#pragma unroll
    for (int i = 0; i < 5; i++) {
      t1[i] = linear_ref(-dff1, q11[i], -df1, q1[i], df2, q2[i], 0.0);
      t2[i] = linear_ref(-dff2, q22[i], -df2, q2[i], df1, q1[i], 0.0);
    }
    if (t1[0] < 0.0) t1[0] = q1[0];
    if (t1[4] < 0.0) t1[4] = q1[4];
    if (t2[0] < 0.0) t2[0] = q2[0];
    if (t2[4] < 0.0) t2[4] = q2[4];
#pragma unroll
    for (int i = 0; i < 5; i++) {
      q1[i] = t1[i];
      q2[i] = t2[i];
    }
  }
  if (q1[0] < 0.0 || q2[0] < 0.0) return -3;  // :414-420
  double dsl, dsp, dsc, F[8];
  int st = chlld_ref(metod, n0, n1, n2, 0.0, q1, q2, &dsl, &dsp, &dsc, F, 0);
  if (st != 0) return st;
  double smax = fabs(dsl) > fabs(dsp) ? fabs(dsl) : fabs(dsp);
  double cand = c.cfl * dist / (smax + fabs(0.0));  // :429
  time = time < cand ? time : cand;
#pragma unroll
  for (int k = 0; k < 8; k++) potok[8 * gr + k] = F[k] * square;
  *time_out = time;
  return 0;
}

// stage 0: Euler, 1/2: Heun.  reduce: fold the CFL candidates into the
// next dt slot.  book: this launch opens a step (t += dt, re-arm slot).
__global__ __launch_bounds__(256) void k_face_pass(UMesh m, UCfg c, const double* __restrict__ S,
                                                   double* __restrict__ potok, DtRing dt, int cur,
                                                   int reduce, int book, int* err) {
  __shared__ double red[4];
  int gr = blockIdx.x * blockDim.x + threadIdx.x;
  double tmin = kDtReset;
  if (book && gr == 0) {
    dt.t[0] = dt.t[0] + dt.slot[cur];       // cuf_kernel.cuf:792-794
    dt.slot[(cur + 2) % 3] = kDtReset;      // :790 for the step after next
  }
  if (gr < m.n_gran) {
    double t;
    int r = face_one(m, c, gr, S, potok, &t);
    if (r < 0) atomicOr(err, r == -2 ? kErrRegion : kErrFaceRho);
    else if (r == 0) tmin = t;
  }
  if (!reduce) return;
  tmin = wave_min(tmin);
  int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) red[wid] = tmin;
  __syncthreads();
  if (threadIdx.x == 0) {
    double b = red[0];
    for (int w = 1; w < (int)(blockDim.x >> 6); w++) b = fmin(b, red[w]);
    if (b < kDtReset) atomic_min_pos(&dt.slot[(cur + 1) % 3], b);
  }
}

__device__ __forceinline__ double emag(double bx, double by, double bz) {
  return 0.5 * (bx * bx + by * by + bz * bz) / kCpi4;
}

// One cell: CUF_MGD_cells (stage 0, in place) / CUF_MGD_cell_RK2 (1: W -> W2,
// 2: 1/2 (W + W2 - dt L(W2)) -> W).
__global__ __launch_bounds__(256) void k_cell_pass(UMesh m, UCfg c, int stage, double* __restrict__ W,
                                                   double* __restrict__ W2,
                                                   const double* __restrict__ H,
                                                   const double* __restrict__ potok, DtRing dt,
                                                   int cur, int* err) {
  int gr = blockIdx.x * blockDim.x + threadIdx.x;
  if (gr >= m.n_cell) return;
  const double ggg = kGamma;
  const bool mhd = !(c.flags & kZeroBW);
  const bool axi = (c.flags & kAxisym) != 0;
  const bool smk = (c.flags & kSourceMK) != 0;
  double time = dt.slot[cur];  // time_step
  double Volume = m.cell_volume[gr];
  const double* rr = &m.cell_center[3 * gr];
  double dist = norm2_3(rr[0], rr[1], rr[2]);
  if (dist < c.r_hold) return;  // :611
  double q[8], qn[8];
  const double* src = (stage == 2) ? W2 : W;
#pragma unroll
  for (int k = 0; k < 8; k++) q[k] = src[8 * gr + k];
  if (stage == 2) {
#pragma unroll
    for (int k = 0; k < 8; k++) qn[k] = W[8 * gr + k];
  }
  double source[3] = {0.0, 0.0, 0.0};
  if (smk) source_mk(q, &H[8 * gr], source);  // :684

  double P[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (int i = 0; i < 6; i++) {  // :691-703
    int j = m.cell_gran[6 * gr + i];
    if (j == 0) continue;
    const double* f = &potok[8 * (j - 1)];
    if (m.gran_nb[2 * (j - 1)] == gr + 1) {
#pragma unroll
      for (int k = 0; k < 8; k++) P[k] = P[k] + f[k];
    } else {
#pragma unroll
      for (int k = 0; k < 8; k++) P[k] = P[k] - f[k];
    }
  }
  double ro3, u3, v3, w3, p3, bx3, by3, bz3;
  double nrm = norm2_3(q[1], q[2], q[3]);
  double E1 = q[4] / (ggg - 1.0) + 0.5 * q[0] * (nrm * nrm);
  if (mhd) E1 = E1 + emag(q[5], q[6], q[7]);
  double axE = 0.0;
  if (axi) axE = time * q[2] * (ggg * q[4] / (ggg - 1) + q[0] * (q[1] * q[1] + q[2] * q[2]) * 0.5) / rr[1];
  if (stage != 2) {  // :705-739
    if (axi) ro3 = q[0] - time * (P[0] / Volume + q[0] * q[2] / rr[1]);
    else ro3 = q[0] - time * (P[0] / Volume);
    if (ro3 <= 0.0) {
      atomicOr(err, kErrRho);
      return;
    }
    double a = q[0] * q[1] - time * (P[1]) / Volume;
    if (smk) a = a + time * source[0];
    if (axi) a = a - time * q[0] * q[1] * q[2] / rr[1];
    u3 = a / ro3;
    a = q[0] * q[2] - time * (P[2]) / Volume;
    if (smk) a = a + time * source[1];
    if (axi) a = a - time * q[0] * q[2] * q[2] / rr[1];
    v3 = a / ro3;
    if (mhd) {
      w3 = (q[0] * q[3] - time * (P[3]) / Volume) / ro3;
      bx3 = q[5] - time * (P[5]) / Volume;
      by3 = q[6] - time * (P[6]) / Volume;
      bz3 = q[7] - time * (P[7]) / Volume;
    } else {
      w3 = bx3 = by3 = bz3 = 0.0;
    }
    a = E1 - time * (P[4]) / Volume;
    if (smk) a = a + time * source[2];
    if (axi) a = a - axE;
    a = a - 0.5 * ro3 * (u3 * u3 + v3 * v3 + w3 * w3);
    if (mhd) a = a - emag(bx3, by3, bz3);
    p3 = a * (ggg - 1.0);
  } else {  // :740-772
    double nrm0 = norm2_3(qn[1], qn[2], qn[3]);
    double E0 = qn[4] / (ggg - 1.0) + 0.5 * qn[0] * (nrm0 * nrm0);
    if (mhd) E0 = E0 + emag(qn[5], qn[6], qn[7]);
    if (axi) ro3 = 0.5 * (qn[0] + q[0] - time * (P[0] / Volume + q[0] * q[2] / rr[1]));
    else ro3 = 0.5 * (qn[0] + q[0] - time * (P[0] / Volume));
    if (ro3 <= 0.0) {
      atomicOr(err, kErrRho);
      return;
    }
    double a = qn[0] * qn[1] + q[0] * q[1] - time * (P[1]) / Volume;
    if (smk) a = a + time * source[0];
    if (axi) a = a - time * q[0] * q[1] * q[2] / rr[1];
    u3 = 0.5 * a / ro3;
    a = qn[0] * qn[2] + q[0] * q[2] - time * (P[2]) / Volume;
    if (smk) a = a + time * source[1];
    if (axi) a = a - time * q[0] * q[2] * q[2] / rr[1];
    v3 = 0.5 * a / ro3;
    if (mhd) {
      w3 = 0.5 * (qn[0] * qn[3] + q[0] * q[3] - time * (P[3]) / Volume) / ro3;
      bx3 = 0.5 * (qn[5] + q[5] - time * (P[5]) / Volume);
      by3 = 0.5 * (qn[6] + q[6] - time * (P[6]) / Volume);
      bz3 = 0.5 * (qn[7] + q[7] - time * (P[7]) / Volume);
    } else {
      w3 = bx3 = by3 = bz3 = 0.0;
    }
    a = E0 + E1 - time * (P[4]) / Volume;
    if (smk) a = a + time * source[2];
    if (axi) a = a - axE;
    a = 0.5 * a - 0.5 * ro3 * (u3 * u3 + v3 * v3 + w3 * w3);
    if (mhd) a = a - emag(bx3, by3, bz3);
    p3 = a * (ggg - 1.0);
  }
  if (p3 <= 0.0) p3 = SETMHD_F4(0.0000001);  // :777-779
  if ((c.flags & kHelioHacks) && rr[0] > SETMHD_F4(0.28) && rr[1] < SETMHD_F4(0.07) && v3 < 0.0)
    v3 = SETMHD_F4(0.01);  // :783-785
  double* out = (stage == 1) ? &W2[8 * gr] : &W[8 * gr];
  out[0] = ro3;
  out[1] = u3;
  out[2] = v3;
  out[3] = w3;
  out[4] = p3;
  out[5] = bx3;
  out[6] = by3;
  out[7] = bz3;
}

// ---- host-side launchers ---------------------------------------------------
hipError_t launch_face_pass(hipStream_t s, const UMesh& m, const UCfg& c, const double* S,
                            double* potok, DtRing dt, int cur, int reduce, int book, int* err) {
  int nb = (m.n_gran + 255) / 256;
  if (nb < 1) nb = 1;
  hipLaunchKernelGGL(k_face_pass, dim3(nb), dim3(256), 0, s, m, c, S, potok, dt, cur, reduce,
                     book, err);
  return hipGetLastError();
}

hipError_t launch_cell_pass(hipStream_t s, const UMesh& m, const UCfg& c, int stage, double* W,
                            double* W2, const double* H, const double* potok, DtRing dt, int cur,
                            int* err) {
  int nb = (m.n_cell + 255) / 256;
  if (nb < 1) return hipSuccess;
  hipLaunchKernelGGL(k_cell_pass, dim3(nb), dim3(256), 0, s, m, c, stage, W, W2, H, potok, dt,
                     cur, err);
  return hipGetLastError();
}

}  // namespace setmhd
