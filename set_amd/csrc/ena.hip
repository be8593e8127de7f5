// ena.hip -- ENA line-of-sight spectra of a heliosphere solution on gfx950.
//
// SET's post-processor (SET_ENA/SET_ENA.cpp:544-956, Culc_ENA and
// Culc_ENA_the) marches each line of sight from 5 AU outwards in steps of
// 1/30 AU.  At every sample it interpolates the 16 cell parameters
// barycentrically in the Delaunay triangle of cell centres that holds the
// sample (:242-339), then, for each of N particle speeds, updates the
// survival probability S_i *= exp(-(nu_ex + nu_ph) ds / v_i) and the ENA
// intensity j_i += nu_H f_p(-v_i l) v_i S_i ds.  The march ends at the first
// sample outside the triangulation.
//
// Here one 64-lane wave owns one line of sight and 64 of its speeds.  The
// sample positions, the triangle walk and the interpolation are the same in
// every lane of the wave (uniform control flow, broadcast loads); only the
// rate arithmetic differs per lane.  The first sample's triangle is found
// by a wave-parallel search over all triangles (the lowest index holding
// it); each later one by walking from the previous triangle across the edge
// beyond which the sample lies (a visibility walk, which ends on a Delaunay
// triangulation; it is bounded by the triangle count all the same).
#include <hip/hip_runtime.h>

#include <cmath>
#include <vector>

#include "../../include/setmhd.h"

namespace setmhd {
namespace ena {

// constants of SET_ENA.cpp: sqrtpi_ (:58, truncated as the reference has
// it), a_2 (:61), Kn_ (:63), 1 AU and 1 km/s in its units (:791-792)
constexpr double kSqrtPi = 1.77245385;
constexpr double kA2 = 0.1307345665;
constexpr double kKn = 0.4326569;
constexpr double kAe = 0.00313166;
constexpr double kKms = 0.0963358;
constexpr double kPi = 3.14159265358979323846;

// The reference's piecewise sixth-degree fits int_1_f1/f2/f3 (:426-536) on
// [0,1], (1,3], (3,5], (5,7]: coefficients of x^0..x^6.  (int_1 is used only
// for u/cp <= 7, so f1's branch beyond 7 is never reached.)
__constant__ double kFit[3][4][7] = {
    {{6.283185155644284, 0.000024846677279866114, 2.0934329078277405, 0.008055998193903208,
      -0.2355169235647438, 0.03820480582423355, 0.006992274370591744},
     {6.437524091973454, -0.6331520099380095, 3.1348881317268997, -0.8454201478027856,
      0.1004702004260311, 0.0009895488638964746, -0.000920750276197054},
     {4.4920780630505135, 2.5133093267020654, 1.1327223176567935, -0.24648691152318875,
      0.031326738629523766, -0.0021366031960331384, 0.00005954097505746697},
     {1.9138683588136232, 5.350374732905213, -0.16380205801427633, 0.06765898334856263,
      -0.011071118267864083, 0.0008673476933852199, -0.00002691859374483661}},
    {{1.328216167939543, -0.000004545681954848391, 2.537368073155103, -0.0020584991728545624,
      -0.03742568018912792, -0.010312136385277346, 0.002767736179209713},
     {1.2959616295629246, 0.1533684067037866, 2.2354849981206106, 0.3113395567715921,
      -0.21656309882941488, 0.041957500887605075, -0.0029978773724628604},
     {1.903643456971281, -1.4801836911099535, 3.973958664572268, -0.6482729779428982,
      0.07665007314658864, -0.005369758193338703, 0.00016605531871992049},
     {-4.484415105552316, 5.3747429756690766, 0.8892806582308143, 0.09767316152573671,
      -0.025704749778475783, 0.0021937998296249206, -0.00006928845984076111}},
    {{1.2938345594193854, -0.000031719847351174835, 1.3183710041280094, -0.014150512069488197,
      0.4226114681928129, -0.06985750969880078, -0.015347864048406958},
     {0.9667460440956788, 1.336271810704016, -0.8687355257991665, 1.7676868273627229,
      -0.2731222764016417, 0.004801770033831665, 0.001780776080720323},
     {4.760566734123174, -5.048204299463048, 3.332342585228025, 0.47584339615235993,
      -0.12072786272726124, 0.011870955604980658, -0.0004580199652402304},
     {9.370493362469261, -10.848615619383615, 6.423326878282571, -0.4148977656870439,
      0.025300923044176957, -0.0010108688120876522, 0.00001864423130429156}}};

// the three fits at x (one branch choice, shared)
__device__ __forceinline__ void fits(double x, double f[3]) {
  const int b = x <= 1.0 ? 0 : x <= 3.0 ? 1 : x <= 5.0 ? 2 : 3;
  const double x2 = x * x, x3 = x2 * x, x4 = x3 * x, x5 = x4 * x, x6 = x5 * x;
#pragma unroll
  for (int k = 0; k < 3; k++) {
    const double* c = kFit[k][b];
    f[k] = c[0] + c[1] * x + c[2] * x2 + c[3] * x3 + c[4] * x4 + c[5] * x5 + c[6] * x6;
  }
}

// charge-exchange rate of a particle at relative speed u in a population of
// density rho and thermal speed cp (:879-910): beyond u/cp = 7 the mean
// relative speed Velosity_1 (:341-351) and the cross-section at it, below
// that the fitted Maxwellian average int_1 (:538-542)
__device__ __forceinline__ double rate(double u, double rho, double cp) {
  const double x = u / cp;
  if (x > 7.0) {
    double uz;
    if (u < 0.00001)
      uz = 2.0 * cp / kSqrtPi + 2.0 * u * u / (3.0 * cp * kSqrtPi) - u * u * u * u / (15.0 * cp * cp * cp * kSqrtPi);
    else
      uz = exp(-u * u / (cp * cp)) * cp / kSqrtPi + (u + cp * cp / (2.0 * u)) * erf(u / cp);
    const double s = 1.0 - kA2 * log(uz);
    return rho * uz * (s * s) / kKn;
  }
  double f[3];
  fits(x, f);
  const double b = 1.0 - kA2 * log(cp);
  const double i1 = (cp / (kSqrtPi * kSqrtPi * kSqrtPi)) * (b * b * f[0] - 2.0 * kA2 * b * f[1] + (kA2 * kA2) * f[2]);
  return rho * i1 / kKn;
}

struct Args {
  int n_cell, n_tri, n_angle, n_v, chunks;
  long long max_samples;  // more samples than the hull's diameter allows: stop, flag
  const double* cc;   // [3 n_cell]
  const double* par;  // [16 n_cell]
  const int* tri;     // [3 n_tri]
  const int* nb;      // [3 n_tri]: neighbour across the edge opposite vertex k, -1 on the hull
  const double* angle;
  double* jena;       // [n_angle n_v]
  int* samples;       // [n_angle]
  int* err;
};

__device__ __forceinline__ double area2(double px, double py, double qx, double qy, double rx, double ry) {
  return ((qx - px) * (ry - py) - (rx - px) * (qy - py)) / 2.0;  // CGAL::area
}

// barycentric coordinates of (x, y) in triangle t
__device__ __forceinline__ void bary(const Args& a, int t, double x, double y, double c[3], int v[3]) {
  v[0] = a.tri[3 * t];
  v[1] = a.tri[3 * t + 1];
  v[2] = a.tri[3 * t + 2];
  const double *p0 = a.cc + 3 * v[0], *p1 = a.cc + 3 * v[1], *p2 = a.cc + 3 * v[2];
  const double ar = area2(p0[0], p0[1], p1[0], p1[1], p2[0], p2[1]);
  c[0] = area2(x, y, p1[0], p1[1], p2[0], p2[1]) / ar;
  c[1] = area2(p0[0], p0[1], x, y, p2[0], p2[1]) / ar;
  c[2] = area2(p0[0], p0[1], p1[0], p1[1], x, y) / ar;
}

constexpr double kInsideEps = 1e-12;  // :316-320

__global__ __launch_bounds__(256) void k_ena_los(Args a) {
  const int wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int lane = threadIdx.x & 63;
  if (wave >= a.n_angle * a.chunks) return;  // whole waves leave together
  const int ang = wave / a.chunks, chunk = wave % a.chunks;
  const int i = chunk * 64 + lane;
  const bool mine = i < a.n_v;

  // the line of sight (:787-826)
  double lx = cos(a.angle[ang]), ly = sin(a.angle[ang]);
  const double nrm = sqrt(0.0 * 0.0 + lx * lx + ly * ly);
  lx /= nrm;
  ly /= nrm;
  const double ds = kAe / 30.0;
  double x0 = 0.0 + 5.0 * lx * kAe, y0 = 0.0 + 5.0 * ly * kAe;
  x0 = x0 + lx * ds / 2.0;
  y0 = y0 + ly * ds / 2.0;
  const double vL = 1 * kKms, vR = 1000 * kKms;
  const double v = vL + (i + 0.5) * ((vR - vL) / a.n_v);

  // first sample: the lowest-numbered triangle holding it
  int t = 0x7fffffff;
  for (int k = lane; k < a.n_tri; k += 64) {
    double c[3];
    int vv[3];
    bary(a, k, x0, y0, c, vv);
    if (c[0] >= -kInsideEps && c[1] >= -kInsideEps && c[2] >= -kInsideEps) {
      t = k;
      break;
    }
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) t = min(t, __shfl_xor(t, o, 64));
  t = __builtin_amdgcn_readfirstlane(t);  // uniform: triangle data by scalar loads

  double S = 1.0, j = 0.0;
  int n = 0;
  while (t != 0x7fffffff) {
    // locate the sample: walk from t (uniform across the wave)
    double c[3];
    int vv[3];
    int moves = 0;
    while (true) {
      bary(a, t, x0, y0, c, vv);
      int k = c[0] <= c[1] ? 0 : 1;
      if (c[2] < c[k]) k = 2;
      if (c[k] >= -kInsideEps) break;
      const int nt = a.nb[3 * t + k];
      if (nt < 0 || ++moves > a.n_tri) {
        if (nt >= 0 && lane == 0) atomicOr(a.err, 1);
        t = -1;
        break;
      }
      t = __builtin_amdgcn_readfirstlane(nt);
    }
    if (t < 0) break;
    // the 16 parameters at the sample; those the rates use (:855-871)
    auto at = [&](int p) {
      return c[0] * a.par[16 * vv[0] + p] + c[1] * a.par[16 * vv[1] + p] + c[2] * a.par[16 * vv[2] + p];
    };
    const double rho_p = at(0), u_p = at(1), v_p = at(2), p_p = at(4);
    const double cp_p = sqrt(p_p / rho_p);
    const double rho_H3 = at(8) * 3.0, u_H3 = at(9), v_H3 = at(10);
    const double cp_H3 = sqrt(0.5 * rho_H3 * at(11) / rho_H3);
    const double rho_H4 = at(12) * 3.0, u_H4 = at(13), v_H4 = at(14);
    const double cp_H4 = sqrt(0.5 * rho_H4 * at(15) / rho_H4);
    const double r = sqrt(x0 * x0 + y0 * y0);
    const double nu_ph = 770.291 * ((kAe / r) * (kAe / r));
    if (++n > a.max_samples) {
      if (lane == 0) atomicOr(a.err, 2);
      break;
    }
    if (mine) {
      const double wx = -v * lx, wy = -v * ly;  // the particle's velocity
      double du = wx - u_p, dv = wy - v_p;
      const double nu_ex = rate(sqrt(du * du + dv * dv + 0.0 * 0.0), rho_p, cp_p);
      du = wx - u_H3;
      dv = wy - v_H3;
      const double nu_3 = rate(sqrt(du * du + dv * dv + 0.0 * 0.0), rho_H3, cp_H3);
      du = wx - u_H4;
      dv = wy - v_H4;
      const double nu_4 = rate(sqrt(du * du + dv * dv + 0.0 * 0.0), rho_H4, cp_H4);
      const double nu_H = nu_3 + nu_4;
      S = S * exp(-(nu_ex + nu_ph) * ds / v);
      // maxwell_distribution of the plasma at the particle's velocity (:353-370)
      const double mx = wx - u_p, my = wy - v_p, mz = 0.0 - 0.0;
      const double v_sq = mx * mx + my * my + mz * mz;
      const double f = rho_p / (kPi * sqrt(kPi) * cp_p * cp_p * cp_p) * exp(-v_sq / (cp_p * cp_p));
      j = j + nu_H * f * v * S * ds;
    }
    x0 = x0 + lx * ds;
    y0 = y0 + ly * ds;
  }
  if (mine) a.jena[(size_t)ang * a.n_v + i] = j;
  if (chunk == 0 && lane == 0) a.samples[ang] = n;
}

}  // namespace ena
}  // namespace setmhd

using namespace setmhd;

extern "C" int setmhd_ena_spectra(int32_t device, int32_t n_cell, const double* cell_center,
                                  const double* cell_par, int32_t n_tri, const int32_t* tri,
                                  const int32_t* tri_nb, int32_t n_angle, const double* angle, int32_t n_v,
                                  double* jena, int32_t* samples) {
  if (n_cell < 3 || n_tri < 1 || n_angle < 1 || n_v < 1 || !cell_center || !cell_par || !tri || !tri_nb ||
      !angle || !jena || !samples)
    return SETMHD_ERR_ARG;
  for (long long k = 0; k < 3LL * n_tri; k++)
    if (tri[k] < 0 || tri[k] >= n_cell || tri_nb[k] < -1 || tri_nb[k] >= n_tri) return SETMHD_ERR_ARG;
  if (hipSetDevice(device) != hipSuccess) return SETMHD_ERR_HIP;
  const size_t b_cc = 3 * (size_t)n_cell * 8, b_par = 16 * (size_t)n_cell * 8, b_tri = 3 * (size_t)n_tri * 4;
  const size_t b_ang = (size_t)n_angle * 8, b_out = (size_t)n_angle * n_v * 8, b_smp = (size_t)n_angle * 4;
  const size_t total = b_cc + b_par + 2 * b_tri + b_ang + b_out + b_smp + 8 + 64;
  char* base = nullptr;
  if (hipMalloc((void**)&base, total) != hipSuccess) return SETMHD_ERR_NOMEM;
  size_t off = 0;
  auto take = [&](size_t n) {
    char* p = base + off;
    off += (n + 7) & ~size_t(7);
    return p;
  };
  ena::Args a{};
  a.n_cell = n_cell;
  a.n_tri = n_tri;
  a.n_angle = n_angle;
  a.n_v = n_v;
  a.chunks = (n_v + 63) / 64;
  {  // a line of sight inside the hull of the cell centres spans at most its diameter
    double lo[2] = {cell_center[0], cell_center[1]}, hi[2] = {cell_center[0], cell_center[1]};
    for (int32_t c = 1; c < n_cell; c++)
      for (int d = 0; d < 2; d++) {
        lo[d] = std::fmin(lo[d], cell_center[3 * (size_t)c + d]);
        hi[d] = std::fmax(hi[d], cell_center[3 * (size_t)c + d]);
      }
    const double diam = std::sqrt((hi[0] - lo[0]) * (hi[0] - lo[0]) + (hi[1] - lo[1]) * (hi[1] - lo[1]));
    const double n = diam / (ena::kAe / 30.0) + 4.0;
    if (!(n < 1e9)) return SETMHD_ERR_ARG;
    a.max_samples = (long long)n;
  }
  double* d_cc = (double*)take(b_cc);
  double* d_par = (double*)take(b_par);
  int* d_tri = (int*)take(b_tri);
  int* d_nb = (int*)take(b_tri);
  double* d_ang = (double*)take(b_ang);
  double* d_out = (double*)take(b_out);
  int* d_smp = (int*)take(b_smp);
  int* d_err = (int*)take(8);
  a.cc = d_cc;
  a.par = d_par;
  a.tri = d_tri;
  a.nb = d_nb;
  a.angle = d_ang;
  a.jena = d_out;
  a.samples = d_smp;
  a.err = d_err;
  int st = SETMHD_OK;
  hipError_t e = hipMemcpy(d_cc, cell_center, b_cc, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(d_par, cell_par, b_par, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(d_tri, tri, b_tri, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(d_nb, tri_nb, b_tri, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(d_ang, angle, b_ang, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemset(d_err, 0, 8);
  if (e == hipSuccess) {
    const long long waves = (long long)n_angle * a.chunks;
    const unsigned blocks = (unsigned)((waves * 64 + 255) / 256);
    hipLaunchKernelGGL(ena::k_ena_los, dim3(blocks), dim3(256), 0, 0, a);
    e = hipGetLastError();
  }
  int h_err = 0;
  if (e == hipSuccess) e = hipMemcpy(jena, d_out, b_out, hipMemcpyDeviceToHost);
  if (e == hipSuccess) e = hipMemcpy(samples, d_smp, b_smp, hipMemcpyDeviceToHost);
  if (e == hipSuccess) e = hipMemcpy(&h_err, d_err, sizeof h_err, hipMemcpyDeviceToHost);
  if (e != hipSuccess) st = SETMHD_ERR_HIP;
  else if (h_err) st = SETMHD_ERR_ARG;  // a walk did not end: not a triangulation with consistent neighbours
  (void)hipFree(base);
  return st;
}
