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
// Four passes.  (0) One thread per (triangle, line of sight) finds the
// lowest-numbered triangle holding each line's first sample.  (1) One wave
// per line of sight records where each segment of seg_len samples starts:
// its position, accumulated sample by sample as the march accumulates it,
// and its triangle, reached by walking from the previous segment's across
// the edge beyond which the point lies (a visibility walk, which ends on a
// Delaunay triangulation; it is bounded by the triangle count all the
// same), and the exact sample count.  (2) One wave per (line of sight,
// segment, 64 SPL speeds) marches its segment from S = 1, walking from
// sample to sample: the positions, the walk, the interpolation and the
// speed-independent parts of the rates are the same in every lane (uniform
// control flow, scalar loads) and are shared by the lane's SPL speeds; the
// rates proper differ per speed.  (3) The survival is a product along the
// line, so the spectrum is the sum over segments of each segment's
// intensity times the survival of the segments before it.  The segments
// give the chip many more waves than lines of sight x speeds, and long and
// short lines balance; only the order of the products differs from one
// march along the whole line.
#include <hip/hip_runtime.h>

#include <algorithm>
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

// A population the rates see at one sample: density, bulk velocity,
// thermal speed, and the speed-independent factors of int_1 (:538-542):
// b = 1 - a_2 ln cp enters as b^2 and 2 a_2 b, the prefactor cp/sqrtpi^3.
struct Pop {
  double rho, u, v, cp, icp, bb, b2a, cpf;
};

__device__ __forceinline__ Pop population(double rho, double u, double v, double cp) {
  Pop p;
  p.rho = rho;
  p.u = u;
  p.v = v;
  p.cp = cp;
  p.icp = 1.0 / cp;
  const double b = 1.0 - kA2 * log(cp);
  p.bb = b * b;
  p.b2a = 2.0 * kA2 * b;
  p.cpf = cp / (kSqrtPi * kSqrtPi * kSqrtPi);
  return p;
}

// charge-exchange rate of a particle at relative speed u in population P
// (:879-910): beyond u/cp = 7 the mean relative speed Velosity_1
// (:341-351) and the cross-section at it, below that the fitted Maxwellian
// average int_1 (:538-542).  The divisions by cp and by Kn_ are
// multiplications by reciprocals formed once (within an ulp of the
// quotient; the spectra move by 1e-15 of their peaks).
constexpr double kInvKn = 1.0 / kKn;
__device__ __forceinline__ double rate(double u, const Pop& P) {
  const double cp = P.cp;
  const double x = u * P.icp;
  if (x > 7.0) {
    double uz;
    if (u < 0.00001)
      uz = 2.0 * cp / kSqrtPi + 2.0 * u * u / (3.0 * cp * kSqrtPi) - u * u * u * u / (15.0 * cp * cp * cp * kSqrtPi);
    else  // (:349) exp(-x^2) cp/sqrt(pi) + (u + cp^2/(2u)) erf(x) at x > 7: the first
          // term is below 3e-22 cp, under half an ulp of the second (> 7 cp),
          // and erf(x) rounds to 1 (erfc(7) = 4e-23; libm and the device
          // library both return 1 from x = 6 on), so this is the same double
      uz = u + cp * cp / (2.0 * u);
    const double s = 1.0 - kA2 * log(uz);
    return P.rho * uz * (s * s) * kInvKn;
  }
  double f[3];
  fits(x, f);
  const double i1 = P.cpf * (P.bb * f[0] - P.b2a * f[1] + (kA2 * kA2) * f[2]);
  return P.rho * i1 * kInvKn;
}

struct Args {
  int n_cell, n_tri, n_angle, n_v, chunks;
  int seg_len, n_seg;     // samples per segment, segments per line of sight
  long long max_samples;  // more samples than the hull's diameter allows: stop, flag
  const double* cc;   // [3 n_cell]
  const double* par;  // [16 n_cell]
  const int* tri;     // [3 n_tri]
  const int* nb;      // [3 n_tri]: neighbour across the edge opposite vertex k, -1 on the hull
  const double* angle;
  int* first;         // [n_angle]: lowest-numbered triangle holding the first sample
  int* seg_tri;       // [n_angle n_seg]: triangle of each segment's first sample
  double* seg_xy;     // [n_angle n_seg 2]: its position
  double* seg_j;      // [n_angle n_seg n_v]: the segment's intensity with S = 1 at its start
  double* seg_s;      // [n_angle n_seg n_v]: the segment's survival factor
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
constexpr int kNone = 0x7fffffff;

// Locate (x, y) by walking from triangle t (wave-uniform) across the edge
// with the most negative barycentric coordinate.  Returns the triangle
// holding the point with its coordinates and vertices, or -1 when the walk
// leaves the hull or makes more moves than there are triangles (then the
// neighbour table is inconsistent: flagged).
__device__ __forceinline__ int walk(const Args& a, int t, double x, double y, double c[3], int v[3]) {
  int moves = 0;
  while (true) {
    bary(a, t, x, y, c, v);
    int k = c[0] <= c[1] ? 0 : 1;
    if (c[2] < c[k]) k = 2;
    if (c[k] >= -kInsideEps) return t;
    const int nt = a.nb[3 * t + k];
    if (nt < 0) return -1;
    if (++moves > a.n_tri) {
      if ((threadIdx.x & 63) == 0) atomicOr(a.err, 1);
      return -1;
    }
    t = __builtin_amdgcn_readfirstlane(nt);
  }
}

// the line of sight of angle `ang` (:787-826): direction and first sample
__device__ __forceinline__ void line_of_sight(const Args& a, int ang, double& lx, double& ly, double& x0,
                                              double& y0) {
  lx = cos(a.angle[ang]);
  ly = sin(a.angle[ang]);
  const double nrm = sqrt(0.0 * 0.0 + lx * lx + ly * ly);
  lx /= nrm;
  ly /= nrm;
  const double ds = kAe / 30.0;
  x0 = 0.0 + 5.0 * lx * kAe;
  y0 = 0.0 + 5.0 * ly * kAe;
  x0 = x0 + lx * ds / 2.0;
  y0 = y0 + ly * ds / 2.0;
}

// Pass 0, one thread per (triangle, line of sight): the lowest-numbered
// triangle holding each line's first sample (first[] starts at kNone).
__global__ __launch_bounds__(256) void k_ena_first(Args a) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= a.n_tri) return;
  for (int ang = blockIdx.y; ang < a.n_angle; ang += gridDim.y) {
    double lx, ly, x0, y0;
    line_of_sight(a, ang, lx, ly, x0, y0);
    double c[3];
    int vv[3];
    bary(a, k, x0, y0, c, vv);
    if (c[0] >= -kInsideEps && c[1] >= -kInsideEps && c[2] >= -kInsideEps) atomicMin(&a.first[ang], k);
  }
}

// Pass 1, one wave per line of sight: where the line lies in the
// triangulation, from the first sample's triangle (pass 0).  Each segment's
// first sample (its position accumulated sample by sample exactly as the
// march accumulates it) is then located by walking from the previous
// segment's first triangle, a few triangles away, until one lies outside:
// the hull is convex, so the line leaves it once, inside the previous
// segment, whose samples are then located one by one for the exact count.
// Samples exactly on an edge may be given the other triangle of the edge
// than a walk from sample to sample gives: the interpolated values agree to
// the inside tolerance.
__global__ __launch_bounds__(256) void k_ena_locate(Args a) {
  const int ang = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int lane = threadIdx.x & 63;
  if (ang >= a.n_angle) return;  // whole waves leave together
  double lx, ly, x0, y0;
  line_of_sight(a, ang, lx, ly, x0, y0);
  const double ds = kAe / 30.0;
  int t = __builtin_amdgcn_readfirstlane(a.first[ang]);  // uniform: triangle data by scalar loads
  long long n = 0;
  if (t != kNone) {
    const size_t g0 = (size_t)ang * a.n_seg;
    int last = 0;  // the last segment whose first sample is inside, and that sample
    int tl = t;
    double xl = x0, yl = y0;
    if (lane == 0) {
      a.seg_tri[g0] = t;
      a.seg_xy[2 * g0] = x0;
      a.seg_xy[2 * g0 + 1] = y0;
    }
    for (int s = 1; s < a.n_seg; s++) {
      for (int m = 0; m < a.seg_len; m++) {
        x0 = x0 + lx * ds;
        y0 = y0 + ly * ds;
      }
      double c[3];
      int vv[3];
      const int ts = walk(a, tl, x0, y0, c, vv);
      if (ts < 0) break;
      last = s;
      tl = ts;
      xl = x0;
      yl = y0;
      if (lane == 0) {
        a.seg_tri[g0 + s] = tl;
        a.seg_xy[2 * (g0 + s)] = xl;
        a.seg_xy[2 * (g0 + s) + 1] = yl;
      }
    }
    // the samples of the last segment, one by one
    t = tl;
    x0 = xl;
    y0 = yl;
    n = (long long)last * a.seg_len;
    while (true) {
      double c[3];
      int vv[3];
      const int ts = walk(a, t, x0, y0, c, vv);
      if (ts < 0) break;
      if (n >= a.max_samples) {  // more samples than the hull allows
        if (lane == 0) atomicOr(a.err, 2);
        break;
      }
      if (n >= (long long)(last + 1) * a.seg_len) break;  // the next segment's start lay outside
      t = ts;
      n++;
      x0 = x0 + lx * ds;
      y0 = y0 + ly * ds;
    }
  }
  if (lane == 0) a.samples[ang] = (int)n;
}

// Pass 2, one wave per (line of sight, segment, 64 SPL speeds): the march
// of Culc_ENA_the (:844-921) over the segment's samples, from S = 1.  The
// walk, the interpolation and the speed-independent parts of the rates are
// the same in every lane and are formed once per sample for the wave's
// 64 SPL speeds (lane l takes speeds l, l + 64, ...); the rates proper are
// per speed.
template <int SPL>
__global__ __launch_bounds__(256) void k_ena_march(Args a) {
  const long long wave = ((long long)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int lane = threadIdx.x & 63;
  if (wave >= (long long)a.n_angle * a.n_seg * a.chunks) return;  // whole waves leave together
  const int chunk = (int)(wave % a.chunks);
  const long long as = wave / a.chunks;
  const int seg = (int)(as % a.n_seg), ang = (int)(as / a.n_seg);
  const int n = a.samples[ang];
  const int k0 = seg * a.seg_len;
  if (k0 >= n) return;
  const int k1 = min(n, k0 + a.seg_len);
  double lx, ly, x0, y0;
  line_of_sight(a, ang, lx, ly, x0, y0);
  const double ds = kAe / 30.0;
  const size_t g = (size_t)ang * a.n_seg + seg;
  int t = __builtin_amdgcn_readfirstlane(a.seg_tri[g]);
  x0 = a.seg_xy[2 * g];
  y0 = a.seg_xy[2 * g + 1];
  const double vL = 1 * kKms, vR = 1000 * kKms;
  double v[SPL], dsv[SPL], S[SPL], j[SPL];
  bool mine[SPL];
#pragma unroll
  for (int q = 0; q < SPL; q++) {
    const int i = (chunk * SPL + q) * 64 + lane;
    mine[q] = i < a.n_v;
    v[q] = vL + (i + 0.5) * ((vR - vL) / a.n_v);
    dsv[q] = ds / v[q];
    S[q] = 1.0;
    j[q] = 0.0;
  }

#pragma unroll 1
  for (int k = k0; k < k1; k++) {
    double c[3];
    int vv[3];
    t = walk(a, t, x0, y0, c, vv);
    if (t < 0) break;  // not reached: pass 1 found the sample inside
    // the 16 parameters at the sample; those the rates use (:855-871)
    auto at = [&](int p) {
      return c[0] * a.par[16 * vv[0] + p] + c[1] * a.par[16 * vv[1] + p] + c[2] * a.par[16 * vv[2] + p];
    };
    const double rho_p = at(0), u_p = at(1), v_p = at(2), p_p = at(4);
    const Pop P = population(rho_p, u_p, v_p, sqrt(p_p / rho_p));
    const double rho_H3 = at(8) * 3.0;
    const Pop H3 = population(rho_H3, at(9), at(10), sqrt(0.5 * rho_H3 * at(11) / rho_H3));
    const double rho_H4 = at(12) * 3.0;
    const Pop H4 = population(rho_H4, at(13), at(14), sqrt(0.5 * rho_H4 * at(15) / rho_H4));
    const double r = sqrt(x0 * x0 + y0 * y0);
    const double nu_ph = 770.291 * ((kAe / r) * (kAe / r));
    // maxwell_distribution of the plasma (:353-370): its prefactor
    const double fpre = P.rho / (kPi * sqrt(kPi) * P.cp * P.cp * P.cp), icp2 = 1.0 / (P.cp * P.cp);
#pragma unroll
    for (int q = 0; q < SPL; q++) {
      if (!mine[q]) continue;
      const double wx = -v[q] * lx, wy = -v[q] * ly;  // the particle's velocity
      double du = wx - P.u, dv = wy - P.v;
      const double nu_ex = rate(sqrt(du * du + dv * dv + 0.0 * 0.0), P);
      du = wx - H3.u;
      dv = wy - H3.v;
      const double nu_3 = rate(sqrt(du * du + dv * dv + 0.0 * 0.0), H3);
      du = wx - H4.u;
      dv = wy - H4.v;
      const double nu_4 = rate(sqrt(du * du + dv * dv + 0.0 * 0.0), H4);
      const double nu_H = nu_3 + nu_4;
      S[q] = S[q] * exp(-(nu_ex + nu_ph) * dsv[q]);
      const double mx = wx - P.u, my = wy - P.v, mz = 0.0 - 0.0;
      const double v_sq = mx * mx + my * my + mz * mz;
      const double f = fpre * exp(-v_sq * icp2);
      j[q] = j[q] + nu_H * f * v[q] * S[q] * ds;
    }
    x0 = x0 + lx * ds;
    y0 = y0 + ly * ds;
  }
#pragma unroll
  for (int q = 0; q < SPL; q++)
    if (mine[q]) {
      const int i = (chunk * SPL + q) * 64 + lane;
      a.seg_j[g * a.n_v + i] = j[q];
      a.seg_s[g * a.n_v + i] = S[q];
    }
}

// Pass 3, one thread per (line of sight, speed): the survival is a product
// along the line, so j = sum over segments of (survival of the segments
// before it) x (the segment's own j from S = 1).
__global__ __launch_bounds__(256) void k_ena_combine(Args a) {
  const long long id = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= (long long)a.n_angle * a.n_v) return;
  const int ang = (int)(id / a.n_v), i = (int)(id % a.n_v);
  const int ns = (a.samples[ang] + a.seg_len - 1) / a.seg_len;
  double p = 1.0, j = 0.0;
  for (int s = 0; s < ns; s++) {
    const size_t g = ((size_t)ang * a.n_seg + s) * a.n_v + i;
    j = j + p * a.seg_j[g];
    p = p * a.seg_s[g];
  }
  a.jena[id] = j;
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
  int prev_device = 0;
  if (hipGetDevice(&prev_device) != hipSuccess || hipSetDevice(device) != hipSuccess) return SETMHD_ERR_HIP;
  // the caller's current device is restored on every return below
  struct Restore {
    int d;
    ~Restore() { (void)hipSetDevice(d); }
  } restore{prev_device};
  const size_t b_cc = 3 * (size_t)n_cell * 8, b_par = 16 * (size_t)n_cell * 8, b_tri = 3 * (size_t)n_tri * 4;
  const size_t b_ang = (size_t)n_angle * 8, b_out = (size_t)n_angle * n_v * 8, b_smp = (size_t)n_angle * 4;
  ena::Args a{};
  a.n_cell = n_cell;
  a.n_tri = n_tri;
  a.n_angle = n_angle;
  a.n_v = n_v;
  // speeds per lane: the per-sample work is shared by 64 SPL speeds (at
  // 180 angles x 1000 speeds, SPL 8 against 4 measured +7 %, 4 against 1
  // +45 %)
  const int spl = n_v >= 512 ? 8 : n_v >= 256 ? 4 : n_v >= 128 ? 2 : 1;
  a.chunks = (n_v + 64 * spl - 1) / (64 * spl);
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
  // segments of 256 samples, fewer and longer where the per-segment spectra
  // would pass 1 GiB
  a.seg_len = 256;
  while ((double)n_angle * ((a.max_samples + a.seg_len - 1) / a.seg_len) * n_v * 16.0 > 1073741824.0 &&
         a.seg_len < (1 << 24))
    a.seg_len *= 2;
  a.n_seg = (int)((a.max_samples + a.seg_len - 1) / a.seg_len);
  const size_t n_sg = (size_t)n_angle * a.n_seg;
  const size_t b_sgt = n_sg * 4, b_sgxy = n_sg * 16, b_sgv = n_sg * n_v * 8;
  const size_t total = b_cc + b_par + 2 * b_tri + b_ang + b_out + 2 * b_smp + b_sgt + b_sgxy + 2 * b_sgv + 8 + 128;
  char* base = nullptr;
  if (hipMalloc((void**)&base, total) != hipSuccess) return SETMHD_ERR_NOMEM;
  size_t off = 0;
  auto take = [&](size_t n) {
    char* p = base + off;
    off += (n + 7) & ~size_t(7);
    return p;
  };
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
  a.first = (int*)take(b_smp);
  a.seg_tri = (int*)take(b_sgt);
  a.seg_xy = (double*)take(b_sgxy);
  a.seg_j = (double*)take(b_sgv);
  a.seg_s = (double*)take(b_sgv);
  int st = SETMHD_OK;
  hipError_t e = hipMemcpy(d_cc, cell_center, b_cc, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(d_par, cell_par, b_par, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(d_tri, tri, b_tri, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(d_nb, tri_nb, b_tri, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(d_ang, angle, b_ang, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemset(d_err, 0, 8);
  if (e == hipSuccess) e = hipMemsetD32((hipDeviceptr_t)a.first, ena::kNone, (size_t)n_angle);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(ena::k_ena_first, dim3((unsigned)((n_tri + 255) / 256), (unsigned)std::min(n_angle, 65535)),
                       dim3(256), 0, 0, a);
    e = hipGetLastError();
  }
  if (e == hipSuccess) {
    hipLaunchKernelGGL(ena::k_ena_locate, dim3((unsigned)(((long long)n_angle * 64 + 255) / 256)), dim3(256), 0, 0,
                       a);
    e = hipGetLastError();
  }
  if (e == hipSuccess) {
    const long long waves = (long long)n_angle * a.n_seg * a.chunks;
    const dim3 grid((unsigned)((waves * 64 + 255) / 256));
    if (spl == 8) hipLaunchKernelGGL(ena::k_ena_march<8>, grid, dim3(256), 0, 0, a);
    else if (spl == 4) hipLaunchKernelGGL(ena::k_ena_march<4>, grid, dim3(256), 0, 0, a);
    else if (spl == 2) hipLaunchKernelGGL(ena::k_ena_march<2>, grid, dim3(256), 0, 0, a);
    else hipLaunchKernelGGL(ena::k_ena_march<1>, grid, dim3(256), 0, 0, a);
    e = hipGetLastError();
  }
  if (e == hipSuccess) {
    hipLaunchKernelGGL(ena::k_ena_combine, dim3((unsigned)(((long long)n_angle * n_v + 255) / 256)), dim3(256), 0,
                       0, a);
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
