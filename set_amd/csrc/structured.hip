// structured.hip -- fused flux + conservative update on a uniform Cartesian
// box, for gfx950.
//
// Semantics are those of SET's face pass + cell pass (cuf_solvers.cuf:62-794)
// with first-order states, an HLLD face flux and the lagged CFL minimum of
// cuf_kernel.cuf:789-790, restated on a structured grid:
//   * one launch per Euler step or Heun stage: face fluxes are formed once in
//     LDS and applied to both neighbour cells in the same kernel, so the only
//     HBM traffic is the state itself (read once, written once);
//   * 2.5-D blocking: a block owns a TX x TY column of cells and marches up a
//     chunk of z-planes.  Each plane (+1-cell halo ring) is staged in LDS as
//     face-ready data (rho, u, b, total pressure, energy, fast speeds); the
//     x- and y-faces of the plane are read from LDS, the z-face below each
//     cell comes from the previous plane kept in registers;
//   * face work is balanced: every thread forms one x-face and one y-face of
//     the plane and one z-face of its column; the halo cells and the TX + TY
//     faces on the far tile edges are dealt to different waves, rotated
//     between odd and even tiles, so no wave carries two extra passes;
//   * the CFL candidate is the per-axis max wave speed (min of c/s over faces
//     = c/max s, division being monotone), reduced per wave by shuffles and
//     per block through LDS, then one atomic per block;
//   * blocks are remapped so each XCD (blockIdx % 8) gets a contiguous range
//     of tiles: neighbouring tiles share halo lines in that XCD's L2.
// Outflow (copy) boundaries come from clamping neighbour indices.  Cells
// with |centre| < r_hold are held (the wind source), as the reference holds
// |centre| < 0.05 (cuf_solvers.cuf:611).  Precision: R = double, or R = float
// for the face states and Riemann solver with the state, the pressure
// recovery, the flux accumulation and the dt reduction kept in FP64.
#include <type_traits>

#include "hlld_fast.hpp"
#include "kernels.hpp"

namespace setmhd {

// face-data slots in LDS
enum { QR = 0, QU = 1, QB = 4, QPT = 7, QE = 8, QCF = 9, NQ = 11 };

template <class R>
struct CellQ {
  R rho, u[3], b[3], pt, e, cf[3];
};

// conserved FP64 state -> face data.  Pressure is recovered in FP64:
// p = (gamma-1)(E - rho|u|^2/2 - |b|^2/2).
template <class R, bool NEED_CFZ>
__device__ __forceinline__ void derive(const double U[8], CellQ<R>& q) {
  const double ir = frcp(U[0]);
  const double u = U[1] * ir, v = U[2] * ir, w = U[3] * ir;
  const double bb = U[5] * U[5] + U[6] * U[6] + U[7] * U[7];
  const double ke2 = U[1] * u + U[2] * v + U[3] * w;
  const double p = (kGamma - 1.0) * (U[4] - 0.5 * ke2 - 0.5 * bb);
  q.rho = R(U[0]);
  q.u[0] = R(u);
  q.u[1] = R(v);
  q.u[2] = R(w);
  q.b[0] = R(U[5]);
  q.b[1] = R(U[6]);
  q.b[2] = R(U[7]);
  q.pt = R(p + 0.5 * bb);
  q.e = R(U[4]);
  const R irr = R(ir);
  const R c2 = R(kGamma) * R(p) * irr;
  const R s = c2 + R(bb) * irr;
  q.cf[0] = cf_cell<R>(c2, s, q.b[0] * q.b[0] * irr);
  q.cf[1] = cf_cell<R>(c2, s, q.b[1] * q.b[1] * irr);
  if (NEED_CFZ) q.cf[2] = cf_cell<R>(c2, s, q.b[2] * q.b[2] * irr);
}

template <class R>
__device__ __forceinline__ Side<R> side_of(const CellQ<R>& q, int d) {
  const int d1 = d == 2 ? 0 : d + 1, d2 = d == 0 ? 2 : d - 1;
  Side<R> s;
  s.rho = q.rho;
  s.vn = q.u[d];
  s.vt1 = q.u[d1];
  s.vt2 = q.u[d2];
  s.bn = q.b[d];
  s.bt1 = q.b[d1];
  s.bt2 = q.b[d2];
  s.pt = q.pt;
  s.e = q.e;
  s.cf = q.cf[d];
  return s;
}

template <class R, int TX, int TY>
struct Smem {
  R q[NQ][TY + 2][TX + 2];
  R fx[8][TY][TX + 1];
  R fy[8][TY + 1][TX];
};

template <class R, int TX, int TY>
__device__ __forceinline__ void put_q(Smem<R, TX, TY>& sm, int ly, int lx, const CellQ<R>& q) {
  sm.q[QR][ly][lx] = q.rho;
#pragma unroll
  for (int c = 0; c < 3; c++) {
    sm.q[QU + c][ly][lx] = q.u[c];
    sm.q[QB + c][ly][lx] = q.b[c];
  }
  sm.q[QPT][ly][lx] = q.pt;
  sm.q[QE][ly][lx] = q.e;
  sm.q[QCF][ly][lx] = q.cf[0];
  sm.q[QCF + 1][ly][lx] = q.cf[1];
}

// face-frame side of the LDS cell at offset off = ly (TX + 2) + lx of each
// face-data slot, for an x- (D = 0) or y- (D = 1) face
template <int D, class R, int TX, int TY>
__device__ __forceinline__ Side<R> get_side(const Smem<R, TX, TY>& sm, int off) {
  constexpr int D1 = D + 1, D2 = D == 0 ? 2 : 0;
  auto at = [&](int f) { return (&sm.q[f][0][0])[off]; };
  Side<R> s;
  s.rho = at(QR);
  s.vn = at(QU + D);
  s.vt1 = at(QU + D1);
  s.vt2 = at(QU + D2);
  s.bn = at(QB + D);
  s.bt1 = at(QB + D1);
  s.bt2 = at(QB + D2);
  s.pt = at(QPT);
  s.e = at(QE);
  s.cf = at(QCF + D);
  return s;
}

// hlld's side accessor for faces whose states sit in LDS: re-reads the
// chosen state (ds_read) instead of selecting it from the two in registers.
// The offset goes through an empty asm so that the compiler cannot fold the
// reads back into selects of the earlier ones.
template <int D, class R, int TX, int TY>
struct PickLds {
  const Smem<R, TX, TY>& sm;
  int offL, offR;
  __device__ __forceinline__ Side<R> operator()(bool left) const {
    int off = left ? offL : offR;
    asm("" : "+v"(off));
    return get_side<D>(sm, off);
  }
};

// One x- (D = 0) or y- (D = 1) face between LDS cells (ly, lx) and
// (ry, rx); its flux, in x,y,z components, goes to the LDS flux slot of the
// face (the face on the low side of cell (ry, rx)).
template <int D, class R, int TX, int TY>
__device__ __forceinline__ void lds_face(Smem<R, TX, TY>& sm, int ly, int lx, int ry, int rx,
                                         R& smax, bool& ok, R out[8], const CellQ<R>* qr = nullptr) {
  constexpr int D1 = D + 1, D2 = D == 0 ? 2 : 0;
  const int offL = ly * (TX + 2) + lx, offR = ry * (TX + 2) + rx;
  const Side<R> L = get_side<D>(sm, offL);
  // the right cell from registers when the caller holds it (its own cell)
  const Side<R> Rr = qr ? side_of(*qr, D) : get_side<D>(sm, offR);
  R F[8], s;
  // FP64 re-reads the chosen side from LDS (a select costs two instructions
  // per value and keeps both sides live); FP32 selects it in registers
  if constexpr (sizeof(R) == 8) ok &= hlld(L, Rr, F, s, PickLds<D, R, TX, TY>{sm, offL, offR});
  else ok &= hlld(L, Rr, F, s);
  smax = fmax(smax, s);
  auto& dst = [&]() -> auto& {
    if constexpr (D == 0) return sm.fx;
    else return sm.fy;
  }();
  out[0] = F[0];
  out[1 + D] = F[1];
  out[1 + D1] = F[2];
  out[1 + D2] = F[3];
  out[4] = F[4];
  out[5 + D] = F[5];
  out[5 + D1] = F[6];
  out[5 + D2] = F[7];
  const int fy = ry - 1, fx = rx - 1;
#pragma unroll
  for (int v = 0; v < 8; v++) dst[v][fy][fx] = out[v];
}

// The 8 fields of cell idx: uniform (SGPR) field base + 32-bit offset.
__device__ __forceinline__ void load8(const double* __restrict__ base, long long nc, unsigned idx,
                                      double U[8]) {
#pragma unroll
  for (int v = 0; v < 8; v++) U[v] = (base + v * nc)[idx];
}

__device__ __forceinline__ bool held_cell(const SGrid& g, double xc, double yc, double zc) {
  // |centre| < r_hold with the centre and norm rounded as the host forms
  // them (no contraction): same held set as the face-list mesh.
  if (fabs(xc) >= g.r_hold || fabs(yc) >= g.r_hold || fabs(zc) >= g.r_hold) return false;
  double r2 = __dadd_rn(__dadd_rn(__dmul_rn(xc, xc), __dmul_rn(yc, yc)), __dmul_rn(zc, zc));
  return __dsqrt_rn(r2) < g.r_hold;
}

// The updated state is written with nontemporal stores: it is read again
// only by the next step, long after it has left L2.
__device__ __forceinline__ void store_state(double* p, double v) { __builtin_nontemporal_store(v, p); }

// Applies the face sum acc to cell idx, whose state U^n is Uold.  The
// marches update the row or plane below the current one, whose U^n they keep
// in registers from the previous iteration.
template <int MODE>
__device__ __forceinline__ void finalize_with(const FusedArgs& a, double dt, unsigned idx, bool held,
                                              const double acc[8], const double Uold[8],
                                              double& rs_d, double& rs_r) {
  const SGrid& g = a.g;
  if (held) {
    if (MODE != kModeHeun2) {
#pragma unroll
      for (int v = 0; v < 8; v++) store_state(a.Uo + v * g.nc + idx, Uold[v]);
    }
    return;
  }
  const double f = dt * g.inv_vol;
  double Un[8];
  if (MODE == kModeHeun2) {
#pragma unroll
    for (int v = 0; v < 8; v++) {
      double un = (a.Un + v * g.nc)[idx];
      Un[v] = 0.5 * (un + Uold[v] - f * acc[v]);
      if (v == 0) {
        rs_d += fabs(Un[0] - un);
        rs_r += fabs(un);
      }
    }
  } else {
#pragma unroll
    for (int v = 0; v < 8; v++) Un[v] = Uold[v] - f * acc[v];
    rs_d += fabs(Un[0] - Uold[0]);
    rs_r += fabs(Uold[0]);
  }
  if (!(Un[0] > 0.0)) atomicOr(a.err, kErrRho);  // cuf_solvers.cuf:708-711
  // p <= 0 -> 1e-7 (cuf_solvers.cuf:777-779), applied to the total energy.
  // For rho > 0, p > 0 <=> 2 E rho - |m|^2 - rho |b|^2 > 0: no division on
  // the common path.
  const double m2 = Un[1] * Un[1] + Un[2] * Un[2] + Un[3] * Un[3];
  const double b2 = Un[5] * Un[5] + Un[6] * Un[6] + Un[7] * Un[7];
  if (!(2.0 * Un[4] * Un[0] - m2 - Un[0] * b2 > 0.0))
    Un[4] = SETMHD_F4(0.0000001) / (kGamma - 1.0) + 0.5 * m2 * frcp(Un[0]) + 0.5 * b2;
#pragma unroll
  for (int v = 0; v < 8; v++) store_state(a.Uo + v * g.nc + idx, Un[v]);
}

// Raw FP64 state of the own cells of one plane of a tile, [field][row][x],
// filled by LDS DMA (global_load_lds).
template <int TX, int TY>
struct Stage {
  double own[8][TY][TX];
};

typedef const void __attribute__((address_space(1)))* gptr_t;
typedef void __attribute__((address_space(3)))* lptr_t;

// Offset (in doubles, within one plane of one field) of the two cells a
// lane moves for chunk c of the tile's row-major cell list: 128 cells per
// wave instruction, 16 bytes (two x-neighbours) per lane.  x is clamped to
// nx - 2 (nx even: pairs never straddle a row end), y to ny - 1, so every
// address is inside the state and every in-grid cell is staged exactly.
template <int TX>
__device__ __forceinline__ unsigned dma_lane_offset(const SGrid& g, int i0, int j0, int c, int lane) {
  const int e = c * 128 + 2 * lane;
  return (unsigned)min(j0 + e / TX, g.ny - 1) * (unsigned)g.nx + (unsigned)min(i0 + e % TX, g.nx - 2);
}

// Issues the DMA of plane k's own cells.  Instruction q moves chunk
// q % NCH of field q / NCH; wave w issues q = w, w + NW, ...  With NW a
// multiple of NCH every instruction of a wave uses the same chunk, whose
// lane offset is off.  Lane l lands at the instruction's LDS base + 16 l.
template <int TX, int TY>
__device__ __forceinline__ void dma_plane(Stage<TX, TY>& st, const FusedArgs& a, int k, int wid,
                                          unsigned off) {
  constexpr int NW = TX * TY / 64, NCH = TX * TY / 128;
  static_assert(NW % NCH == 0, "one chunk per wave");
  const double* base = a.Uf + (size_t)a.g.nx * a.g.ny * k + off;
#pragma unroll
  for (int q = 0; q < 8 * NCH / NW; q++) {
    const int qq = wid + q * NW, v = qq / NCH, c = qq % NCH;
    __builtin_amdgcn_global_load_lds((gptr_t)(base + v * a.g.nc),
                                     (lptr_t)(&st.own[0][0][0] + v * TX * TY + c * 128), 16, 0, 0);
  }
}

// vmcnt(0) only (expcnt and lgkmcnt left at their maxima)
__device__ __forceinline__ void wait_vm() { __builtin_amdgcn_s_waitcnt(0x0F70); }

// Block-level results of one tile, valid in thread 0: the tile's CFL
// candidate for the next step and its density-residual partial sums.
struct TileRed {
  double dt, rd, rr;
};

// One 3-D tile (TX x TY columns, the planes of one z-chunk) of one grid:
// stage the planes, form the faces, update the cells, and fold the tile's
// CFL candidate into the grid's next dt slot.  book: this tile also does the
// step's bookkeeping (t += dt, re-arm the dt slot after next).
template <class R, int TX, int TY, int DIM, int MODE, bool DMA = false>
__device__ __forceinline__ TileRed fused_tile(const FusedArgs& a, int b, bool book,
                                              Smem<R, TX, TY>& sm, double* red,
                                              Stage<TX, TY>* st = nullptr) {
  static_assert(DIM == 3, "2-D grids march rows (row_tile, wrow_march)");
  constexpr int NH = 2 * TX + 2 * TY;
  constexpr int NW = TX * TY / 64;  // waves per block
  static_assert(TX <= 64 && TY <= 64 && NW >= 2 && NH <= (NW - 1) * 64 + 64, "extra-work layout");
  const SGrid& g = a.g;
  const int tid = threadIdx.x, tx = tid % TX, ty = tid / TX;
  // Work beyond one cell per lane -- the halo ring's cells and the faces on
  // the tile's far x and far y edges -- is spread over the waves by role:
  // roles 0.. take 64 halo cells each, role NW-1 the far-x faces, role NW-2
  // the far-y faces.  Odd tiles rotate the roles by NW/2 so that the two
  // blocks sharing a CU load its SIMDs evenly.
  const int lane = tid & 63;
  const int wid = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int role = (wid + ((b & 1) ? NW / 2 : 0)) % NW;
  const int hid = role * 64 + lane;  // halo ring index of this lane

  const int ntx = (g.nx + TX - 1) / TX, nty = (g.ny + TY - 1) / TY;
  const int tile_x = b % ntx, rest = b / ntx, tile_y = rest % nty, kc = rest / nty;
  const int i0 = tile_x * TX, j0 = tile_y * TY;
  const int i = i0 + tx, j = j0 + ty;
  const bool own = i < g.nx && j < g.ny;
  const int ic = min(i, g.nx - 1), jc = min(j, g.ny - 1);
  const int k0 = kc * a.kchunk;
  const int k1 = min(k0 + a.kchunk, g.nz);
  const unsigned plane = (unsigned)g.nx * (unsigned)g.ny;
  const unsigned col0 = (unsigned)ic + (unsigned)g.nx * (unsigned)jc;
  const double dt = a.dt.slot[a.cur];
  if (MODE != kModeHeun2 && book && tid == 0) {
    a.dt.t[0] = a.dt.t[0] + dt;                 // cuf_kernel.cuf:792-794
    a.dt.slot[(a.cur + 2) % 3] = kDtReset;      // :790, for step n+2
  }
  const double xc = __dadd_rn(g.x0, __dmul_rn((double)ic + 0.5, g.dx));
  const double yc = __dadd_rn(g.y0, __dmul_rn((double)jc + 0.5, g.dy));

  // halo cell of this thread (ring around the tile, corners excluded)
  int hly = 0, hlx = 0, hgi = 0, hgj = 0;
  const bool has_h = hid < NH;
  if (has_h) {
    if (hid < TX) { hlx = hid + 1; hly = 0; hgi = i0 + hid; hgj = j0 - 1; }
    else if (hid < 2 * TX) { hlx = hid - TX + 1; hly = TY + 1; hgi = i0 + hid - TX; hgj = j0 + TY; }
    else if (hid < 2 * TX + TY) { hlx = 0; hly = hid - 2 * TX + 1; hgi = i0 - 1; hgj = j0 + hid - 2 * TX; }
    else { hlx = TX + 1; hly = hid - 2 * TX - TY + 1; hgi = i0 + TX; hgj = j0 + hid - 2 * TX - TY; }
    hgi = min(max(hgi, 0), g.nx - 1);
    hgj = min(max(hgj, 0), g.ny - 1);
  }

  R smax[3] = {R(0), R(0), R(0)};  // wave speeds in R: max commutes with the exact widening
  double rs_d = 0.0, rs_r = 0.0;
  double accp[8];
  Side<R> zp;
  bool ok = true;

  // Software pipeline (FP64): the next plane's own and halo cells are loaded
  // while the current plane's faces are formed.  The mixed form skips it:
  // its 32 extra VGPRs would cost a wave per SIMD.
  // With DMA (3-D, even nx) the next plane goes global -> LDS directly
  // while the faces are formed, using no registers.
  constexpr bool PF = sizeof(R) == 8 && !DMA;
  const unsigned hcol = (unsigned)hgi + (unsigned)g.nx * (unsigned)hgj;
  // staged positions of this lane's own and halo cells (clamped as above)
  const int sox = ic - i0, soy = jc - j0;
  unsigned dma_off = 0;
  if constexpr (DMA) {
    dma_off = dma_lane_offset<TX>(g, i0, j0, wid % (TX * TY / 128), lane);
    dma_plane<TX, TY>(*st, a, k0, wid, dma_off);
    wait_vm();
    __syncthreads();
  }
  // Ukeep: U^n of this column's previous plane, kept in registers (a
  // re-read would make its wait drain the next plane's prefetch as well)
  double Unx[8], Hnx[8], Ukeep[8] = {};
  if constexpr (PF) load8(a.Uf, g.nc, col0 + plane * (unsigned)max(k0 - 1, 0), Unx);

  // Planes k0 - 1 (the plane below the chunk: its z-side only) through k1
  // (the closing z-face above the chunk's last plane, from plane
  // min(k1, nz - 1)) take the same code as the owned planes, so the result
  // does not depend on the chunking (the loop is not peeled: a peeled copy
  // of the flux code may contract multiply-adds differently).
#pragma clang loop unroll(disable)
  for (int k = k0 - 1; k <= k1; ++k) {
    const bool below = k < k0, top = k == k1;
    const unsigned kk = (unsigned)min(max(k, 0), g.nz - 1);  // the plane read
    double Uc[8], Uh[8];
    if constexpr (PF) {
#pragma unroll
      for (int v = 0; v < 8; v++) {
        Uc[v] = Unx[v];
        Uh[v] = Hnx[v];
      }
      if (!top) {  // plane k+1: own cells, and the halo unless it is the closing plane
        const int kn = min(k + 1, g.nz - 1);
        load8(a.Uf, g.nc, col0 + plane * kn, Unx);
        if (has_h && k + 1 < k1) load8(a.Uf, g.nc, hcol + plane * kn, Hnx);
      }
    } else if constexpr (DMA) {
      if (below) {
        load8(a.Uf, g.nc, col0 + plane * kk, Uc);
      } else {
        if (has_h && !top) load8(a.Uf, g.nc, hcol + plane * kk, Uh);
#pragma unroll
        for (int v = 0; v < 8; v++) Uc[v] = st->own[v][soy][sox];
      }
    } else {
      load8(a.Uf, g.nc, col0 + plane * kk, Uc);
      if (has_h && !below && !top) load8(a.Uf, g.nc, hcol + plane * kk, Uh);
    }
    double Uprev[8];  // U^n of the cell below, updated once the z-face between them is known
#pragma unroll
    for (int v = 0; v < 8; v++) {
      Uprev[v] = Ukeep[v];
      Ukeep[v] = Uc[v];
    }
    CellQ<R> qc;
    derive<R, true>(Uc, qc);
    const Side<R> zc = side_of(qc, 2);
    if (!below) {
      // Stage the plane (the previous plane's faces were all read before
      // its flux barrier).
      if (!top) {
        put_q(sm, ty + 1, tx + 1, qc);
        if (has_h) {
          CellQ<R> qh;
          derive<R, false>(Uh, qh);
          put_q(sm, hly, hlx, qh);
        }
      }
      // The z-face below this cell and the update of the cell below, whose
      // faces are then all known, need registers only: they run before the
      // staging barrier.
      R Fz[8];
      {
        R s;
        ok &= hlld(zp, zc, Fz, s);
        smax[2] = fmax(smax[2], s);
      }
      // z-face: rotate frame (z, x, y) back to x, y, z
      R fz[8];
      fz[0] = Fz[0];
      fz[3] = Fz[1];
      fz[1] = Fz[2];
      fz[2] = Fz[3];
      fz[4] = Fz[4];
      fz[7] = Fz[5];
      fz[5] = Fz[6];
      fz[6] = Fz[7];
      if (k > k0 && own) {
#pragma unroll
        for (int v = 0; v < 8; v++) accp[v] = accp[v] + (double)fz[v] * g.az;
        const double zcen = __dadd_rn(g.z0, __dmul_rn((double)(k - 1) + 0.5, g.dz));
        finalize_with<MODE>(a, dt, col0 + plane * (unsigned)(k - 1), held_cell(g, xc, yc, zcen), accp, Uprev,
                            rs_d, rs_r);
      }
      if (!top) {
        __syncthreads();
        // every wave has read the staged plane: stage plane k+1 (the closing
        // plane when k+1 == k1)
        if constexpr (DMA) dma_plane<TX, TY>(*st, a, min(k + 1, g.nz - 1), wid, dma_off);

        // x- and y-faces of the plane.  Each thread forms the faces on the
        // low side of its cell and keeps their fluxes (fxo, fyo); the TY faces
        // on the tile's far x edge go to lanes 0..TY-1 of the role NW-1 wave,
        // the TX faces on the far y edge to lanes 0..TX-1 of the role NW-2
        // wave, each as a second iteration of its loop.
        R fxo[8], fyo[8], fe[8];
        {
          lds_face<0>(sm, ty + 1, tx, ty + 1, tx + 1, smax[0], ok, fxo, &qc);
          if (role == NW - 1 && lane < TY) lds_face<0>(sm, lane + 1, TX, lane + 1, TX + 1, smax[0], ok, fe);
          lds_face<1>(sm, ty, tx + 1, ty + 1, tx + 1, smax[1], ok, fyo, &qc);
          if (role == NW - 2 && lane < TX) lds_face<1>(sm, TY, lane + 1, TY + 1, lane + 1, smax[1], ok, fe);
        }
        // DMA: this wave's share of plane k+1 has landed before the barrier
        if constexpr (DMA) wait_vm();
        __syncthreads();
        // signed face sum in the face-slot order (x-, x+, y-, y+, z-, z+);
        // the x+ and y+ fluxes are the neighbours'
#pragma unroll
        for (int v = 0; v < 8; v++) {
          double s = -((double)fxo[v] * g.ax);
          s = s + (double)sm.fx[v][ty][tx + 1] * g.ax;
          s = s - (double)fyo[v] * g.ay;
          s = s + (double)sm.fy[v][ty + 1][tx] * g.ay;
          accp[v] = s - (double)fz[v] * g.az;
        }
      }
    }
    zp = zc;
  }
  if (!ok) atomicOr(a.err, kErrRegion);

  // Reductions over the block: CFL minimum for the next step and, when
  // asked for, the density-residual sums (one LDS round, no atomics here).
  const bool want_rs = a.resid != nullptr && MODE != kModeHeun1;
  double cand = kDtReset;
  if (MODE != kModeHeun2) {
    cand = g.cfl * g.distx / (double)smax[0];
    cand = fmin(cand, g.cfl * g.disty / (double)smax[1]);
    cand = fmin(cand, g.cfl * g.distz / (double)smax[2]);
    cand = fmin(cand, g.dt_cap);
    cand = wave_min(cand);
  }
  if (want_rs) {
    rs_d = wave_sum(rs_d);
    rs_r = wave_sum(rs_r);
  }
  if (lane == 0) {
    red[wid] = cand;
    red[NW + wid] = rs_d;
    red[2 * NW + wid] = rs_r;
  }
  __syncthreads();
  TileRed out{kDtReset, 0.0, 0.0};
  if (tid == 0) {
    out = TileRed{red[0], red[NW], red[2 * NW]};
    for (int w = 1; w < NW; w++) {
      out.dt = fmin(out.dt, red[w]);
      out.rd += red[NW + w];
      out.rr += red[2 * NW + w];
    }
  }
  return out;
}

// ---- 2-D: row march --------------------------------------------------------
// A 2-D block is a segment of RW = 256 columns (lane l holds column
// x0 + l - 1: lanes 1..RW-2 own their column, lanes 0 and RW-1 are the
// halo) marching up a chunk of rows.  Per row every lane derives its cell
// once; lane l forms the x-face between itself and lane l+1 (its own data
// from registers, the neighbour's from LDS) and the y-face below its cell
// from the previous row held in registers.  No halo pass and no extra edge
// pass; the next row is prefetched into registers.  Overlapping segments
// (254 owned columns each) re-derive one column on each side.
enum { XR = 0, XU = 1, XB = 4, XPT = 7, XE = 8, XCF = 9, NQX = 10 };

template <class R>
struct RowSmem {
  R q[NQX][RW];   // x-face data of the row
  R fx[8][RW];    // flux of the x-face on the right of lane l (x, y, z components)
};

template <class R>
__device__ __forceinline__ Side<R> row_side(const RowSmem<R>& sm, int l) {
  Side<R> s;
  s.rho = sm.q[XR][l];
  s.vn = sm.q[XU][l];
  s.vt1 = sm.q[XU + 1][l];
  s.vt2 = sm.q[XU + 2][l];
  s.bn = sm.q[XB][l];
  s.bt1 = sm.q[XB + 1][l];
  s.bt2 = sm.q[XB + 2][l];
  s.pt = sm.q[XPT][l];
  s.e = sm.q[XE][l];
  s.cf = sm.q[XCF][l];
  return s;
}

// y-face flux in the face frame (n = y, t1 = z, t2 = x) -> x, y, z components
template <class R>
__device__ __forceinline__ void from_y_frame(const R F[8], double f[8]) {
  f[0] = F[0];
  f[2] = F[1];
  f[3] = F[2];
  f[1] = F[3];
  f[4] = F[4];
  f[6] = F[5];
  f[7] = F[6];
  f[5] = F[7];
}

template <class R, int MODE>
__device__ __forceinline__ TileRed row_tile(const FusedArgs& a, int b, bool book, RowSmem<R>& sm,
                                            double* red) {
  constexpr int NW = RW / 64;
  const SGrid& g = a.g;
  const int l = threadIdx.x, lane = l & 63;
  const int wid = __builtin_amdgcn_readfirstlane(l >> 6);
  const int ntx = (g.nx + ROWN - 1) / ROWN;
  const int seg = b % ntx, rc = b / ntx;
  const int c = seg * ROWN + l - 1;
  const int cc = min(max(c, 0), g.nx - 1);
  const bool own = l >= 1 && l <= ROWN && c < g.nx;
  const int j0 = rc * a.kchunk, j1 = min(j0 + a.kchunk, g.ny);
  const unsigned row = (unsigned)g.nx;
  const double dt = a.dt.slot[a.cur];
  if (MODE != kModeHeun2 && book && l == 0) {
    a.dt.t[0] = a.dt.t[0] + dt;                 // cuf_kernel.cuf:792-794
    a.dt.slot[(a.cur + 2) % 3] = kDtReset;      // :790, for step n+2
  }
  const double xc = __dadd_rn(g.x0, __dmul_rn((double)cc + 0.5, g.dx));
  R smax[2] = {R(0), R(0)};
  double rs_d = 0.0, rs_r = 0.0;
  double accp[8];
  bool ok = true;
  // Rows j0 - 1 (the row below the chunk: its y-side only) through j1 (the
  // closing y-face above the chunk's last row, from row min(j1, ny - 1))
  // take the same code as the owned rows, so the result does not depend on
  // the chunking (the loop is not peeled: a peeled copy of the flux code
  // may contract multiply-adds differently).
  Side<R> yp;
  double Unx[8], Uprev[8];  // Uprev: U^n of the cell below, updated once the y-face between them is known
  load8(a.Uf, g.nc, (unsigned)cc + row * (unsigned)max(j0 - 1, 0), Unx);
#pragma clang loop unroll(disable)
  for (int j = j0 - 1; j <= j1; ++j) {
    const bool below = j < j0, top = j == j1;
    const unsigned idx = (unsigned)cc + row * (unsigned)j;
    double Uc[8];
#pragma unroll
    for (int v = 0; v < 8; v++) Uc[v] = Unx[v];
    if (!top) load8(a.Uf, g.nc, (unsigned)cc + row * (unsigned)min(j + 1, g.ny - 1), Unx);  // next row
    CellQ<R> qc;
    derive<R, false>(Uc, qc);
    const Side<R> yc = side_of(qc, 1);
    if (!below) {
      if (!top) {
        sm.q[XR][l] = qc.rho;
#pragma unroll
        for (int k = 0; k < 3; k++) {
          sm.q[XU + k][l] = qc.u[k];
          sm.q[XB + k][l] = qc.b[k];
        }
        sm.q[XPT][l] = qc.pt;
        sm.q[XE][l] = qc.e;
        sm.q[XCF][l] = qc.cf[0];
      }
      if (!top) {
        __syncthreads();
        if (l < RW - 1) {  // x-face on the right of this lane's cell
          R F[8], s;
          ok &= hlld(side_of(qc, 0), row_side(sm, l + 1), F, s);
          smax[0] = fmax(smax[0], s);
#pragma unroll
          for (int v = 0; v < 8; v++) sm.fx[v][l] = F[v];
        }
      }
      R Fy[8];
      {  // y-face below this lane's cell
        R s;
        ok &= hlld(yp, yc, Fy, s);
        smax[1] = fmax(smax[1], s);
      }
      double fy[8];
      from_y_frame(Fy, fy);
      if (j > j0 && own) {
#pragma unroll
        for (int v = 0; v < 8; v++) accp[v] = accp[v] + fy[v] * g.ay;
        const double ycen = __dadd_rn(g.y0, __dmul_rn((double)(j - 1) + 0.5, g.dy));
        finalize_with<MODE>(a, dt, idx - row, held_cell(g, xc, ycen, 0.0), accp, Uprev, rs_d, rs_r);
      }
      if (!top) {
        __syncthreads();
        // signed face sum in the face-slot order (x-, x+, y-, y+)
        const int lm = max(l - 1, 0);
#pragma unroll
        for (int v = 0; v < 8; v++) {
          double s = -((double)sm.fx[v][lm] * g.ax);
          s = s + (double)sm.fx[v][l] * g.ax;
          accp[v] = s - fy[v] * g.ay;
        }
      }
    }
    yp = yc;
#pragma unroll
    for (int v = 0; v < 8; v++) Uprev[v] = Uc[v];
  }
  if (!ok) atomicOr(a.err, kErrRegion);

  const bool want_rs = a.resid != nullptr && MODE != kModeHeun1;
  double cand = kDtReset;
  if (MODE != kModeHeun2) {
    cand = g.cfl * g.distx / (double)smax[0];
    cand = fmin(cand, g.cfl * g.disty / (double)smax[1]);
    cand = fmin(cand, g.dt_cap);
    cand = wave_min(cand);
  }
  if (want_rs) {
    rs_d = wave_sum(rs_d);
    rs_r = wave_sum(rs_r);
  }
  if (lane == 0) {
    red[wid] = cand;
    red[NW + wid] = rs_d;
    red[2 * NW + wid] = rs_r;
  }
  __syncthreads();
  TileRed out{kDtReset, 0.0, 0.0};
  if (l == 0) {
    out = TileRed{red[0], red[NW], red[2 * NW]};
    for (int w = 1; w < NW; w++) {
      out.dt = fmin(out.dt, red[w]);
      out.rd += red[NW + w];
      out.rr += red[2 * NW + w];
    }
  }
  return out;
}

// One 2-D grid: one block per segment x row chunk (XCD-contiguous ranges).
template <class R, int MODE>
__global__ __launch_bounds__(RW) void k_rows(FusedArgs a) {
  __shared__ RowSmem<R> sm;
  __shared__ double red[3 * RW / 64];
  int b = blockIdx.x;
  const int nblk = gridDim.x;
  if ((nblk & 7) == 0) b = (b & 7) * (nblk >> 3) + (b >> 3);
  const TileRed r = row_tile<R, MODE>(a, b, blockIdx.x == 0, sm, red);
  if (threadIdx.x == 0) {
    if (MODE != kModeHeun2) atomic_min_pos(&a.dt.slot[(a.cur + 1) % 3], r.dt);
    if (a.resid != nullptr && MODE != kModeHeun1) {
      atomicAdd(&a.resid[0], r.rd);
      atomicAdd(&a.resid[1], r.rr);
    }
  }
}

// Wave row march: every wave marches its own WRN = 62-column segment (lanes
// 1..62 own, 0 and 63 are the halo) up its own chunk of rows, with the same
// arithmetic in the same order as the block row march.  The neighbour's
// x-face data and the left x-face's flux cross lanes by shuffles, so the
// march has no barrier and no LDS, and narrow grids waste fewer lanes.

template <class R>
__device__ __forceinline__ Side<R> shfl_down1(Side<R> s) {
  s.rho = __shfl_down(s.rho, 1, 64);
  s.vn = __shfl_down(s.vn, 1, 64);
  s.vt1 = __shfl_down(s.vt1, 1, 64);
  s.vt2 = __shfl_down(s.vt2, 1, 64);
  s.bn = __shfl_down(s.bn, 1, 64);
  s.bt1 = __shfl_down(s.bt1, 1, 64);
  s.bt2 = __shfl_down(s.bt2, 1, 64);
  s.pt = __shfl_down(s.pt, 1, 64);
  s.e = __shfl_down(s.e, 1, 64);
  s.cf = __shfl_down(s.cf, 1, 64);
  return s;
}

// One wave tile t (segment t % nseg, row chunk t / nseg) of one grid; an
// inactive wave marches nothing.  Returns the tile's CFL candidate and
// density-residual sums, reduced over the wave (valid in every lane).
template <class R, int MODE>
__device__ __forceinline__ TileRed wrow_march(const FusedArgs& a, int t, bool active, bool book) {
  const SGrid& g = a.g;
  const int l = threadIdx.x, lane = l & 63;
  const int ntx = (g.nx + WRN - 1) / WRN;
  const int seg = t % ntx, rc = t / ntx;
  const int c = seg * WRN + lane - 1;
  const int cc = min(max(c, 0), g.nx - 1);
  const bool own = lane >= 1 && lane <= WRN && c < g.nx;
  const int j0 = rc * a.kchunk, j1 = active ? min(j0 + a.kchunk, g.ny) : j0 - 2;
  const unsigned row = (unsigned)g.nx;
  const double dt = a.dt.slot[a.cur];
  if (MODE != kModeHeun2 && book && l == 0) {
    a.dt.t[0] = a.dt.t[0] + dt;                 // cuf_kernel.cuf:792-794
    a.dt.slot[(a.cur + 2) % 3] = kDtReset;      // :790, for step n+2
  }
  const double xc = __dadd_rn(g.x0, __dmul_rn((double)cc + 0.5, g.dx));
  R smax[2] = {R(0), R(0)};
  double rs_d = 0.0, rs_r = 0.0;
  double accp[8];
  bool ok = true;
  Side<R> yp;
  double Unx[8], Uprev[8];
  if (active) load8(a.Uf, g.nc, (unsigned)cc + row * (unsigned)max(j0 - 1, 0), Unx);
#pragma clang loop unroll(disable)
  for (int j = j0 - 1; j <= j1; ++j) {
    const bool below = j < j0, top = j == j1;
    const unsigned idx = (unsigned)cc + row * (unsigned)j;
    double Uc[8];
#pragma unroll
    for (int v = 0; v < 8; v++) Uc[v] = Unx[v];
    if (!top) load8(a.Uf, g.nc, (unsigned)cc + row * (unsigned)min(j + 1, g.ny - 1), Unx);  // next row
    CellQ<R> qc;
    derive<R, false>(Uc, qc);
    const Side<R> yc = side_of(qc, 1);
    if (!below) {
      R Fx[8];
      if (!top) {  // x-face on the right of this lane's cell (lane 63's is not used)
        const Side<R> xs = side_of(qc, 0);
        const Side<R> xr = shfl_down1(xs);
        R s;
        const bool okx = hlld(xs, xr, Fx, s);
        if (lane < 63) {
          ok &= okx;
          smax[0] = fmax(smax[0], s);
        }
      }
      R Fy[8];
      {  // y-face below this lane's cell
        R s;
        ok &= hlld(yp, yc, Fy, s);
        smax[1] = fmax(smax[1], s);
      }
      double fy[8];
      from_y_frame(Fy, fy);
      if (j > j0 && own) {
#pragma unroll
        for (int v = 0; v < 8; v++) accp[v] = accp[v] + fy[v] * g.ay;
        const double ycen = __dadd_rn(g.y0, __dmul_rn((double)(j - 1) + 0.5, g.dy));
        finalize_with<MODE>(a, dt, idx - row, held_cell(g, xc, ycen, 0.0), accp, Uprev, rs_d, rs_r);
      }
      if (!top) {
        // signed face sum in the face-slot order (x-, x+, y-, y+); the x-face
        // on the left comes from lane - 1 (the x frame is the global order)
#pragma unroll
        for (int v = 0; v < 8; v++) {
          const R left = __shfl_up(Fx[v], 1, 64);
          double s = -((double)left * g.ax);
          s = s + (double)Fx[v] * g.ax;
          accp[v] = s - fy[v] * g.ay;
        }
      }
    }
    yp = yc;
#pragma unroll
    for (int v = 0; v < 8; v++) Uprev[v] = Uc[v];
  }
  if (!ok) atomicOr(a.err, kErrRegion);

  const bool want_rs = a.resid != nullptr && MODE != kModeHeun1;
  double cand = kDtReset;
  if (MODE != kModeHeun2 && active) {
    cand = g.cfl * g.distx / (double)smax[0];
    cand = fmin(cand, g.cfl * g.disty / (double)smax[1]);
    cand = fmin(cand, g.dt_cap);
  }
  cand = wave_min(cand);
  if (want_rs) {
    rs_d = wave_sum(rs_d);
    rs_r = wave_sum(rs_r);
  }
  return TileRed{cand, rs_d, rs_r};
}

// One 2-D grid, wave row march: block b runs wave tiles WPB b .. WPB b + WPB - 1
// (XCD-contiguous block ranges).
template <class R, int MODE>
__global__ __launch_bounds__(RW) void k_wrows(FusedArgs a) {
  __shared__ double red[3 * WPB];
  int b = blockIdx.x;
  const int nblk = gridDim.x;
  if ((nblk & 7) == 0) b = (b & 7) * (nblk >> 3) + (b >> 3);
  const int wid = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int t = b * WPB + wid;
  const bool active = t < wave_row_tiles(a.g, a.kchunk);
  const TileRed w = wrow_march<R, MODE>(a, active ? t : 0, active, blockIdx.x == 0);
  if ((threadIdx.x & 63) == 0) {
    red[wid] = w.dt;
    red[WPB + wid] = w.rd;
    red[2 * WPB + wid] = w.rr;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    TileRed r{red[0], red[WPB], red[2 * WPB]};
    for (int v = 1; v < WPB; v++) {
      r.dt = fmin(r.dt, red[v]);
      r.rd += red[WPB + v];
      r.rr += red[2 * WPB + v];
    }
    if (MODE != kModeHeun2) atomic_min_pos(&a.dt.slot[(a.cur + 1) % 3], r.dt);
    if (a.resid != nullptr && MODE != kModeHeun1) {
      atomicAdd(&a.resid[0], r.rd);
      atomicAdd(&a.resid[1], r.rr);
    }
  }
}

// One grid: one block per tile, tiles remapped so each XCD (blockIdx % 8)
// gets a contiguous range (neighbouring tiles share halo lines in its L2).
template <class R, int TX, int TY, int DIM, int MODE, bool DMA>
__global__ __launch_bounds__(TX* TY) __attribute__((amdgpu_waves_per_eu(sizeof(R) == 4 ? 3 : 2))) void k_fused(FusedArgs a) {
  __shared__ Smem<R, TX, TY> sm;
  __shared__ double red[3 * TX * TY / 64];
  int b = blockIdx.x;
  const int nblk = gridDim.x;
  if ((nblk & 7) == 0) b = (b & 7) * (nblk >> 3) + (b >> 3);
  TileRed r;
  if constexpr (DMA) {
    __shared__ Stage<TX, TY> st;
    r = fused_tile<R, TX, TY, DIM, MODE, true>(a, b, blockIdx.x == 0, sm, red, &st);
  } else {
    r = fused_tile<R, TX, TY, DIM, MODE>(a, b, blockIdx.x == 0, sm, red);
  }
  if (threadIdx.x == 0) {
    if (MODE != kModeHeun2) atomic_min_pos(&a.dt.slot[(a.cur + 1) % 3], r.dt);
    if (a.resid != nullptr && MODE != kModeHeun1) {
      atomicAdd(&a.resid[0], r.rd);
      atomicAdd(&a.resid[1], r.rr);
    }
  }
}

// The launch arguments of grid gi of a batch at its current step: its step
// index on the device selects the dt slot and the ping-pong buffer.
__device__ __forceinline__ FusedArgs batch_args(const BatchGrid* __restrict__ grids, int gi) {
  const BatchGrid& bg = grids[gi];
  const int n = *bg.nstep;
  FusedArgs a;
  a.g = bg.g;
  a.Uf = bg.buf[n & 1];
  a.Un = a.Uf;
  a.Uo = bg.buf[(n & 1) ^ 1];
  a.dt = bg.dt;
  a.cur = n % 3;
  a.kchunk = bg.kchunk;
  a.resid = bg.resid;
  a.err = bg.err;
  return a;
}

// Publishes one grid's folded CFL candidate and residual sums.
__device__ __forceinline__ void batch_publish(const BatchGrid& g, const TileRed& r) {
  atomic_min_pos(&g.dt.slot[(*g.nstep + 1) % 3], r.dt);
  atomicAdd(&g.resid[0], r.rd);
  atomicAdd(&g.resid[1], r.rr);
}

// 2-D grids of a batch: wave tiles.  Each wave of the block marches global
// tile t0 + wave (when below t1) of whatever grid it belongs to; the waves'
// results go through LDS and thread 0 folds them, in tile order, into acc
// for grid *cur, publishing acc whenever the grid changes.
template <class R>
__device__ __forceinline__ void wave_tiles_step(const BatchGrid* __restrict__ grids,
                                                const int* __restrict__ tile_grid,
                                                const int* __restrict__ tile_first, int t0, int t1,
                                                double* red, int* gid, TileRed& acc, int& cur) {
  const int wid = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int t = t0 + wid;
  const bool active = t < t1;
  const int gi = active ? tile_grid[t] : -1;
  TileRed w{kDtReset, 0.0, 0.0};
  if (active) {
    const FusedArgs a = batch_args(grids, gi);
    w = wrow_march<R, kModeEuler>(a, t - tile_first[gi], true, false);
  }
  if ((threadIdx.x & 63) == 0) {
    gid[wid] = gi;
    red[wid] = w.dt;
    red[WPB + wid] = w.rd;
    red[2 * WPB + wid] = w.rr;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int v = 0; v < WPB && gid[v] >= 0; v++) {
      if (gid[v] != cur) {
        if (cur >= 0) batch_publish(grids[cur], acc);
        cur = gid[v];
        acc = TileRed{kDtReset, 0.0, 0.0};
      }
      acc.dt = fmin(acc.dt, red[v]);
      acc.rd += red[WPB + v];
      acc.rr += red[2 * WPB + v];
    }
  }
  __syncthreads();  // red and gid are reused
}

// Many grids, one persistent launch per step.  Block b walks a contiguous
// range of the global tile list, cut so that every range carries about the
// same number of planes or rows (ranges in block order: dealing them out per
// XCD measured 25 % slower for this persistent launch).  Each grid's step
// index lives on the device (BatchGrid::nstep): it selects the dt slot and
// the ping-pong buffer, so the launch has no per-step arguments and many
// steps can be captured in one graph.  k_batch_book, launched after it on
// the same stream, closes the step of every grid.  3-D grids: z-march tiles.
template <class R, int TX, int TY, int DIM>
__global__ __launch_bounds__(TX* TY) __attribute__((amdgpu_waves_per_eu(sizeof(R) == 4 ? 3 : 2))) void k_fused_batch(
    const BatchGrid* __restrict__ grids, const int* __restrict__ tile_grid,
    const int* __restrict__ tile_first, const int* __restrict__ blk_first) {
  static_assert(DIM == 3, "2-D grids of a batch march wave rows (k_wrow_batch)");
  __shared__ Smem<R, TX, TY> sm;
  __shared__ double red[3 * TX * TY / 64];
  // tables on the device: the batch can change between graph replays
  const int t0 = blk_first[blockIdx.x], t1 = blk_first[blockIdx.x + 1];
  // thread 0 folds the block's results over consecutive tiles of one grid
  // and publishes them (one atomic each) when the grid changes
  int cur = -1;
  TileRed acc{kDtReset, 0.0, 0.0};
  for (int t = t0; t < t1; t++) {
    const int gi = tile_grid[t];
    if (threadIdx.x == 0 && gi != cur && cur >= 0) {
      batch_publish(grids[cur], acc);
      acc = TileRed{kDtReset, 0.0, 0.0};
    }
    cur = gi;
    const FusedArgs a = batch_args(grids, gi);
    const TileRed r = fused_tile<R, TX, TY, DIM, kModeEuler>(a, t - tile_first[gi], false, sm, red);
    if (threadIdx.x == 0) {
      acc.dt = fmin(acc.dt, r.dt);
      acc.rd += r.rd;
      acc.rr += r.rr;
    }
    __syncthreads();  // LDS is reused by the next tile
  }
  if (threadIdx.x == 0 && cur >= 0) batch_publish(grids[cur], acc);
}

// Many grids, one block per tile (the other batch launch shape): the
// hardware balances tiles of unequal cost as blocks finish, as for a single
// grid, with the same XCD-contiguous tile ranges; the launch size is the
// tile count, so the captured graph is re-made when that changes.  3-D grids.
template <class R, int TX, int TY, int DIM>
__global__ __launch_bounds__(TX* TY) __attribute__((amdgpu_waves_per_eu(sizeof(R) == 4 ? 3 : 2))) void k_fused_tiles(
    const BatchGrid* __restrict__ grids, const int* __restrict__ tile_grid,
    const int* __restrict__ tile_first) {
  static_assert(DIM == 3, "2-D grids of a batch march wave rows (k_wrow_tiles)");
  __shared__ Smem<R, TX, TY> sm;
  __shared__ double red[3 * TX * TY / 64];
  int t = blockIdx.x;
  const int nblk = gridDim.x;
  if ((nblk & 7) == 0) t = (t & 7) * (nblk >> 3) + (t >> 3);
  const int gi = tile_grid[t];
  const FusedArgs a = batch_args(grids, gi);
  const TileRed r = fused_tile<R, TX, TY, DIM, kModeEuler>(a, t - tile_first[gi], false, sm, red);
  if (threadIdx.x == 0) {
    atomic_min_pos(&a.dt.slot[(a.cur + 1) % 3], r.dt);
    atomicAdd(&a.resid[0], r.rd);
    atomicAdd(&a.resid[1], r.rr);
  }
}

// The 2-D grids of a batch in block rows (the mixed form), one block per
// tile, XCD-contiguous tile ranges.
template <class R>
__global__ __launch_bounds__(RW) __attribute__((amdgpu_waves_per_eu(sizeof(R) == 4 ? 3 : 2))) void k_brow_tiles(
    const BatchGrid* __restrict__ grids, const int* __restrict__ tile_grid,
    const int* __restrict__ tile_first) {
  __shared__ RowSmem<R> sm;
  __shared__ double red[3 * WPB];
  int t = blockIdx.x;
  const int nblk = gridDim.x;
  if ((nblk & 7) == 0) t = (t & 7) * (nblk >> 3) + (t >> 3);
  const int gi = tile_grid[t];
  const TileRed r = row_tile<R, kModeEuler>(batch_args(grids, gi), t - tile_first[gi], false, sm, red);
  if (threadIdx.x == 0) batch_publish(grids[gi], r);
}

// The 2-D grids of a batch in block rows, persistent blocks: block b walks
// its range of tiles, folding consecutive tiles of one grid.
template <class R>
__global__ __launch_bounds__(RW) __attribute__((amdgpu_waves_per_eu(sizeof(R) == 4 ? 3 : 2))) void k_brow_batch(
    const BatchGrid* __restrict__ grids, const int* __restrict__ tile_grid,
    const int* __restrict__ tile_first, const int* __restrict__ blk_first) {
  __shared__ RowSmem<R> sm;
  __shared__ double red[3 * WPB];
  const int t0 = blk_first[blockIdx.x], t1 = blk_first[blockIdx.x + 1];
  int cur = -1;
  TileRed acc{kDtReset, 0.0, 0.0};
  for (int t = t0; t < t1; t++) {
    const int gi = tile_grid[t];
    if (threadIdx.x == 0 && gi != cur && cur >= 0) {
      batch_publish(grids[cur], acc);
      acc = TileRed{kDtReset, 0.0, 0.0};
    }
    cur = gi;
    const TileRed r = row_tile<R, kModeEuler>(batch_args(grids, gi), t - tile_first[gi], false, sm, red);
    if (threadIdx.x == 0) {
      acc.dt = fmin(acc.dt, r.dt);
      acc.rd += r.rd;
      acc.rr += r.rr;
    }
    __syncthreads();  // LDS is reused by the next tile
  }
  if (threadIdx.x == 0 && cur >= 0) batch_publish(grids[cur], acc);
}

// The 2-D grids of a batch, persistent blocks: block b walks its range of
// wave tiles WPB at a time.
template <class R>
__global__ __launch_bounds__(RW) __attribute__((amdgpu_waves_per_eu(sizeof(R) == 4 ? 3 : 2))) void k_wrow_batch(
    const BatchGrid* __restrict__ grids, const int* __restrict__ tile_grid,
    const int* __restrict__ tile_first, const int* __restrict__ blk_first) {
  __shared__ double red[3 * WPB];
  __shared__ int gid[WPB];
  const int t0 = blk_first[blockIdx.x], t1 = blk_first[blockIdx.x + 1];
  int cur = -1;
  TileRed acc{kDtReset, 0.0, 0.0};
  for (int t = t0; t < t1; t += WPB)
    wave_tiles_step<R>(grids, tile_grid, tile_first, t, t1, red, gid, acc, cur);
  if (threadIdx.x == 0 && cur >= 0) batch_publish(grids[cur], acc);
}

// The 2-D grids of a batch, one block per WPB consecutive wave tiles
// (XCD-contiguous block ranges); counts[0] is the tile count.
template <class R>
__global__ __launch_bounds__(RW) __attribute__((amdgpu_waves_per_eu(sizeof(R) == 4 ? 3 : 2))) void k_wrow_tiles(
    const BatchGrid* __restrict__ grids, const int* __restrict__ tile_grid,
    const int* __restrict__ tile_first, const int* __restrict__ counts) {
  __shared__ double red[3 * WPB];
  __shared__ int gid[WPB];
  int b = blockIdx.x;
  const int nblk = gridDim.x;
  if ((nblk & 7) == 0) b = (b & 7) * (nblk >> 3) + (b >> 3);
  const int t0 = b * WPB;
  int cur = -1;
  TileRed acc{kDtReset, 0.0, 0.0};
  wave_tiles_step<R>(grids, tile_grid, tile_first, t0, min(t0 + WPB, counts[0]), red, gid, acc, cur);
  if (threadIdx.x == 0 && cur >= 0) batch_publish(grids[cur], acc);
}

// Closes one batch step for each grid: t += dt (cuf_kernel.cuf:792-794),
// re-arms the dt slot after next (:790), publishes the step's density
// residual and advances the grid's step index.
__global__ void k_batch_book(const BatchGrid* __restrict__ grids, const int* __restrict__ counts) {
  const int gi = blockIdx.x * blockDim.x + threadIdx.x;
  if (gi >= counts[1]) return;
  const BatchGrid& bg = grids[gi];
  const int n = *bg.nstep;
  bg.dt.t[0] = bg.dt.t[0] + bg.dt.slot[n % 3];
  bg.dt.slot[(n + 2) % 3] = kDtReset;
  bg.resid_last[0] = bg.resid[0];
  bg.resid_last[1] = bg.resid[1];
  bg.resid[0] = 0.0;
  bg.resid[1] = 0.0;
  *bg.nstep = n + 1;
}

template <class R, int TX, int TY, int DIM, int MODE>
static hipError_t launch_one(hipStream_t s, const FusedArgs& a) {
  if constexpr (DIM == 2) {  // 2-D grids march rows whatever the tile shape
    if (alone_wave_rows(a.g)) {
      const int nb = (wave_row_tiles(a.g, a.kchunk) + WPB - 1) / WPB;
      hipLaunchKernelGGL((k_wrows<R, MODE>), dim3((unsigned)nb), dim3(RW), 0, s, a);
    } else {
      hipLaunchKernelGGL((k_rows<R, MODE>), dim3((unsigned)row_tiles(a.g, a.kchunk)), dim3(RW), 0, s, a);
    }
  } else {
    const int ntx = (a.g.nx + TX - 1) / TX, nty = (a.g.ny + TY - 1) / TY;
    const int nkz = (a.g.nz + a.kchunk - 1) / a.kchunk;
    const long long nb = (long long)ntx * nty * nkz;
    // plane DMA: mixed form, 32x8 tiles (the staging fits 3 blocks/CU
    // beside the face data), rows of even length (16-byte pairs)
    constexpr bool kDma = sizeof(R) == 4 && TX == 32 && TY == 8;
    if (kDma && a.g.nx % 2 == 0)
      hipLaunchKernelGGL((k_fused<R, TX, TY, DIM, MODE, kDma>), dim3((unsigned)nb), dim3(TX * TY), 0, s, a);
    else
      hipLaunchKernelGGL((k_fused<R, TX, TY, DIM, MODE, false>), dim3((unsigned)nb), dim3(TX * TY), 0, s, a);
  }
  return hipGetLastError();
}

template <class R, int TX, int TY>
static hipError_t launch_tile(hipStream_t s, const FusedArgs& a, int mode) {
  const bool d3 = a.g.nz > 1;
  switch (mode) {
    case kModeEuler:
      return d3 ? launch_one<R, TX, TY, 3, kModeEuler>(s, a) : launch_one<R, TX, TY, 2, kModeEuler>(s, a);
    case kModeHeun1:
      return d3 ? launch_one<R, TX, TY, 3, kModeHeun1>(s, a) : launch_one<R, TX, TY, 2, kModeHeun1>(s, a);
    default:
      return d3 ? launch_one<R, TX, TY, 3, kModeHeun2>(s, a) : launch_one<R, TX, TY, 2, kModeHeun2>(s, a);
  }
}

template <class R>
static hipError_t launch_prec(hipStream_t s, const FusedArgs& a, int mode, int tx, int ty) {
  if (tx == 32 && ty == 8) return launch_tile<R, 32, 8>(s, a, mode);
  if (tx == 16 && ty == 16) return launch_tile<R, 16, 16>(s, a, mode);
  if (tx == 64 && ty == 4) return launch_tile<R, 64, 4>(s, a, mode);
  if (tx == 16 && ty == 8) return launch_tile<R, 16, 8>(s, a, mode);
  return hipErrorInvalidValue;
}

template <class R, int TX, int TY>
static hipError_t launch_batch_tile(hipStream_t s, const BatchLaunch& L) {
  if (L.dim == 2) {  // wave rows (FP64) or block rows (mixed), whatever the tile shape
    static_assert(batch_wave_rows(0) && !batch_wave_rows(1), "FP64 = double, mixed = float");
    if constexpr (sizeof(R) == 8) {
      if (L.per_tile)
        hipLaunchKernelGGL((k_wrow_tiles<R>), dim3(L.nblocks), dim3(RW), 0, s, L.grids, L.tile_grid,
                           L.tile_first, L.counts);
      else
        hipLaunchKernelGGL((k_wrow_batch<R>), dim3(L.nblocks), dim3(RW), 0, s, L.grids, L.tile_grid,
                           L.tile_first, L.blk_first);
    } else {
      if (L.per_tile)
        hipLaunchKernelGGL((k_brow_tiles<R>), dim3(L.nblocks), dim3(RW), 0, s, L.grids, L.tile_grid,
                           L.tile_first);
      else
        hipLaunchKernelGGL((k_brow_batch<R>), dim3(L.nblocks), dim3(RW), 0, s, L.grids, L.tile_grid,
                           L.tile_first, L.blk_first);
    }
  } else if (L.per_tile) {
    hipLaunchKernelGGL((k_fused_tiles<R, TX, TY, 3>), dim3(L.nblocks), dim3(TX * TY), 0, s, L.grids,
                       L.tile_grid, L.tile_first);
  } else {
    hipLaunchKernelGGL((k_fused_batch<R, TX, TY, 3>), dim3(L.nblocks), dim3(TX * TY), 0, s, L.grids,
                       L.tile_grid, L.tile_first, L.blk_first);
  }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_batch_book, dim3((L.cap_grids + 63) / 64), dim3(64), 0, s, L.grids, L.counts);
  return hipGetLastError();
}

hipError_t launch_batch(hipStream_t s, const BatchLaunch& L, int precision, int tile_x, int tile_y) {
  if (L.nblocks <= 0 || L.cap_grids <= 0) return hipSuccess;
  if (L.dim == 2) {  // wave rows: the tile shape does not apply
    tile_x = 32;
    tile_y = 8;
  }
#define SETMHD_BT(T)                                                         \
  if (tile_x == 32 && tile_y == 8) return launch_batch_tile<T, 32, 8>(s, L);   \
  if (tile_x == 16 && tile_y == 16) return launch_batch_tile<T, 16, 16>(s, L); \
  if (tile_x == 64 && tile_y == 4) return launch_batch_tile<T, 64, 4>(s, L);   \
  if (tile_x == 16 && tile_y == 8) return launch_batch_tile<T, 16, 8>(s, L);   \
  return hipErrorInvalidValue;
  if (precision == 0) {
    SETMHD_BT(double)
  } else {
    SETMHD_BT(float)
  }
#undef SETMHD_BT
}

hipError_t launch_fused(hipStream_t s, const FusedArgs& a, int mode, int precision, int tile_x,
                        int tile_y) {
  if (a.kchunk < 1) return hipErrorInvalidValue;
  return precision == 0 ? launch_prec<double>(s, a, mode, tile_x, tile_y)
                        : launch_prec<float>(s, a, mode, tile_x, tile_y);
}

}  // namespace setmhd
