// capi.hip -- the C ABI of include/setmhd.h: context, device state, step
// loop and error protocol.
//
// The reference keeps its state in module-global device arrays and drives
// four synchronous launches per step from the host, with the dt scalar
// round-tripped through the host every step (cuf_kernel.cuf:780-885).  Here
// a context owns its arrays and one stream; a step is a sequence of
// asynchronous launches whose dt and time live on the device (a three-slot
// dt ring: step n reads slot n%3, min-reduces into slot (n+1)%3), so nothing
// crosses to the host until setmhd_sync / setmhd_download.  The device
// error word replaces the reference's in-kernel STOP.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/setmhd.h"
#include "kernels.hpp"

using namespace setmhd;

struct setmhd_ctx {
  setmhd_config cfg;
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  std::string msg;
  int kind = 0;  // 0 none, 1 unstructured, 2 structured
  long long n_cell = 0, n_gran = 0;
  long long step = 0;  // steps done since upload (selects the dt slot)
  bool per_cell_ok = false;  // every face is in its first cell's Cell_gran (SETMHD_F_PER_CELL)
  // device scalars: dt ring [3], time [1], residual [2]; error words: [0] the
  // steps', [1] the batch admissions' (their scans run beside a step)
  double* d_scal = nullptr;
  int* d_err = nullptr;
  // unstructured
  UMesh um{};
  double *d_W = nullptr, *d_W2 = nullptr, *d_H = nullptr, *d_potok = nullptr;
  void* d_mesh = nullptr;
  // SETMHD_F_REORDER: device place -> caller's cell / face (0-based), and
  // caller's cell -> device place; empty when the caller's order is kept
  std::vector<int32_t> cperm, fperm, cinv;
  const int* d_cperm = nullptr;  // cperm on the device (inside d_mesh)
  // structured
  SGrid sg{};
  double *d_U = nullptr, *d_U2 = nullptr;  // Euler: ping-pong; Heun: U^n, U^1
  int cur_buf = 0;                         // Euler: which buffer holds U^n
  double* d_H16 = nullptr;                 // Cell_par(9:16) kept for download
  int kchunk = 32;  // z-planes per block of the single structured grid
  // batch of structured grids (kind 3)
  struct Slot {
    bool used = false;     // in the batch
    bool retired = false;  // removed, state kept readable until the next wait
    bool fresh = false;    // admitted since the last completed step (no mirror entry)
    bool in_step = false;  // part of the step in flight
    SGrid g{};
    double* buf[2] = {nullptr, nullptr};
    double* h16 = nullptr;
    long long steps = 0;  // steps since admission (mirrors the device nstep)
  };
  std::vector<Slot> slots;
  void* d_slot_scal = nullptr;  // SlotScal[cap]
  int slot_cap = 0;
  bool dirty = true;            // tables / graph need a rebuild
  BatchGrid* d_bgrid[2] = {nullptr, nullptr};  // per dimensionality (2-D, 3-D)
  int* d_tgrid[2] = {nullptr, nullptr};
  int* d_tfirst[2] = {nullptr, nullptr};
  int* d_counts[2] = {nullptr, nullptr};
  int* d_bfirst[2] = {nullptr, nullptr};
  BatchLaunch bl[2]{};
  bool tables_dirty = true;     // tables need re-writing (no new graph)
  std::vector<BatchGrid> stage_g[2];
  std::vector<int> stage_t[2], stage_f[2], stage_b[2];
  int stage_n[2][2] = {{0, 0}, {0, 0}};
  std::vector<struct SlotScal> scal_host;  // host mirror after each batch step
  bool mirror_ok = false;
  // overlap of admissions and downloads with a running batch step: a side
  // stream for their copies and conversions, pinned memory for the step's
  // mirror copy, grow-only device staging
  hipStream_t side = nullptr;
  hipEvent_t ev_side = nullptr, ev_step = nullptr;
  bool in_flight = false;
  struct SlotScal* h_mirror = nullptr;  // pinned, slot_cap entries
  // one staging buffer for admissions' uploads and retirements' downloads,
  // both on the side stream in its order; sized by the largest grid
  // admitted, so a download never grows it.  A grown-out buffer is freed
  // when the GPU is idle (a hipFree would wait for the running step).
  double* up_stage = nullptr;
  size_t up_cap = 0;  // doubles
  std::vector<double*> stage_old;
  hipGraph_t graph1 = nullptr, graphN = nullptr;
  hipGraphExec_t exec1 = nullptr, execN = nullptr;
  int graph_steps = 8;
  // batch launch per dimensionality ([0] 2-D, [1] 3-D): tile shape and
  // blocks per CU (0: one block per tile)
  int bt_x[2] = {32, 32}, bt_y[2] = {8, 8}, bt_bpc[2] = {0, 0};
  hipEvent_t ev_mid = nullptr;  // autotune: splits a probe between the two launches
  void* d_probe = nullptr;      // setmhd_probe: cell ids, then 8 doubles per cell
  int probe_cap = 0;
  // setmhd_download_async: device and pinned snapshot (Cell_par, then t), grow-only
  double* snap_dev = nullptr;
  double* snap_host = nullptr;
  size_t snap_cap = 0;  // doubles
  long long snap_n = 0;  // cells of the snapshot in flight
  bool snap_pending = false;
  hipEvent_t ev_snapc = nullptr, ev_snap = nullptr;
  int n_cu = 256;
  // setmhd_batch_download_async: queued copies of retired (or idle) slots to
  // pinned host memory on the side stream, one per ticket; pinned buffers
  // are kept for reuse when a ticket is collected
  struct Ticket {
    bool used = false;
    long long nc = 0;
    double t = 0.0;
    double* host = nullptr;  // pinned, cap doubles
    size_t cap = 0;
    hipEvent_t ev = nullptr;
  };
  std::vector<Ticket> tickets;
};

// per-slot device scalars of a batch
struct SlotScal {
  double slot[3];
  double t;
  double resid[2];
  double resid_last[2];
  int nstep;
  int pad[3];
};

namespace {

int fail(setmhd_ctx* c, int code, const std::string& m) {
  if (c) c->msg = m;
  return code;
}

#define HIPCHK(c, call)                                                                   \
  do {                                                                                    \
    hipError_t e_ = (call);                                                               \
    if (e_ != hipSuccess)                                                                 \
      return fail((c), SETMHD_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

template <class T>
int dalloc(setmhd_ctx* c, T** p, size_t n) {
  if (n == 0) n = 1;
  hipError_t e = hipMalloc((void**)p, n * sizeof(T));
  if (e != hipSuccess) return fail(c, SETMHD_ERR_NOMEM, "hipMalloc failed");
  return SETMHD_OK;
}

// A temporary device buffer freed on every return path (hipFree waits for
// the work queued on it).
struct StageBuf {
  double* p = nullptr;
  StageBuf() = default;
  StageBuf(const StageBuf&) = delete;
  StageBuf& operator=(const StageBuf&) = delete;
  ~StageBuf() {
    if (p) hipFree(p);
  }
};

void free_graphs(setmhd_ctx* c) {
  if (c->exec1) hipGraphExecDestroy(c->exec1);
  if (c->execN) hipGraphExecDestroy(c->execN);
  if (c->graph1) hipGraphDestroy(c->graph1);
  if (c->graphN) hipGraphDestroy(c->graphN);
  c->exec1 = c->execN = nullptr;
  c->graph1 = c->graphN = nullptr;
}

void free_batch(setmhd_ctx* c) {
  if (c->stream) hipStreamSynchronize(c->stream);
  if (c->side) hipStreamSynchronize(c->side);
  free_graphs(c);
  for (auto& tk : c->tickets) {
    if (tk.host) hipHostFree(tk.host);
    if (tk.ev) hipEventDestroy(tk.ev);
  }
  c->tickets.clear();
  for (auto& sl : c->slots) {
    for (double* p : {sl.buf[0], sl.buf[1], sl.h16})
      if (p) hipFree(p);
  }
  c->slots.clear();
  for (int d = 0; d < 2; d++) {
    for (void* p : {(void*)c->d_bgrid[d], (void*)c->d_tgrid[d], (void*)c->d_tfirst[d], (void*)c->d_counts[d],
                    (void*)c->d_bfirst[d]})
      if (p) hipFree(p);
    c->d_bfirst[d] = nullptr;
    c->d_bgrid[d] = nullptr;
    c->d_tgrid[d] = nullptr;
    c->d_tfirst[d] = nullptr;
    c->d_counts[d] = nullptr;
    c->bl[d] = BatchLaunch{};
  }
  if (c->d_slot_scal) hipFree(c->d_slot_scal);
  c->d_slot_scal = nullptr;
  c->slot_cap = 0;
  if (c->h_mirror) hipHostFree(c->h_mirror);
  c->h_mirror = nullptr;
  if (c->up_stage) hipFree(c->up_stage);
  for (double* p : c->stage_old) hipFree(p);
  c->stage_old.clear();
  c->up_stage = nullptr;
  c->up_cap = 0;
  c->in_flight = false;
  c->dirty = true;
  c->tables_dirty = true;
  c->mirror_ok = false;
}

void free_state(setmhd_ctx* c) {
  if (c->snap_pending) {  // a snapshot's copy may still read its staging
    hipEventSynchronize(c->ev_snap);
    c->snap_pending = false;
  }
  for (double** p : {&c->d_W, &c->d_W2, &c->d_H, &c->d_potok, &c->d_U, &c->d_U2, &c->d_H16}) {
    if (*p) hipFree(*p);
    *p = nullptr;
  }
  if (c->d_mesh) hipFree(c->d_mesh);
  c->d_mesh = nullptr;
  c->d_cperm = nullptr;
  c->cperm.clear();
  c->fperm.clear();
  c->cinv.clear();
  free_batch(c);
  c->kind = 0;
}

DtRing ring(setmhd_ctx* c) { return DtRing{c->d_scal, c->d_scal + 3}; }

// Reads and clears the device error word; maps it to a status.
int check_err(setmhd_ctx* c) {
  int h = 0;
  HIPCHK(c, hipMemcpyAsync(&h, c->d_err, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (h == 0) return SETMHD_OK;
  HIPCHK(c, hipMemsetAsync(c->d_err, 0, sizeof(int), c->stream));
  if (h & kErrRho) return fail(c, SETMHD_ERR_RHO, "density <= 0 after update");
  if (h & kErrRegion) return fail(c, SETMHD_ERR_REGION, "HLLD: no wave region (non-finite state)");
  return fail(c, SETMHD_ERR_FACE_RHO, "negative face density after reconstruction");
}

__global__ void k_prim_to_cons(const double* __restrict__ cp, long long nc, double* __restrict__ U) {
#pragma clang fp contract(off)
  long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nc) return;
  const double* w = cp + 16 * c;
  const double spi4 = sqrt(kCpi4);
  double rho = w[0], u = w[1], v = w[2], ww = w[3], p = w[4];
  double bx = w[5] / spi4, by = w[6] / spi4, bz = w[7] / spi4;
  U[c] = rho;
  U[nc + c] = rho * u;
  U[2 * nc + c] = rho * v;
  U[3 * nc + c] = rho * ww;
  U[4 * nc + c] = p / (kGamma - 1.0) + 0.5 * rho * ((u * u + v * v) + ww * ww) +
                  0.5 * ((bx * bx + by * by) + bz * bz);
  U[5 * nc + c] = bx;
  U[6 * nc + c] = by;
  U[7 * nc + c] = bz;
}

// Cell_par(1:8) primitives of cell c from the structured conserved state
// (SoA, b in Heaviside-Lorentz units) into o[0..7]
__device__ __forceinline__ void cons_to_prim_cell(const double* __restrict__ U, long long nc, long long c,
                                                  double* __restrict__ o) {
#pragma clang fp contract(off)
  const double spi4 = sqrt(kCpi4);
  double rho = U[c];
  double u = U[nc + c] / rho, v = U[2 * nc + c] / rho, w = U[3 * nc + c] / rho;
  double bx = U[5 * nc + c], by = U[6 * nc + c], bz = U[7 * nc + c];
  double p = (kGamma - 1.0) * (U[4 * nc + c] - 0.5 * rho * ((u * u + v * v) + w * w) -
                               0.5 * ((bx * bx + by * by) + bz * bz));
  o[0] = rho;
  o[1] = u;
  o[2] = v;
  o[3] = w;
  o[4] = p;
  o[5] = bx * spi4;
  o[6] = by * spi4;
  o[7] = bz * spi4;
}

__global__ void k_cons_to_prim(const double* __restrict__ U, long long nc, double* __restrict__ cp) {
  long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nc) return;
  cons_to_prim_cell(U, nc, c, cp + 16 * c);
}

// Probes: Cell_par(1:8) of n cells (0-based ids) from the face-list state W
// (8 primitives per cell) or, when W is null, the structured state U.
__global__ void k_probe(const double* __restrict__ W, const double* __restrict__ U, long long nc,
                        const int* __restrict__ cells, int n, double* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const long long c = cells[i];
  if (W) {
#pragma unroll
    for (int k = 0; k < 8; k++) out[8 * i + k] = W[8 * c + k];
  } else {
    cons_to_prim_cell(U, nc, c, out + 8 * i);
  }
}

// AoS16 <-> AoS8 column split for the unstructured state
// Cell_par (caller's numbering) <-> W, H (device numbering): device cell c
// is the caller's cell perm[c], or c when perm is null.
__global__ void k_split16(const double* __restrict__ cp, long long n, const int* __restrict__ perm,
                          double* __restrict__ W, double* __restrict__ H) {
  long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
  const long long o = perm ? perm[c] : c;
#pragma unroll
  for (int k = 0; k < 8; k++) {
    W[8 * c + k] = cp[16 * o + k];
    H[8 * c + k] = cp[16 * o + 8 + k];
  }
}
__global__ void k_join16(const double* __restrict__ W, const double* __restrict__ H, long long n,
                         const int* __restrict__ perm, double* __restrict__ cp) {
  long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
  const long long o = perm ? perm[c] : c;
#pragma unroll
  for (int k = 0; k < 8; k++) {
    cp[16 * o + k] = W[8 * c + k];
    cp[16 * o + 8 + k] = H[8 * c + k];
  }
}

// structured: keep Cell_par(9:16) for download
__global__ void k_take_h(const double* __restrict__ cp, long long n, double* __restrict__ H) {
  long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
#pragma unroll
  for (int k = 0; k < 8; k++) H[8 * c + k] = cp[16 * c + 8 + k];
}
__global__ void k_put_h(const double* __restrict__ H, long long n, double* __restrict__ cp) {
  long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
#pragma unroll
  for (int k = 0; k < 8; k++) cp[16 * c + 8 + k] = H[8 * c + k];
}

unsigned nblocks(long long n) { return (unsigned)((n + 255) / 256); }

// dt ring re-armed for "next step uses dt": slot (step%3) = dt,
// slot ((step+1)%3) = 1e15.
int arm_dt(setmhd_ctx* c, double dt) {
  double h[3];
  int cur = (int)(c->step % 3);
  HIPCHK(c, hipMemcpyAsync(h, c->d_scal, 3 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  h[cur] = dt;
  h[(cur + 1) % 3] = kDtReset;
  HIPCHK(c, hipMemcpyAsync(c->d_scal, h, 3 * sizeof(double), hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return SETMHD_OK;
}

int upload_scalars(setmhd_ctx* c, double t0) {
  double h[6] = {2.0e-9, kDtReset, kDtReset, t0, 0.0, 0.0};  // cuf_kernel.cuf:758
  c->step = 0;
  HIPCHK(c, hipMemcpyAsync(c->d_scal, h, sizeof h, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemsetAsync(c->d_err, 0, sizeof(int), c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return SETMHD_OK;
}

UCfg ucfg(const setmhd_config& k) {
  return UCfg{k.riemann, k.recon, k.flags, k.cfl, k.dt_cap, k.r_skip_face, k.r_hold_cell};
}

// planes (3-D) or rows (2-D) one tile marches, the configured value if set.
// Results do not depend on it (the marches run a chunk's boundary rows and
// planes through their loop bodies).  3-D alone: enough blocks for about
// four per CU with 32x8 columns, within [8, 32] planes (each chunk re-reads
// two boundary planes).  In a batch (the grids fill the chip together):
// 24 planes in 3-D, 16 rows in 2-D.  A 2-D grid alone: a tile
// re-reads one row below its chunk and starts up in about two row-times,
// and the launch runs in rounds of the resident tiles (two blocks per CU,
// the FP64 occupancy; a third mixed-form block adds little; a block is one
// tile of block rows or WPB tiles of wave rows), so take the chunk in
// [1, 32] with the fewest row-times over its rounds, (k + 3) x rounds,
// among those whose tiles fill half a round (one row when none does).
int kchunk_for(const setmhd_config& cfg, int n_cu, int nx, int ny, int nz, bool batch) {
  if (cfg.kchunk > 0) return cfg.kchunk;
  if (batch && nz > 1) return 24;
  if (nz <= 1) {
    if (batch) return 16;
    const bool waves = nx <= kWaveRowsMaxNx;
    const long long nseg = waves ? (nx + WRN - 1) / WRN : (nx + ROWN - 1) / ROWN;
    const long long slots = (waves ? 2LL * WPB : 2LL) * n_cu;
    int best = 1;
    long long best_cost = -1;
    for (int k = 1; k <= 32; k++) {
      const long long nb = nseg * ((ny + k - 1) / k);
      if (2 * nb < slots) break;  // nb does not grow with k
      const long long cost = (nb + slots - 1) / slots * (k + 3);
      if (best_cost < 0 || cost <= best_cost) {
        best = k;
        best_cost = cost;
      }
    }
    return best;
  }
  const long long txy = (long long)((nx + 31) / 32) * ((ny + 7) / 8);
  const long long kc = (nz * txy + 4LL * n_cu - 1) / (4LL * n_cu);
  return (int)std::min(32LL, std::max(8LL, kc));
}

int grid_kchunk(const setmhd_ctx* c, int nx, int ny, int nz, bool batch) {
  return kchunk_for(c->cfg, c->n_cu, nx, ny, nz, batch);
}

// Default z-march tiles of a grid alone: 16x16 in FP64 (256^3: +1.5 %
// against 32x8); in the mixed form 32x8, whose planes arrive by LDS DMA,
// where rows have even length (256^3: +3 % against 64x4), else 64x4
int tile_x_for(const setmhd_config& cfg, int nx) {
  if (cfg.tile_x > 0) return cfg.tile_x;
  if (cfg.precision == SETMHD_PREC_FP64) return 16;
  return nx % 2 == 0 ? 32 : 64;
}
int tile_y_for(const setmhd_config& cfg, int nx) {
  if (cfg.tile_y > 0) return cfg.tile_y;
  if (cfg.precision == SETMHD_PREC_FP64) return 16;
  return nx % 2 == 0 ? 8 : 4;
}
int tile_x(const setmhd_ctx* c) { return tile_x_for(c->cfg, c->sg.nx); }
int tile_y(const setmhd_ctx* c) { return tile_y_for(c->cfg, c->sg.nx); }

// drop-in kernels: the exact build, or the fast one under SETMHD_F_FAST
template <class... A>
hipError_t face_pass(const setmhd_ctx* c, A... a) {
  return (c->cfg.flags & SETMHD_F_FAST) ? fast::launch_face_pass(a...) : exact::launch_face_pass(a...);
}
template <class... A>
hipError_t cell_pass(const setmhd_ctx* c, A... a) {
  return (c->cfg.flags & SETMHD_F_FAST) ? fast::launch_cell_pass(a...) : exact::launch_cell_pass(a...);
}
template <class... A>
hipError_t cell_fused(const setmhd_ctx* c, A... a) {
  return (c->cfg.flags & SETMHD_F_FAST) ? fast::launch_cell_fused(a...) : exact::launch_cell_fused(a...);
}

// The uniform box of nx x ny x nz cells spanning box[6] (x0,x1,y0,y1,z0,z1;
// z ignored in 2-D), with the run's CFL number, dt cap and held radius.
SGrid make_sgrid(const setmhd_config& cfg, int nx, int ny, int nz, const double* box) {
  SGrid g{};
  g.nx = nx;
  g.ny = ny;
  g.nz = nz;
  g.nc = (long long)nx * ny * nz;
  g.x0 = box[0];
  g.y0 = box[2];
  g.z0 = nz > 1 ? box[4] : 0.0;
  g.dx = (box[1] - box[0]) / nx;
  g.dy = (box[3] - box[2]) / ny;
  g.dz = nz > 1 ? (box[5] - box[4]) / nz : 1.0;
  if (nz > 1) {
    g.ax = g.dy * g.dz;
    g.ay = g.dx * g.dz;
    g.az = g.dx * g.dy;
    g.inv_vol = 1.0 / (g.dx * g.dy * g.dz);
  } else {
    g.ax = g.dy;
    g.ay = g.dx;
    g.az = 0.0;
    g.inv_vol = 1.0 / (g.dx * g.dy);
  }
  g.distx = 0.5 * g.dx;
  g.disty = 0.5 * g.dy;
  g.distz = 0.5 * g.dz;
  g.cfl = cfg.cfl;
  g.dt_cap = cfg.dt_cap;
  g.r_hold = cfg.r_hold_cell;
  return g;
}

// one structured step; dt-only scan is mode -1 (no write)
int structured_step(setmhd_ctx* c, double* resid) {
  FusedArgs a{};
  a.g = c->sg;
  a.dt = ring(c);
  a.cur = (int)(c->step % 3);
  a.kchunk = c->kchunk;
  a.resid = resid;
  a.err = c->d_err;
  const int prec = c->cfg.precision;
  hipError_t e;
  if (c->cfg.time_scheme == SETMHD_TIME_EULER) {
    double* in = c->cur_buf == 0 ? c->d_U : c->d_U2;
    double* out = c->cur_buf == 0 ? c->d_U2 : c->d_U;
    a.Uf = in;
    a.Un = in;
    a.Uo = out;
    e = launch_fused(c->stream, a, kModeEuler, prec, tile_x(c), tile_y(c));
    if (e != hipSuccess) return fail(c, SETMHD_ERR_HIP, std::string("fused: ") + hipGetErrorString(e));
    c->cur_buf ^= 1;
  } else {
    a.Uf = c->d_U;
    a.Un = c->d_U;
    a.Uo = c->d_U2;
    a.resid = nullptr;
    e = launch_fused(c->stream, a, kModeHeun1, prec, tile_x(c), tile_y(c));
    if (e != hipSuccess) return fail(c, SETMHD_ERR_HIP, std::string("fused: ") + hipGetErrorString(e));
    a.Uf = c->d_U2;
    a.Un = c->d_U;
    a.Uo = c->d_U;
    a.resid = resid;
    e = launch_fused(c->stream, a, kModeHeun2, prec, tile_x(c), tile_y(c));
    if (e != hipSuccess) return fail(c, SETMHD_ERR_HIP, std::string("fused: ") + hipGetErrorString(e));
  }
  c->step++;
  return SETMHD_OK;
}

// Spreads the low 21 bits of v to every third bit (Morton interleave).
uint64_t spread3(uint64_t v) {
  v &= 0x1fffffULL;
  v = (v | v << 32) & 0x1f00000000ffffULL;
  v = (v | v << 16) & 0x1f0000ff0000ffULL;
  v = (v | v << 8) & 0x100f00f00f00f00fULL;
  v = (v | v << 4) & 0x10c30c30c30c30c3ULL;
  v = (v | v << 2) & 0x1249249249249249ULL;
  return v;
}

// Device numbering of a face-list mesh (setmhd_mesh_order): cells by the
// Morton key of their centre; faces grouped by the axis their normal is
// closest to, then by the new place of their first cell.  The grouping
// keeps the faces of one wave alike (on the heliosphere meshes the x-faces
// are mostly supersonic, the others not), so the Riemann solver's branches
// diverge less.  gran_neighbour(1,:) must lie in 1..n_cell.
void mesh_order(int32_t nc, int32_t nf, const double* cc, const int32_t* nb, const double* gn,
                int32_t* cperm, int32_t* fperm) {
  double lo[3] = {HUGE_VAL, HUGE_VAL, HUGE_VAL}, hi[3] = {-HUGE_VAL, -HUGE_VAL, -HUGE_VAL};
  for (int32_t i = 0; i < nc; i++)
    for (int d = 0; d < 3; d++) {
      lo[d] = std::min(lo[d], cc[3 * (size_t)i + d]);
      hi[d] = std::max(hi[d], cc[3 * (size_t)i + d]);
    }
  const double top = (double)((1 << 21) - 1);
  std::vector<uint64_t> key(nc);
  for (int32_t i = 0; i < nc; i++) {
    uint64_t k = 0;
    for (int d = 0; d < 3; d++) {
      double w = hi[d] > lo[d] ? (cc[3 * (size_t)i + d] - lo[d]) / (hi[d] - lo[d]) : 0.0;
      if (!(w >= 0.0)) w = 0.0;  // also NaN
      if (w > 1.0) w = 1.0;
      k |= spread3((uint64_t)(w * top)) << d;
    }
    key[i] = k;
  }
  for (int32_t i = 0; i < nc; i++) cperm[i] = i;
  std::sort(cperm, cperm + nc, [&](int32_t a, int32_t b) { return key[a] != key[b] ? key[a] < key[b] : a < b; });
  std::vector<int32_t> cinv(nc);
  for (int32_t i = 0; i < nc; i++) cinv[cperm[i]] = i;
  std::vector<uint8_t> axis(nf);
  for (int32_t f = 0; f < nf; f++) {
    const double* n = gn + 3 * (size_t)f;
    const double ax = fabs(n[0]), ay = fabs(n[1]), az = fabs(n[2]);
    axis[f] = ax >= ay && ax >= az ? 0 : (ay >= az ? 1 : 2);  // NaN: z
  }
  for (int32_t f = 0; f < nf; f++) fperm[f] = f;
  std::sort(fperm, fperm + nf, [&](int32_t a, int32_t b) {
    if (axis[a] != axis[b]) return axis[a] < axis[b];
    const int32_t ca = cinv[nb[2 * (size_t)a] - 1], cb = cinv[nb[2 * (size_t)b] - 1];
    return ca != cb ? ca < cb : a < b;
  });
}

int unstructured_step(setmhd_ctx* c) {
  const UCfg k = ucfg(c->cfg);
  const int cur = (int)(c->step % 3);
  DtRing r = ring(c);
  hipError_t e;
  if (c->cfg.time_scheme == SETMHD_TIME_EULER) {  // cuf_kernel.cuf:837-846
    e = face_pass(c, c->stream, c->um, k, c->d_W, c->d_potok, r, cur, 1, 1, c->d_err);
    if (e == hipSuccess)
      e = cell_pass(c, c->stream, c->um, k, 0, c->d_W, c->d_W2, c->d_H, c->d_potok, r, cur,
                           c->d_err);
  } else if ((c->cfg.flags & SETMHD_F_PER_CELL) && c->per_cell_ok) {  // :847-877 in two launches
    e = cell_fused(c, c->stream, c->um, k, 1, c->d_W, c->d_W2, c->d_H, c->d_potok, r, cur, 1, 1, c->d_err);
    if (e == hipSuccess)
      e = cell_fused(c, c->stream, c->um, k, 2, c->d_W, c->d_W2, c->d_H, c->d_potok, r, cur, 0, 0, c->d_err);
  } else {  // :847-877
    e = face_pass(c, c->stream, c->um, k, c->d_W, c->d_potok, r, cur, 1, 1, c->d_err);
    if (e == hipSuccess)
      e = cell_pass(c, c->stream, c->um, k, 1, c->d_W, c->d_W2, c->d_H, c->d_potok, r, cur,
                           c->d_err);
    if (e == hipSuccess)
      e = face_pass(c, c->stream, c->um, k, c->d_W2, c->d_potok, r, cur, 0, 0, c->d_err);
    if (e == hipSuccess)
      e = cell_pass(c, c->stream, c->um, k, 2, c->d_W, c->d_W2, c->d_H, c->d_potok, r, cur,
                           c->d_err);
  }
  if (e != hipSuccess) return fail(c, SETMHD_ERR_HIP, std::string("launch: ") + hipGetErrorString(e));
  c->step++;
  return SETMHD_OK;
}

}  // namespace

namespace {
int ensure_side(setmhd_ctx* c);  // batch_impl.inc: the context's second stream
hipError_t convert_state(setmhd_ctx* c, double* dst);
}  // namespace

extern "C" {

void setmhd_config_reference(setmhd_config* k) {
  std::memset(k, 0, sizeof *k);
  k->riemann = SETMHD_RIEMANN_HLLC;  // cuf_solvers.cuf:260
  k->recon = 1;                      // :363-412
  k->time_scheme = SETMHD_TIME_HEUN; // cuf_kernel.cuf:847-877
  k->flags = SETMHD_F_ZERO_BW | SETMHD_F_AXISYM | SETMHD_F_SOURCE_MK | SETMHD_F_HELIO_HACKS;
  k->cfl = SETMHD_F4(0.2);           // cuf_solvers.cuf:429
  k->dt_cap = 1.0e5;                 // :259
  k->r_skip_face = SETMHD_F4(0.02);  // :282
  k->r_hold_cell = SETMHD_F4(0.05);  // :611
  k->precision = SETMHD_PREC_FP64;
}

void setmhd_config_cartesian(setmhd_config* k, double cfl, double r_hold) {
  std::memset(k, 0, sizeof *k);
  k->riemann = SETMHD_RIEMANN_HLLD;
  k->recon = 0;
  k->time_scheme = SETMHD_TIME_EULER;
  k->flags = 0;
  // the reference's dt formula cfl*dist/max(|SL|,|SR|) (cuf_solvers.cuf:429)
  // with its constant 0.2 replaced by `cfl`; dist = h/2 on a uniform grid
  k->cfl = cfl;
  k->dt_cap = 1.0e5;
  k->r_skip_face = -1.0;
  k->r_hold_cell = r_hold;
  k->precision = SETMHD_PREC_FP64;
}

int setmhd_create(setmhd_ctx** out, const setmhd_config* cfg) {
  if (!out || !cfg) return SETMHD_ERR_ARG;
  *out = nullptr;
  if (cfg->riemann < 0 || cfg->riemann > 3 || cfg->recon < 0 || cfg->recon > 1 || cfg->kchunk < 0 ||
      cfg->time_scheme < 0 || cfg->time_scheme > 1 || cfg->precision < 0 || cfg->precision > 1)
    return SETMHD_ERR_ARG;
  setmhd_ctx* c = new setmhd_ctx();
  c->cfg = *cfg;
  c->device = cfg->device;
  if (cfg->kchunk > 0) c->kchunk = cfg->kchunk;
  if (cfg->precision == SETMHD_PREC_MIXED) {  // 3-D batch tiles: 64x4 in the mixed form
    c->bt_x[1] = 64;                         // (12 serving-mix grids: +6 % against 32x8)
    c->bt_y[1] = 4;
  }
  auto bad = [&](int code) {
    if (c->stream) hipStreamDestroy(c->stream);
    delete c;
    return code;
  };
  if (hipSetDevice(c->device) != hipSuccess) return bad(SETMHD_ERR_HIP);
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) return bad(SETMHD_ERR_HIP);
  if (hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess)
    return bad(SETMHD_ERR_HIP);
  if (hipMalloc((void**)&c->d_scal, 8 * sizeof(double)) != hipSuccess) return bad(SETMHD_ERR_NOMEM);
  if (hipMalloc((void**)&c->d_err, 2 * sizeof(int)) != hipSuccess) return bad(SETMHD_ERR_NOMEM);
  if (hipMemset(c->d_err, 0, 2 * sizeof(int)) != hipSuccess) return bad(SETMHD_ERR_HIP);
  if (upload_scalars(c, 0.0) != SETMHD_OK) return bad(SETMHD_ERR_HIP);
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, c->device) == hipSuccess && prop.multiProcessorCount > 0)
    c->n_cu = prop.multiProcessorCount;
  *out = c;
  return SETMHD_OK;
}

void setmhd_destroy(setmhd_ctx* c) {
  if (!c) return;
  hipSetDevice(c->device);
  if (c->stream) hipStreamSynchronize(c->stream);
  free_state(c);
  if (c->d_scal) hipFree(c->d_scal);
  if (c->d_err) hipFree(c->d_err);
  if (c->d_probe) hipFree(c->d_probe);
  if (c->side) hipStreamSynchronize(c->side);
  if (c->snap_dev) hipFree(c->snap_dev);
  if (c->snap_host) hipHostFree(c->snap_host);
  if (c->ev_snapc) hipEventDestroy(c->ev_snapc);
  if (c->ev_snap) hipEventDestroy(c->ev_snap);
  if (c->ev0) hipEventDestroy(c->ev0);
  if (c->ev1) hipEventDestroy(c->ev1);
  if (c->ev_side) hipEventDestroy(c->ev_side);
  if (c->ev_step) hipEventDestroy(c->ev_step);
  if (c->ev_mid) hipEventDestroy(c->ev_mid);
  if (c->side) hipStreamDestroy(c->side);
  if (c->stream) hipStreamDestroy(c->stream);
  delete c;
}

int setmhd_mesh_order(int32_t n_cell, int32_t n_gran, const double* cell_center,
                      const double* gran_normal, const int32_t* gran_neighbour, int32_t* cell_perm,
                      int32_t* gran_perm) {
  if (n_cell <= 0 || n_gran <= 0 || !cell_center || !gran_normal || !gran_neighbour || !cell_perm ||
      !gran_perm)
    return SETMHD_ERR_ARG;
  for (int32_t f = 0; f < n_gran; f++)
    if (gran_neighbour[2 * (size_t)f] < 1 || gran_neighbour[2 * (size_t)f] > n_cell) return SETMHD_ERR_ARG;
  mesh_order(n_cell, n_gran, cell_center, gran_neighbour, gran_normal, cell_perm, gran_perm);
  return SETMHD_OK;
}

int setmhd_upload_mesh(setmhd_ctx* c, int32_t n_cell, int32_t n_gran, const double* cell_par,
                       const double* cell_center, const double* cell_volume,
                       const int32_t* cell_gran, const double* gran_normal,
                       const double* gran_square, const double* gran_center,
                       const int32_t* gran_neighbour, const int32_t* gran_neighbour_tvd,
                       const int32_t* gran_type, const double* gran_potok, double t0) {
  if (!c) return SETMHD_ERR_ARG;
  if (n_cell <= 0 || n_gran <= 0 || !cell_par || !cell_center || !cell_volume || !cell_gran ||
      !gran_normal || !gran_square || !gran_center || !gran_neighbour || !gran_neighbour_tvd ||
      !gran_type)
    return fail(c, SETMHD_ERR_ARG, "upload_mesh: null array or empty mesh");
  if (c->cfg.precision != SETMHD_PREC_FP64)
    return fail(c, SETMHD_ERR_UNSUPPORTED, "the face-list path is FP64 only");
  // The validation of Set_Storage (Storage_module.f90:92-156): neighbour 1
  // in 1..N_cell, neighbour 2 and TVD neighbours <= N_cell, the neutral
  // moments Cell_par(9), (12), (13), (16) >= 0 and density >= 1e-9, compared
  // as the reference compares them (its 1e-9 is a default-REAL literal).
  for (int32_t f = 0; f < n_gran; f++) {
    int32_t s1 = gran_neighbour[2 * f], s2 = gran_neighbour[2 * f + 1];
    if (s1 < 1 || s1 > n_cell || s2 > n_cell || s2 == 0 || s2 < -1 ||
        gran_neighbour_tvd[2 * f] > n_cell || gran_neighbour_tvd[2 * f + 1] > n_cell)
      return fail(c, SETMHD_ERR_ARG, "upload_mesh: face neighbour out of range");
  }
  for (int32_t i = 0; i < n_cell; i++) {
    const double* cp = cell_par + 16 * (size_t)i;
    if (cp[8] < 0.0 || cp[11] < 0.0 || cp[12] < 0.0 || cp[15] < 0.0)
      return fail(c, SETMHD_ERR_ARG, "upload_mesh: Cell_par(9/12/13/16) < 0");
    if (cp[0] < SETMHD_F4(0.000000001))
      return fail(c, SETMHD_ERR_ARG, "upload_mesh: rho < 1e-9");
    for (int s = 0; s < 6; s++)
      if (cell_gran[6 * (size_t)i + s] < 0 || cell_gran[6 * (size_t)i + s] > n_gran)
        return fail(c, SETMHD_ERR_ARG, "upload_mesh: cell face number out of range");
  }
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  free_state(c);
  const size_t nc = n_cell, nf = n_gran;
  // SETMHD_F_REORDER: renumber the geometry on the host; the state is
  // permuted on the device (k_split16 / k_join16)
  const bool reorder = (c->cfg.flags & SETMHD_F_REORDER) != 0;
  std::vector<int32_t> cperm, fperm, cinv, finv, r_cg, r_nb, r_tvd, r_ty;
  std::vector<double> r_cc, r_cv, r_gn, r_gs, r_gc;
  if (reorder) {
    cperm.resize(nc);
    fperm.resize(nf);
    mesh_order(n_cell, n_gran, cell_center, gran_neighbour, gran_normal, cperm.data(), fperm.data());
    cinv.resize(nc);
    finv.resize(nf);
    for (size_t i = 0; i < nc; i++) cinv[cperm[i]] = (int32_t)i;
    for (size_t f = 0; f < nf; f++) finv[fperm[f]] = (int32_t)f;
    auto cmap = [&](int32_t v) { return v >= 1 ? cinv[v - 1] + 1 : v; };  // -1 / 0 stay
    r_cc.resize(nc * 3);
    r_cv.resize(nc);
    r_cg.resize(nc * 6);
    for (size_t i = 0; i < nc; i++) {
      const size_t o = cperm[i];
      for (int d = 0; d < 3; d++) r_cc[3 * i + d] = cell_center[3 * o + d];
      r_cv[i] = cell_volume[o];
      for (int k = 0; k < 6; k++) {
        const int32_t g = cell_gran[6 * o + k];
        r_cg[6 * i + k] = g >= 1 ? finv[g - 1] + 1 : g;
      }
    }
    r_gn.resize(nf * 3);
    r_gs.resize(nf);
    r_gc.resize(nf * 3);
    r_nb.resize(nf * 2);
    r_tvd.resize(nf * 2);
    r_ty.resize(nf);
    for (size_t f = 0; f < nf; f++) {
      const size_t o = fperm[f];
      for (int d = 0; d < 3; d++) {
        r_gn[3 * f + d] = gran_normal[3 * o + d];
        r_gc[3 * f + d] = gran_center[3 * o + d];
      }
      r_gs[f] = gran_square[o];
      for (int k = 0; k < 2; k++) {
        r_nb[2 * f + k] = cmap(gran_neighbour[2 * o + k]);
        r_tvd[2 * f + k] = cmap(gran_neighbour_tvd[2 * o + k]);
      }
      r_ty[f] = gran_type[o];
    }
    cell_center = r_cc.data();
    cell_volume = r_cv.data();
    cell_gran = r_cg.data();
    gran_normal = r_gn.data();
    gran_square = r_gs.data();
    gran_center = r_gc.data();
    gran_neighbour = r_nb.data();
    gran_neighbour_tvd = r_tvd.data();
    gran_type = r_ty.data();
  }
  // SETMHD_F_PER_CELL needs every face in its first cell's Cell_gran: that
  // cell writes the face's Gran_POTOK row and its CFL candidate counts
  c->per_cell_ok = true;
  for (size_t f = 0; f < nf && c->per_cell_ok; f++) {
    const int32_t* g = cell_gran + 6 * (size_t)(gran_neighbour[2 * f] - 1);
    bool listed = false;
    for (int k = 0; k < 6; k++) listed |= g[k] == (int32_t)f + 1;
    c->per_cell_ok = listed;
  }
  // the cell pass's gather table: each Cell_gran slot signed by the cell's
  // side of the face
  std::vector<int32_t> gsig(nc * 6);
  for (size_t i = 0; i < nc; i++)
    for (int k = 0; k < 6; k++) {
      const int32_t j = cell_gran[6 * i + k];
      gsig[6 * i + k] = j == 0 ? 0 : gran_neighbour[2 * (size_t)(j - 1)] == (int32_t)i + 1 ? j : -j;
    }
  // one block for the geometry (and the cell permutation)
  size_t bytes = nc * (3 + 1) * 8 + nf * (3 + 1 + 3) * 8 + 2 * nc * 6 * 4 + nf * (2 + 2 + 1) * 4 + 72 +
                 (reorder ? nc * 4 + 8 : 0);
  char* base = nullptr;
  if (hipMalloc((void**)&base, bytes) != hipSuccess) return fail(c, SETMHD_ERR_NOMEM, "hipMalloc mesh");
  c->d_mesh = base;
  size_t off = 0;
  auto take = [&](size_t n) {
    char* p = base + off;
    off += (n + 7) & ~size_t(7);
    return p;
  };
  double* cc = (double*)take(nc * 3 * 8);
  double* cv = (double*)take(nc * 8);
  double* gn = (double*)take(nf * 3 * 8);
  double* gs = (double*)take(nf * 8);
  double* gc = (double*)take(nf * 3 * 8);
  int* cg = (int*)take(nc * 6 * 4);
  int* cgs = (int*)take(nc * 6 * 4);
  int* gnb = (int*)take(nf * 2 * 4);
  int* gtv = (int*)take(nf * 2 * 4);
  int* gty = (int*)take(nf * 4);
  int* cpm = reorder ? (int*)take(nc * 4) : nullptr;
  HIPCHK(c, hipMemcpyAsync(cc, cell_center, nc * 3 * 8, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(cv, cell_volume, nc * 8, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(gn, gran_normal, nf * 3 * 8, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(gs, gran_square, nf * 8, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(gc, gran_center, nf * 3 * 8, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(cg, cell_gran, nc * 6 * 4, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(cgs, gsig.data(), nc * 6 * 4, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(gnb, gran_neighbour, nf * 2 * 4, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(gtv, gran_neighbour_tvd, nf * 2 * 4, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(gty, gran_type, nf * 4, hipMemcpyHostToDevice, c->stream));
  if (reorder) HIPCHK(c, hipMemcpyAsync(cpm, cperm.data(), nc * 4, hipMemcpyHostToDevice, c->stream));
  c->d_cperm = cpm;
  c->um = UMesh{n_cell, n_gran, cc, cv, cg, gn, gs, gc, gnb, gtv, gty, cgs};
  int st;
  if ((st = dalloc(c, &c->d_W, nc * 8)) || (st = dalloc(c, &c->d_W2, nc * 8)) ||
      (st = dalloc(c, &c->d_H, nc * 8)) || (st = dalloc(c, &c->d_potok, nf * 8)))
    return st;
  StageBuf stage;
  if ((st = dalloc(c, &stage.p, nc * 16))) return st;
  HIPCHK(c, hipMemcpyAsync(stage.p, cell_par, nc * 16 * 8, hipMemcpyHostToDevice, c->stream));
  hipLaunchKernelGGL(k_split16, dim3(nblocks(nc)), dim3(256), 0, c->stream, stage.p, (long long)nc,
                     c->d_cperm, c->d_W, c->d_H);
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipMemcpyAsync(c->d_W2, c->d_W, nc * 8 * 8, hipMemcpyDeviceToDevice, c->stream));  // :361
  if (gran_potok) {
    std::vector<double> p8(nf * 8);
    for (size_t f = 0; f < nf; f++) {
      const size_t o = reorder ? (size_t)fperm[f] : f;
      for (int k = 0; k < 8; k++) p8[8 * f + k] = gran_potok[9 * o + k];
    }
    HIPCHK(c, hipMemcpyAsync(c->d_potok, p8.data(), nf * 8 * 8, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
  } else {
    HIPCHK(c, hipMemsetAsync(c->d_potok, 0, nf * 8 * 8, c->stream));
  }
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->n_cell = n_cell;
  c->n_gran = n_gran;
  c->cperm = std::move(cperm);
  c->fperm = std::move(fperm);
  c->cinv = std::move(cinv);
  c->kind = 1;
  return upload_scalars(c, t0);
}

int setmhd_upload_structured(setmhd_ctx* c, int32_t nx, int32_t ny, int32_t nz, const double* box,
                             const double* cell_par, double t0) {
  if (!c) return SETMHD_ERR_ARG;
  if (nx < 1 || ny < 1 || nz < 1 || !box || !cell_par)
    return fail(c, SETMHD_ERR_ARG, "upload_structured: bad size or null array");
  if (c->cfg.riemann != SETMHD_RIEMANN_HLLD || c->cfg.recon != 0 ||
      (c->cfg.flags & (SETMHD_F_ZERO_BW | SETMHD_F_AXISYM | SETMHD_F_SOURCE_MK | SETMHD_F_HELIO_HACKS)))
    return fail(c, SETMHD_ERR_UNSUPPORTED,
                "structured path: first-order HLLD ideal MHD without heliosphere switches");
  const long long nc = (long long)nx * ny * nz;
  if (nc > (1LL << 31) - 1) return fail(c, SETMHD_ERR_ARG, "upload_structured: grid too large");
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  free_state(c);
  int st;
  if ((st = dalloc(c, &c->d_U, (size_t)nc * 8)) || (st = dalloc(c, &c->d_U2, (size_t)nc * 8)) ||
      (st = dalloc(c, &c->d_H16, (size_t)nc * 8)))
    return st;
  // Cell_par (16 doubles per cell) is staged in a temporary device buffer
  // and converted there (conserved state, neutral moments).
  StageBuf stage;
  if ((st = dalloc(c, &stage.p, (size_t)nc * 16))) return st;
  HIPCHK(c, hipMemcpyAsync(stage.p, cell_par, (size_t)nc * 16 * 8, hipMemcpyHostToDevice, c->stream));
  hipLaunchKernelGGL(k_prim_to_cons, dim3(nblocks(nc)), dim3(256), 0, c->stream, stage.p, nc, c->d_U);
  HIPCHK(c, hipGetLastError());
  hipLaunchKernelGGL(k_take_h, dim3(nblocks(nc)), dim3(256), 0, c->stream, stage.p, nc, c->d_H16);
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipMemcpyAsync(c->d_U2, c->d_U, (size_t)nc * 8 * 8, hipMemcpyDeviceToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->sg = make_sgrid(c->cfg, nx, ny, nz, box);
  c->kchunk = grid_kchunk(c, nx, ny, nz, false);
  c->cur_buf = 0;
  c->n_cell = nc;
  c->n_gran = nz > 1 ? (long long)(nx + 1) * ny * nz + (long long)nx * (ny + 1) * nz +
                           (long long)nx * ny * (nz + 1)
                     : (long long)(nx + 1) * ny + (long long)nx * (ny + 1);
  c->kind = 2;
  return upload_scalars(c, t0);
}

int setmhd_set_dt(setmhd_ctx* c, double dt) {
  if (!c || c->kind == 0) return SETMHD_ERR_ARG;
  if (!(dt > 0.0)) return fail(c, SETMHD_ERR_ARG, "set_dt: dt must be > 0");
  HIPCHK(c, hipSetDevice(c->device));
  return arm_dt(c, dt);
}

int setmhd_dt_from_state(setmhd_ctx* c, double* dt_out) {
  if (!c || c->kind == 0) return SETMHD_ERR_ARG;
  HIPCHK(c, hipSetDevice(c->device));
  // Run the face reduction on the current state without an update: the
  // unstructured face pass writes only fluxes; for the structured path an
  // Euler pass into the spare buffer (its output is discarded).
  HIPCHK(c, hipStreamSynchronize(c->stream));
  double save[6];
  HIPCHK(c, hipMemcpyAsync(save, c->d_scal, sizeof save, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  const int cur = (int)(c->step % 3);
  double h[6] = {save[0], save[1], save[2], save[3], save[4], save[5]};
  // the discarded update runs with dt = 0, so a large armed dt cannot drive
  // it to rho <= 0 and raise an error for a sound state
  h[cur] = 0.0;
  h[(cur + 1) % 3] = kDtReset;
  HIPCHK(c, hipMemcpyAsync(c->d_scal, h, sizeof h, hipMemcpyHostToDevice, c->stream));
  hipError_t e;
  if (c->kind == 1) {
    e = face_pass(c, c->stream, c->um, ucfg(c->cfg), c->d_W, c->d_potok, ring(c), cur, 1, 0,
                         c->d_err);
  } else {
    FusedArgs a{};
    a.g = c->sg;
    a.dt = ring(c);
    a.cur = cur;
    a.kchunk = c->kchunk;
    a.err = c->d_err;
    double* in = c->cur_buf == 0 ? c->d_U : c->d_U2;
    double* spare = c->cur_buf == 0 ? c->d_U2 : c->d_U;
    if (c->cfg.time_scheme == SETMHD_TIME_HEUN) {
      in = c->d_U;
      spare = c->d_U2;
    }
    a.Uf = in;
    a.Un = in;
    a.Uo = spare;
    e = launch_fused(c->stream, a, kModeHeun1, c->cfg.precision, tile_x(c), tile_y(c));
  }
  if (e != hipSuccess) return fail(c, SETMHD_ERR_HIP, hipGetErrorString(e));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  double r[6];
  HIPCHK(c, hipMemcpyAsync(r, c->d_scal, sizeof r, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  double dt = r[(cur + 1) % 3];
  // restore the ring and time, then arm dt for the next step
  save[cur] = dt;
  save[(cur + 1) % 3] = kDtReset;
  HIPCHK(c, hipMemcpyAsync(c->d_scal, save, sizeof save, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (dt_out) *dt_out = dt;
  return check_err(c);
}

int setmhd_step(setmhd_ctx* c, int32_t nsteps) {
  if (!c || c->kind == 0 || nsteps < 0) return SETMHD_ERR_ARG;
  HIPCHK(c, hipSetDevice(c->device));
  for (int32_t s = 0; s < nsteps; s++) {
    int st = c->kind == 1 ? unstructured_step(c) : structured_step(c, nullptr);
    if (st) return st;
  }
  return check_err(c);
}

int setmhd_step_timed(setmhd_ctx* c, int32_t nsteps, float* ms) {
  if (!c || c->kind == 0 || nsteps < 0) return SETMHD_ERR_ARG;
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, hipEventRecord(c->ev0, c->stream));
  for (int32_t s = 0; s < nsteps; s++) {
    int st = c->kind == 1 ? unstructured_step(c) : structured_step(c, nullptr);
    if (st) return st;
  }
  HIPCHK(c, hipEventRecord(c->ev1, c->stream));
  HIPCHK(c, hipEventSynchronize(c->ev1));
  if (ms) HIPCHK(c, hipEventElapsedTime(ms, c->ev0, c->ev1));
  return check_err(c);
}

int setmhd_residual(setmhd_ctx* c, double* res) {
  if (!c || c->kind != 2 || !res) return SETMHD_ERR_ARG;
  HIPCHK(c, hipSetDevice(c->device));
  double* r = c->d_scal + 4;
  HIPCHK(c, hipMemsetAsync(r, 0, 2 * sizeof(double), c->stream));
  int st = structured_step(c, r);
  if (st) return st;
  double h[2];
  HIPCHK(c, hipMemcpyAsync(h, r, sizeof h, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  *res = h[1] > 0.0 ? h[0] / h[1] : 0.0;
  return check_err(c);
}

int setmhd_sync(setmhd_ctx* c, double* t, double* dt_next) {
  if (!c) return SETMHD_ERR_ARG;
  HIPCHK(c, hipSetDevice(c->device));
  double h[4];
  HIPCHK(c, hipMemcpyAsync(h, c->d_scal, sizeof h, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (t) *t = h[3];
  if (dt_next) *dt_next = h[c->step % 3];
  return check_err(c);
}

int setmhd_download(setmhd_ctx* c, double* cell_par, double* t) {
  if (!c || c->kind == 0 || !cell_par) return SETMHD_ERR_ARG;
  HIPCHK(c, hipSetDevice(c->device));
  const long long nc = c->n_cell;
  double* stage = nullptr;
  int st;
  if ((st = dalloc(c, &stage, (size_t)nc * 16))) return st;
  hipError_t e = convert_state(c, stage);
  if (e == hipSuccess) e = hipMemcpyAsync(cell_par, stage, (size_t)nc * 16 * 8, hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  hipFree(stage);
  if (e != hipSuccess) return fail(c, SETMHD_ERR_HIP, std::string("download: ") + hipGetErrorString(e));
  return setmhd_sync(c, t, nullptr);
}

}  // extern "C"

namespace {
// Cell_par of the current state into dst (16 doubles per cell) on the
// library's stream.
hipError_t convert_state(setmhd_ctx* c, double* dst) {
  const long long nc = c->n_cell;
  if (c->kind == 1) {
    hipLaunchKernelGGL(k_join16, dim3(nblocks(nc)), dim3(256), 0, c->stream, c->d_W, c->d_H, nc, c->d_cperm,
                       dst);
  } else {
    const double* U = (c->cfg.time_scheme == SETMHD_TIME_EULER && c->cur_buf == 1) ? c->d_U2 : c->d_U;
    hipLaunchKernelGGL(k_cons_to_prim, dim3(nblocks(nc)), dim3(256), 0, c->stream, U, nc, dst);
    hipLaunchKernelGGL(k_put_h, dim3(nblocks(nc)), dim3(256), 0, c->stream, c->d_H16, nc, dst);
  }
  return hipGetLastError();
}
}  // namespace

extern "C" {

int setmhd_download_async(setmhd_ctx* c) {
  if (!c) return SETMHD_ERR_ARG;
  if (c->kind != 1 && c->kind != 2) return fail(c, SETMHD_ERR_ARG, "download_async: no mesh or grid uploaded");
  if (c->snap_pending) return fail(c, SETMHD_ERR_ARG, "download_async: the previous snapshot was not collected");
  HIPCHK(c, hipSetDevice(c->device));
  int st = ensure_side(c);
  if (st) return st;
  const size_t nc = (size_t)c->n_cell, nd = nc * 16 + 1;
  if (nd > c->snap_cap) {
    HIPCHK(c, hipStreamSynchronize(c->side));
    if (c->snap_dev) (void)hipFree(c->snap_dev);
    if (c->snap_host) (void)hipHostFree(c->snap_host);
    c->snap_dev = c->snap_host = nullptr;
    c->snap_cap = 0;
    if (hipMalloc((void**)&c->snap_dev, nd * 8) != hipSuccess)
      return fail(c, SETMHD_ERR_NOMEM, "download_async: hipMalloc");
    if (hipHostMalloc((void**)&c->snap_host, nd * 8, hipHostMallocDefault) != hipSuccess)
      return fail(c, SETMHD_ERR_NOMEM, "download_async: hipHostMalloc");
    c->snap_cap = nd;
  }
  if (!c->ev_snapc) HIPCHK(c, hipEventCreateWithFlags(&c->ev_snapc, hipEventDisableTiming));
  if (!c->ev_snap) HIPCHK(c, hipEventCreateWithFlags(&c->ev_snap, hipEventDisableTiming));
  // the conversion and the time follow the queued steps on the library's
  // stream (later steps cannot overwrite the state before it is taken); the
  // copy to the host runs on the side stream beside those later steps
  HIPCHK(c, convert_state(c, c->snap_dev));
  HIPCHK(c, hipMemcpyAsync(c->snap_dev + nc * 16, c->d_scal + 3, sizeof(double), hipMemcpyDeviceToDevice,
                           c->stream));
  HIPCHK(c, hipEventRecord(c->ev_snapc, c->stream));
  HIPCHK(c, hipStreamWaitEvent(c->side, c->ev_snapc, 0));
  HIPCHK(c, hipMemcpyAsync(c->snap_host, c->snap_dev, nd * 8, hipMemcpyDeviceToHost, c->side));
  HIPCHK(c, hipEventRecord(c->ev_snap, c->side));
  c->snap_n = (long long)nc;
  c->snap_pending = true;
  return SETMHD_OK;
}

int setmhd_download_wait(setmhd_ctx* c, double* cell_par, double* t) {
  if (!c || !cell_par) return SETMHD_ERR_ARG;
  if (!c->snap_pending) return fail(c, SETMHD_ERR_ARG, "download_wait: no snapshot queued");
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, hipEventSynchronize(c->ev_snap));
  c->snap_pending = false;
  std::memcpy(cell_par, c->snap_host, (size_t)c->snap_n * 16 * sizeof(double));
  if (t) *t = c->snap_host[(size_t)c->snap_n * 16];
  return SETMHD_OK;
}

int setmhd_probe(setmhd_ctx* c, int32_t n, const int32_t* cells, double* out, double* t) {
  if (!c) return SETMHD_ERR_ARG;
  if (c->kind != 1 && c->kind != 2) return fail(c, SETMHD_ERR_ARG, "probe: no mesh or grid uploaded");
  if (n < 1 || !cells || !out) return fail(c, SETMHD_ERR_ARG, "probe: no cells or null output");
  std::vector<int> ids(n);
  for (int32_t i = 0; i < n; i++) {
    if (cells[i] < 1 || cells[i] > c->n_cell) return fail(c, SETMHD_ERR_ARG, "probe: cell number out of range");
    ids[i] = c->cinv.empty() ? cells[i] - 1 : c->cinv[cells[i] - 1];
  }
  HIPCHK(c, hipSetDevice(c->device));
  if (n > c->probe_cap) {  // grow-only: ids, then 8 doubles per cell
    if (c->d_probe) (void)hipFree(c->d_probe);
    c->d_probe = nullptr;
    c->probe_cap = 0;
    if (hipMalloc(&c->d_probe, (size_t)n * (sizeof(int) + 8 * sizeof(double)) + 8) != hipSuccess)
      return fail(c, SETMHD_ERR_NOMEM, "probe: hipMalloc");
    c->probe_cap = n;
  }
  int* d_ids = (int*)c->d_probe;
  double* d_out = (double*)((char*)c->d_probe + (size_t)c->probe_cap * sizeof(int));
  d_out = (double*)(((uintptr_t)d_out + 7) & ~(uintptr_t)7);
  HIPCHK(c, hipMemcpyAsync(d_ids, ids.data(), (size_t)n * sizeof(int), hipMemcpyHostToDevice, c->stream));
  const double* W = c->kind == 1 ? c->d_W : nullptr;
  const double* U = c->kind == 2 ? ((c->cfg.time_scheme == SETMHD_TIME_EULER && c->cur_buf == 1) ? c->d_U2 : c->d_U)
                                 : nullptr;
  hipLaunchKernelGGL(k_probe, dim3((n + 255) / 256), dim3(256), 0, c->stream, W, U, (long long)c->n_cell, d_ids, n,
                     d_out);
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipMemcpyAsync(out, d_out, (size_t)n * 8 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return setmhd_sync(c, t, nullptr);
}

int setmhd_download_potok(setmhd_ctx* c, double* gran_potok) {
  if (!c || c->kind != 1 || !gran_potok) return SETMHD_ERR_ARG;
  HIPCHK(c, hipSetDevice(c->device));
  std::vector<double> p8((size_t)c->n_gran * 8);
  HIPCHK(c, hipMemcpyAsync(p8.data(), c->d_potok, p8.size() * 8, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  for (long long f = 0; f < c->n_gran; f++) {
    const long long o = c->fperm.empty() ? f : c->fperm[f];  // the caller's face
    for (int k = 0; k < 8; k++) gran_potok[9 * o + k] = p8[8 * f + k];
    gran_potok[9 * o + 8] = 0.0;  // cuf_solvers.cuf:433
  }
  return SETMHD_OK;
}

int setmhd_sizes(const setmhd_ctx* c, int64_t* n_cell, int64_t* n_gran) {
  if (!c) return SETMHD_ERR_ARG;
  if (n_cell) *n_cell = c->n_cell;
  if (n_gran) *n_gran = c->n_gran;
  return SETMHD_OK;
}

int setmhd_last_error(const setmhd_ctx* c, char* buf, size_t n) {
  if (!c || !buf || n == 0) return SETMHD_ERR_ARG;
  std::snprintf(buf, n, "%s", c->msg.c_str());
  return SETMHD_OK;
}

}  // extern "C"

#include "batch_impl.inc"

// ---- one structured step without allocation (setmhd_step_noalloc) --------
// The step's scalars -- dt slots, time, density-residual sums and the error
// word -- live in `scratch`, one copy per device, 64 bytes.  Host code never
// names the array in a device call: it uses the address hipGetSymbolAddress
// returns on each device, taken once per device with that device current.
namespace {

constexpr int kNoallocScratchDoubles = 8;  // [0..2] dt slots, [3] t, [4..5] residual sums, [6] error word
constexpr int kNoallocMaxDevices = 64;
static __device__ __align__(256) double scratch[kNoallocScratchDoubles];

double* g_scratch_ptr[kNoallocMaxDevices];
hipError_t g_scratch_err[kNoallocMaxDevices];
std::once_flag g_scratch_once[kNoallocMaxDevices];

// The device address of this device's scratch (device `dev` current).
double* scratch_on(int dev, hipError_t* err) {
  std::call_once(g_scratch_once[dev], [dev]() {
    void* p = nullptr;
    g_scratch_err[dev] = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    g_scratch_ptr[dev] = g_scratch_err[dev] == hipSuccess ? static_cast<double*>(p) : nullptr;
  });
  *err = g_scratch_err[dev];
  return g_scratch_ptr[dev];
}

}  // namespace

extern "C" int setmhd_step_noalloc(const setmhd_config* cfg, int32_t nx, int32_t ny, int32_t nz,
                                   const double* box, const double* d_in, double* d_out, double dt,
                                   void* stream, double* dt_next, double* residual) {
  if (!cfg || !box || !d_in || !d_out || d_in == d_out) return SETMHD_ERR_ARG;
  if (nx < 1 || ny < 1 || nz < 1 || (long long)nx * ny * nz > (1LL << 31) - 1) return SETMHD_ERR_ARG;
  {  // the step reads d_in while it writes d_out: the two states must not overlap
    const uintptr_t n = (uintptr_t)nx * ny * nz * 8 * sizeof(double);
    const uintptr_t i0 = (uintptr_t)d_in, o0 = (uintptr_t)d_out;
    if (i0 < o0 + n && o0 < i0 + n) return SETMHD_ERR_ARG;
  }
  if (!(dt > 0.0) || cfg->precision < 0 || cfg->precision > 1 || cfg->kchunk < 0) return SETMHD_ERR_ARG;
  if (cfg->riemann != SETMHD_RIEMANN_HLLD || cfg->recon != 0 || cfg->time_scheme != SETMHD_TIME_EULER ||
      (cfg->flags & (SETMHD_F_ZERO_BW | SETMHD_F_AXISYM | SETMHD_F_SOURCE_MK | SETMHD_F_HELIO_HACKS)))
    return SETMHD_ERR_UNSUPPORTED;
  int dev = 0, n_cu = 0;
  if (hipGetDevice(&dev) != hipSuccess) return SETMHD_ERR_HIP;
  if (dev < 0 || dev >= kNoallocMaxDevices) return SETMHD_ERR_UNSUPPORTED;
  if (hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n_cu < 1)
    return SETMHD_ERR_HIP;
  hipError_t e;
  double* sp = scratch_on(dev, &e);
  if (e != hipSuccess || !sp) return SETMHD_ERR_HIP;
  hipStream_t s = static_cast<hipStream_t>(stream);
  // dt of this step in slot 0, the face minimum reduced into slot 1; time,
  // residual sums and error word cleared (a copy, not a memset node)
  const double hin[kNoallocScratchDoubles] = {dt, kDtReset, kDtReset, 0.0, 0.0, 0.0, 0.0, 0.0};
  double h[kNoallocScratchDoubles];
  if (hipMemcpyAsync(sp, hin, sizeof hin, hipMemcpyHostToDevice, s) != hipSuccess) return SETMHD_ERR_HIP;
  FusedArgs a{};
  a.g = make_sgrid(*cfg, nx, ny, nz, box);
  a.Uf = d_in;
  a.Un = d_in;
  a.Uo = d_out;
  a.dt = DtRing{sp, sp + 3};
  a.cur = 0;
  a.kchunk = kchunk_for(*cfg, n_cu, nx, ny, nz, false);
  a.resid = sp + 4;
  a.err = reinterpret_cast<int*>(sp + 6);
  e = launch_fused(s, a, kModeEuler, cfg->precision, tile_x_for(*cfg, nx), tile_y_for(*cfg, nx));
  if (e != hipSuccess) return e == hipErrorInvalidValue ? SETMHD_ERR_ARG : SETMHD_ERR_HIP;
  if (hipMemcpyAsync(h, sp, sizeof h, hipMemcpyDeviceToHost, s) != hipSuccess) return SETMHD_ERR_HIP;
  if (hipStreamSynchronize(s) != hipSuccess) return SETMHD_ERR_HIP;
  if (dt_next) *dt_next = h[1];
  if (residual) *residual = h[5] > 0.0 ? h[4] / h[5] : 0.0;
  int err = 0;
  std::memcpy(&err, &h[6], sizeof err);
  if (err & kErrRho) return SETMHD_ERR_RHO;
  if (err & kErrRegion) return SETMHD_ERR_REGION;
  return SETMHD_OK;
}
