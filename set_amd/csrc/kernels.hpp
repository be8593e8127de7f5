// kernels.hpp -- device-side data descriptors and launchers shared by the
// C ABI (capi.hip) and the kernel translation units.
#pragma once
#include "common.hpp"

namespace setmhd {

// SET's face-list mesh on the device (Storage_module.f90:16-37), 1-based.
struct UMesh {
  int n_cell, n_gran;
  const double* cell_center;  // [3*Nc]
  const double* cell_volume;  // [Nc]
  const int* cell_gran;       // [6*Nc]
  const double* gran_normal;  // [3*Nf]
  const double* gran_square;  // [Nf]
  const double* gran_center;  // [3*Nf]
  const int* gran_nb;         // [2*Nf]
  const int* gran_tvd;        // [2*Nf]
  const int* gran_type;       // [Nf]
  // [6*Nc]: Cell_gran with the cell's side as the sign, +face where the cell
  // is the face's first neighbour, -face otherwise, 0 for an empty slot (the
  // cell pass's gather table: no gather of Gran_neighbour per slot)
  const int* cell_gs;
};

struct UCfg {
  int riemann, recon, flags;
  double cfl, dt_cap, r_skip, r_hold;
};

namespace exact {
hipError_t launch_face_pass(hipStream_t s, const UMesh& m, const UCfg& c, const double* S,
                            double* potok, DtRing dt, int cur, int reduce, int book, int* err);
hipError_t launch_cell_fused(hipStream_t s, const UMesh& m, const UCfg& c, int stage, double* W, double* W2,
                             const double* H, double* potok, DtRing dt, int cur, int reduce, int book, int* err);
hipError_t launch_cell_pass(hipStream_t s, const UMesh& m, const UCfg& c, int stage, double* W,
                            double* W2, const double* H, const double* potok, DtRing dt, int cur,
                            int* err);
}  // namespace exact
namespace fast {
hipError_t launch_face_pass(hipStream_t s, const UMesh& m, const UCfg& c, const double* S,
                            double* potok, DtRing dt, int cur, int reduce, int book, int* err);
hipError_t launch_cell_fused(hipStream_t s, const UMesh& m, const UCfg& c, int stage, double* W, double* W2,
                             const double* H, double* potok, DtRing dt, int cur, int reduce, int book, int* err);
hipError_t launch_cell_pass(hipStream_t s, const UMesh& m, const UCfg& c, int stage, double* W,
                            double* W2, const double* H, const double* potok, DtRing dt, int cur,
                            int* err);
}  // namespace fast

// A uniform Cartesian box: state is FP64 conserved, one field of nc
// doubles per variable (rho, mx, my, mz, E, bx, by, bz), x fastest.
struct SGrid {
  int nx, ny, nz;       // nz == 1 for 2-D
  long long nc;         // nx*ny*nz, field stride
  double x0, y0, z0;    // box origin
  double dx, dy, dz;    // spacing (dz = 1 in 2-D)
  double ax, ay, az;    // face areas (2-D: ax = dy, ay = dx)
  double inv_vol;       // 1 / cell volume
  double distx, disty, distz;  // face-to-centre distance (CFL length)
  double cfl, dt_cap, r_hold;
};

enum : int { kModeEuler = 0, kModeHeun1 = 1, kModeHeun2 = 2 };

struct FusedArgs {
  SGrid g;
  const double* Uf;  // state whose face fluxes are taken
  const double* Un;  // U^n (Heun stage 2)
  double* Uo;        // output state
  DtRing dt;
  int cur;           // dt slot of this step
  int kchunk;        // z-planes per block (3-D)
  double* resid;     // [2]: sum |drho|, sum |rho| (or null)
  int* err;
};

// One grid of a batch (device-resident; see k_fused_batch).
struct BatchGrid {
  SGrid g;
  double* buf[2];     // ping-pong state buffers; U^n is buf[nstep & 1]
  DtRing dt;          // this grid's dt slots and time
  int* nstep;         // steps done since admission
  int ntiles;         // tiles of this grid
  int kchunk;
  double* resid;      // [2] accumulators of the running step
  double* resid_last; // [2] sums of the last finished step
  int* err;
};

// largest persistent grid of a batch launch, in blocks per compute unit
constexpr int kMaxBlocksPerCu = 64;

struct BatchLaunch {
  const BatchGrid* grids;  // device array [cap_grids]
  const int* tile_grid;    // [cap_tiles] grid of each tile
  const int* tile_first;   // [cap_grids] first global tile of each grid
  const int* counts;       // device [2]: tiles, grids in use
  const int* blk_first;    // [nblocks + 1] tile range of each block, split by cost
  int cap_tiles, cap_grids;
  int nblocks;             // persistent grid size, or the block count when per_tile
  int per_tile;            // 1: one block per tile (2-D: per WPB wave tiles), 0: persistent blocks
  int dim;                 // 2 or 3: all grids of one launch share it
};

// 2-D grids march rows, either in blocks of RW lanes that own RW - 2
// columns each (block rows: wide grids alone, and the mixed form's batches),
// or in waves of 64 lanes that own WRN = 62 columns each, WPB waves per block
// (wave rows: grids up to kWaveRowsMaxNx columns alone, and the FP64 form's
// batches)
constexpr int RW = 256, ROWN = RW - 2;
constexpr int WRN = 62, WPB = RW / 64;
constexpr int kWaveRowsMaxNx = 1024;

__host__ __device__ inline int row_tiles(const SGrid& g, int kchunk) {  // block rows
  return ((g.nx + ROWN - 1) / ROWN) * ((g.ny + kchunk - 1) / kchunk);
}
__host__ __device__ inline int wave_row_tiles(const SGrid& g, int kchunk) {  // wave rows
  return ((g.nx + WRN - 1) / WRN) * ((g.ny + kchunk - 1) / kchunk);
}
// a 2-D grid stepped alone marches wave rows
__host__ __device__ inline bool alone_wave_rows(const SGrid& g) { return g.nx <= kWaveRowsMaxNx; }
// the 2-D grids of a batch march wave rows in FP64, block rows in the mixed
// form (measured on 64 grids of 128-1024 columns: +5 % and -4 %)
constexpr bool batch_wave_rows(int precision) { return precision == 0; }

// tiles of grid g in a batch: 3-D, TX x TY columns x chunks of kchunk
// planes; 2-D, row segments (wave or block rows) x chunks of kchunk rows
inline int tiles_of(const SGrid& g, int tx, int ty, int kchunk, bool waves) {
  if (g.nz == 1) return waves ? wave_row_tiles(g, kchunk) : row_tiles(g, kchunk);
  const int nkz = (g.nz + kchunk - 1) / kchunk;
  return ((g.nx + tx - 1) / tx) * ((g.ny + ty - 1) / ty) * nkz;
}

hipError_t launch_batch(hipStream_t s, const BatchLaunch& L, int precision, int tile_x, int tile_y);

// precision: 0 FP64, 1 mixed.  Returns hipErrorInvalidValue for an
// unsupported tile shape.
hipError_t launch_fused(hipStream_t s, const FusedArgs& a, int mode, int precision, int tile_x,
                        int tile_y);

}  // namespace setmhd
