/*
 * setmhd.h -- C ABI of the MI355X (gfx950) ideal-MHD finite-volume step.
 *
 * Drop-in for SET's CUDA-Fortran hot path (SergeyMSU/SET,
 * SET_Fortran_Cuda_MHD/).  The reference's boundary is a set of module-global
 * device arrays plus chevron launches:
 *   Set_CUDA            cuf_kernel.cuf:351-382   -> setmhd_upload_mesh
 *   time_step2 = 2e-9   cuf_kernel.cuf:758       -> setmhd_set_dt
 *   step loop, 4 launches per step
 *                       cuf_kernel.cuf:780-885   -> setmhd_step
 *   Send_data_to_Host   cuf_kernel.cuf:384-400   -> setmhd_download
 * The kernels CUF_MGD_grans(_RK2) / CUF_MGD_cell(s)(_RK2)
 * (cuf_solvers.cuf:62-794) and chlld (Solvers.f90:1143-1859) run inside
 * setmhd_step.  Every call returns an int status (SETMHD_OK = 0) where the
 * reference printed an error string or executed STOP.
 *
 * Arrays cross the boundary in the reference's own layout
 * (Storage_module.f90:16-37): Fortran column-major, component index
 * fastest, 1-based cell/face numbers, -1 = no neighbour, 0 = no face slot.
 * Cell_par holds primitives rho,u,v,w,p,Bx,By,Bz (B in SET's stored units,
 * chlld divides by sqrt(4 pi)) followed by 8 neutral-H moments.
 *
 * One context owns its device memory and one HIP stream; contexts are
 * independent and may be used from different host threads.
 */
#ifndef SETMHD_H
#define SETMHD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum setmhd_status {
  SETMHD_OK = 0,
  SETMHD_ERR_ARG = 1,        /* bad argument / state                         */
  SETMHD_ERR_HIP = 2,        /* a HIP runtime call failed                    */
  SETMHD_ERR_NOMEM = 3,      /* device allocation failed                     */
  SETMHD_ERR_RHO = 4,        /* density <= 0 after an update (ref: STOP,
                                cuf_solvers.cuf:708-711)                    */
  SETMHD_ERR_REGION = 5,     /* HLLD found no wave region (ref: STOP,
                                Solvers.f90:1816-1825)                      */
  SETMHD_ERR_FACE_RHO = 6,   /* negative face density after TVD (ref: STOP,
                                cuf_solvers.cuf:414-420)                    */
  SETMHD_ERR_UNSUPPORTED = 7 /* option not available on this path            */
};

/* flags (bitmask): the reference's heliosphere-specific semantics */
#define SETMHD_F_ZERO_BW 1     /* w = B = 0 after update, hydro energy (cuf_solvers.cuf:725-733) */
#define SETMHD_F_AXISYM 2      /* cylindrical y-sources /rr(2) (cuf_solvers.cuf:706-736)        */
#define SETMHD_F_SOURCE_MK 4   /* charge-exchange source (cuf_kernel.cuf:215-288)               */
#define SETMHD_F_HELIO_HACKS 8 /* LAX wedge :296-298, u<=-0.1 clamp :320-322, v clamp :783-785 */
#define SETMHD_F_FAST 16       /* drop-in kernels with FMA contraction and reciprocal-based
                                  division: same algorithm, agrees to rounding, not bit for bit */
#define SETMHD_F_REORDER 32    /* face-list meshes: the device holds the cells in Morton order
                                  of their centres and the faces by normal axis and first cell
                                  (setmhd_mesh_order); every call still takes and returns the
                                  caller's numbering.  Results are bit for bit the same. */
#define SETMHD_F_PER_CELL 64   /* face-list meshes, Heun: one launch per stage, each cell forming
                                  the faces of its Cell_gran slots itself (cuf_solvers.cuf:243-794
                                  without the Gran_POTOK round trip of cuf_kernel.cuf:849-876);
                                  every interior face is formed by both its cells.  Results are
                                  bit for bit the same.  Ignored (two launches per stage) when a
                                  face is missing from its first cell's Cell_gran. */

#define SETMHD_RIEMANN_LAX 0
#define SETMHD_RIEMANN_HLL 1
#define SETMHD_RIEMANN_HLLC 2
#define SETMHD_RIEMANN_HLLD 3

#define SETMHD_TIME_EULER 0 /* CUF_MGD_grans + CUF_MGD_cells (cuf_kernel.cuf:837-846) */
#define SETMHD_TIME_HEUN 1  /* CUF_MGD_*_RK2 stages 1 and 2 (cuf_kernel.cuf:847-877) */

#define SETMHD_PREC_FP64 0
#define SETMHD_PREC_MIXED 1 /* FP32 face states / Riemann solver, FP64 state, p and dt */

typedef struct setmhd_config {
  int32_t riemann;     /* chlld n_state (cuf_solvers.cuf:260 uses 2 = HLLC) */
  int32_t recon;       /* 0 first order, 1 minmod TVD (cuf_solvers.cuf:363-412) */
  int32_t time_scheme; /* SETMHD_TIME_*                                      */
  int32_t flags;       /* SETMHD_F_*                                         */
  double cfl;          /* dt_face = cfl*dist/max(|SL|,|SR|); ref 0.2 (REAL*4) at :429 */
  double dt_cap;       /* start value of the face min ('time = 100000.0', :259) */
  double r_skip_face;  /* faces with |centre(s1)| < r skipped; ref 0.02 (:282-284) */
  double r_hold_cell;  /* cells with |centre| < r held; ref 0.05 (:611)       */
  int32_t precision;   /* SETMHD_PREC_* (structured path)                    */
  int32_t device;      /* HIP device ordinal                                 */
  int32_t tile_x;      /* structured 3-D tile tile_x x tile_y: 32x8, 16x16,  */
  int32_t tile_y;      /* 64x4 or 16x8 (0 = 16x16 in FP64, 64x4 in the mixed
                          form; 2-D grids march rows whatever the tile)       */
  int32_t kchunk;      /* structured: z-planes (3-D) or rows (2-D) per tile
                          (0 = chosen per grid; results do not depend on it) */
} setmhd_config;

typedef struct setmhd_ctx setmhd_ctx;

/* Fills *cfg with the reference's production semantics (cuf_solvers.cuf:
 * HLLC, minmod TVD, Heun RK2, axisymmetric + charge-exchange sources,
 * w = B = 0, heliosphere clamps, CFL 0.2, skip radii 0.02/0.05). */
void setmhd_config_reference(setmhd_config *cfg);

/* Fills *cfg with plain ideal MHD on a Cartesian box: first-order HLLD,
 * forward Euler, no sources, the reference's dt formula
 * cfl*dist/max(|SL|,|SR|) (cuf_solvers.cuf:429) with its 0.2 replaced by
 * `cfl` (dist = half the cell width, so 0.4 gives dt = 0.2 h/max|S|), cells
 * inside r_hold held (the wind source). */
void setmhd_config_cartesian(setmhd_config *cfg, double cfl, double r_hold);

/* Creates a context on cfg->device.  Replaces the allocation half of
 * Set_CUDA (cuf_kernel.cuf:351-382). */
int setmhd_create(setmhd_ctx **out, const setmhd_config *cfg);
void setmhd_destroy(setmhd_ctx *ctx);

/* Unstructured drop-in: uploads SET's mesh and state exactly as Set_CUDA
 * does (cuf_kernel.cuf:357-381; arrays of Storage_module.f90:16-37).
 * gran_potok may be NULL (zeros).  Resets time to t0. */
int setmhd_upload_mesh(setmhd_ctx *ctx, int32_t n_cell, int32_t n_gran,
                       const double *cell_par /*[16*n_cell]*/,
                       const double *cell_center /*[3*n_cell]*/,
                       const double *cell_volume /*[n_cell]*/,
                       const int32_t *cell_gran /*[6*n_cell]*/,
                       const double *gran_normal /*[3*n_gran]*/,
                       const double *gran_square /*[n_gran]*/,
                       const double *gran_center /*[3*n_gran]*/,
                       const int32_t *gran_neighbour /*[2*n_gran]*/,
                       const int32_t *gran_neighbour_tvd /*[2*n_gran]*/,
                       const int32_t *gran_type /*[n_gran]*/,
                       const double *gran_potok /*[9*n_gran] or NULL*/,
                       double t0);

/* The numbering SETMHD_F_REORDER gives a face-list mesh on the device, for
 * callers that renumber their own files: cell_perm[i] (0-based) is the
 * caller's cell placed i-th, cells sorted by the Morton (Z-order) key of
 * their centre in the mesh's bounding box, 21 bits per axis, ties by
 * number; gran_perm[f] is the caller's face placed f-th, faces grouped by
 * the axis their normal is closest to (x, y, z; ties to the lower axis),
 * then sorted by the new place of their first cell (gran_neighbour(1,:)),
 * ties by number.  Host only (no device call).  No counterpart in the
 * reference, whose kernels take the mesh in the order Setka.cpp:1158-1567
 * writes it. */
int setmhd_mesh_order(int32_t n_cell, int32_t n_gran, const double *cell_center /*[3*n_cell]*/,
                      const double *gran_normal /*[3*n_gran]*/,
                      const int32_t *gran_neighbour /*[2*n_gran]*/,
                      int32_t *cell_perm /*[n_cell]*/, int32_t *gran_perm /*[n_gran]*/);

/* Structured Cartesian box nx*ny*nz (nz = 1: 2-D) on
 * [box0,box1]x[box2,box3]x[box4,box5], outflow (copy) boundaries on every
 * side.  cell_par is Cell_par for the cells numbered c = i + nx*(j+ny*k).
 * The state is kept on the device as FP64 conserved variables, SoA. */
int setmhd_upload_structured(setmhd_ctx *ctx, int32_t nx, int32_t ny,
                             int32_t nz, const double *box /*[6]*/,
                             const double *cell_par /*[16*nx*ny*nz]*/,
                             double t0);

/* time_step2 := dt (the dt the next step uses; reference 2e-9 at
 * cuf_kernel.cuf:758). */
int setmhd_set_dt(setmhd_ctx *ctx, double dt);

/* CFL minimum over all faces of the current state (the face pass's
 * reduction without an update); stores it as the next step's dt. */
int setmhd_dt_from_state(setmhd_ctx *ctx, double *dt_out);

/* Advances nsteps steps (cuf_kernel.cuf:780-885 without the dumps): dt of
 * step n is the face minimum found in step n-1 (lagged, :789-790).  All
 * steps are queued on the context's stream with no host synchronisation
 * between them (dt never leaves the device); the call returns when they
 * have finished, with the device error word checked once. */
int setmhd_step(setmhd_ctx *ctx, int32_t nsteps);

/* As setmhd_step, bracketed by HIP events on the context's stream;
 * *ms = elapsed device time of the nsteps steps. */
int setmhd_step_timed(setmhd_ctx *ctx, int32_t nsteps, float *ms);

/* Waits for the stream; returns the simulated time and the next dt. */
int setmhd_sync(setmhd_ctx *ctx, double *t, double *dt_next);

/* Send_data_to_Host (cuf_kernel.cuf:384-400): Cell_par (16*n_cell,
 * variables 9..16 as uploaded) and the simulated time. */
int setmhd_download(setmhd_ctx *ctx, double *cell_par, double *t);

/* Gran_POTOK of the last face pass (9*n_gran, flux x area; drop-in only). */
int setmhd_download_potok(setmhd_ctx *ctx, double *gran_potok);

/* Number of cells / faces of the uploaded problem. */
int setmhd_sizes(const setmhd_ctx *ctx, int64_t *n_cell, int64_t *n_gran);

/* Relative density change of the last step, ||rho^{n+1}-rho^n||_1 /
 * ||rho^n||_1 (structured path; serving convergence test). */
int setmhd_residual(setmhd_ctx *ctx, double *res);

/* ---- batches of structured grids (continuous batching) -------------------
 * A context can instead hold many independent structured grids ("slots").
 * All admitted grids advance together: one fused launch per step per
 * dimensionality, captured in a HIP graph; each grid keeps its own
 * lagged dt, time, step count and density residual on the device.  Grids
 * may be admitted and retired between calls.  Forward Euler only.
 * No counterpart in the reference (its solver holds one mesh); this is the
 * batching / serving form of the same step. */

/* Admits a grid (as setmhd_upload_structured) with its CFL-sized first dt;
 * *slot receives its id. */
int setmhd_batch_add(setmhd_ctx *ctx, int32_t nx, int32_t ny, int32_t nz,
                     const double *box /*[6]*/, const double *cell_par,
                     double t0, int32_t *slot);

/* Retires a grid: it takes part in no further step.  Its state stays
 * readable (setmhd_batch_query / setmhd_batch_download) until the next
 * setmhd_batch_wait or setmhd_batch_step returns; its memory is freed
 * there. */
int setmhd_batch_remove(setmhd_ctx *ctx, int32_t slot);

/* Advances every admitted grid by nsteps steps and waits for them. */
int setmhd_batch_step(setmhd_ctx *ctx, int32_t nsteps);

/* Queues nsteps steps of the admitted grids and returns at once.  While
 * they run, setmhd_batch_add (the new grid joins the next step) and
 * setmhd_batch_download / _query of retired or newly admitted grids proceed
 * on a second stream; calls that need a grid of the running step wait for
 * it. */
int setmhd_batch_step_async(setmhd_ctx *ctx, int32_t nsteps);

/* *done = 1 when the queued steps have finished (or none are queued),
 * else 0; never blocks. */
int setmhd_batch_poll(setmhd_ctx *ctx, int32_t *done);

/* Completes the queued steps: per-grid bookkeeping becomes current, device
 * errors are reported, retired grids are freed. */
int setmhd_batch_wait(setmhd_ctx *ctx);

/* Steps since admission, time, next dt and ||rho^{n+1}-rho^n||_1/||rho^n||_1
 * of the slot's last step (any pointer may be NULL). */
int setmhd_batch_query(setmhd_ctx *ctx, int32_t slot, int64_t *steps,
                       double *t, double *dt_next, double *residual);

/* Cell_par of a slot (16 doubles per cell) and its time. */
int setmhd_batch_download(setmhd_ctx *ctx, int32_t slot, double *cell_par,
                          double *t);

/* Queues a copy of a slot's Cell_par (16 doubles per cell) and time to
 * pinned host memory and returns at once (after the running step, if the
 * slot takes part in it); *ticket names the copy.  The copy runs on the
 * context's second stream, beside steps queued meanwhile; a retired slot's
 * memory is freed only after its queued copies have run. */
int setmhd_batch_download_async(setmhd_ctx *ctx, int32_t slot, int32_t *ticket);

/* Collects a queued copy.  wait = 0: returns at once, *done = 0 while the
 * copy runs; otherwise waits for it.  When done (*done = 1) cell_par and *t
 * (either may be NULL: the copy is then dropped, e.g. after
 * setmhd_batch_download_view) hold the slot's state and time and the
 * ticket is free. */
int setmhd_batch_download_collect(setmhd_ctx *ctx, int32_t ticket, int32_t wait, double *cell_par,
                                  double *t, int32_t *done);

/* The landed copy in place, without a host copy: never blocks; when the
 * copy has landed *done = 1 and *data points at its Cell_par (16 doubles
 * per cell, pinned host memory of the library) until the ticket is
 * collected; else *done = 0. */
int setmhd_batch_download_view(setmhd_ctx *ctx, int32_t ticket, const double **data, double *t,
                               int32_t *done);

/* Times probe_steps steps of the admitted batch for each launch shape
 * (persistent grids of 1, 2, 4, 8, 64 blocks per CU, or one block per tile,
 * reported as 0) and, for the 3-D grids, each tile shape (32x8, 16x16,
 * 64x4, 16x8; 2-D grids march 256-cell row segments whatever the tile and
 * report 32x8), the 2-D and the 3-D launch separately; restores the state
 * after each probe, keeps the fastest of each for this context and caches
 * them per (precision, dimensionalities present, size bucket).  Outputs, if
 * not NULL, are arrays of 2: [0] the 2-D launch, [1] the 3-D launch. */
int setmhd_batch_autotune(setmhd_ctx *ctx, int32_t probe_steps,
                          int32_t *tile_x, int32_t *tile_y,
                          int32_t *blocks_per_cu);

/* Sets the batch launch of 2-D (dim 2), 3-D (dim 3) or all (dim 0) grids
 * without timing: tile shape (32x8, 16x16, 64x4 or 16x8; 3-D grids only,
 * 2-D grids keep 32x8) and blocks per CU (1..64 persistent blocks, or 0 for
 * one block per tile).  Defaults: one block per tile; 32x8 tiles for the
 * 3-D grids in FP64, 64x4 in the mixed form. */
int setmhd_batch_set_launch(setmhd_ctx *ctx, int32_t dim, int32_t tile_x, int32_t tile_y,
                            int32_t blocks_per_cu);

/* Free and total device memory (admission control). */
int setmhd_mem_info(setmhd_ctx *ctx, int64_t *free_bytes, int64_t *total_bytes);

/* Snapshot of Cell_par and the time for a dump that overlaps the next steps
 * (the periodic Send_data_to_Host of cuf_kernel.cuf:912-919 blocks instead):
 * download_async queues the conversion behind the queued steps and the copy
 * to pinned host memory on a second stream, and returns; steps queued after
 * it do not change the snapshot.  download_wait waits for that copy and
 * fills cell_par (16 doubles per cell) and *t (may be NULL).  One snapshot
 * at a time: download_async fails while one is not yet collected. */
int setmhd_download_async(setmhd_ctx *ctx);
int setmhd_download_wait(setmhd_ctx *ctx, double *cell_par, double *t);

/* Cell_par(1:8) (rho, u, v, w, p, Bx, By, Bz) of n cells, 1-based cell
 * numbers, from the current state of a face-list mesh or a structured grid:
 * one gather and one copy of 8 n doubles into out (cell after cell), the
 * same values a download gives; *t the simulated time (may be NULL).  In
 * place of the probe series' single-element device reads and their syncs
 * (cuf_kernel.cuf:886-906). */
int setmhd_probe(setmhd_ctx *ctx, int32_t n, const int32_t *cells, double *out, double *t);

/* One forward-Euler step of a structured box held in the caller's device
 * memory, allocating nothing: no counterpart in the reference; for callers
 * that cannot allocate device memory (HIP graph capture, tight serving
 * loops).  d_in and d_out (not overlapping, on the current device) hold the FP64
 * conserved state as the structured path keeps it: 8 fields (rho, rho u,
 * rho v, rho w, E, bx, by, bz with b = B/sqrt(4 pi)) of nx*ny*nz doubles
 * each, cell c = i + nx*(j + ny*k); box and cfg as for
 * setmhd_upload_structured (first-order HLLD, forward Euler, no heliosphere
 * switches).  The step uses time step dt and writes U^{n+1} to d_out;
 * *dt_next receives the CFL minimum of d_in's faces (the next step's dt
 * under the reference's lagged rule, cuf_kernel.cuf:789-790) and
 * *residual ||rho^{n+1}-rho^n||_1 / ||rho^n||_1 (either may be NULL).  Runs
 * on `stream` (a hipStream_t; NULL: the null stream) and waits for it.
 * The step's scalars live in a 64-byte static device array shared by every
 * call on that device: calls that overlap on one device race.  Grids of
 * more than 2^31-1 cells are rejected (SETMHD_ERR_ARG). */
int setmhd_step_noalloc(const setmhd_config *cfg, int32_t nx, int32_t ny, int32_t nz,
                        const double *box /*[6]*/, const double *d_in, double *d_out, double dt,
                        void *stream, double *dt_next, double *residual);

/* ENA line-of-sight spectra of a solution (SET_ENA/SET_ENA.cpp:544-956,
 * Culc_ENA / Culc_ENA_the): for each angle (radians from the x-axis, in
 * the x-y plane of the axisymmetric solution) the line of sight from 5 AU
 * outwards, sampled every 1/30 AU until it leaves the triangulation, and
 * for each of n_v particle speeds v_i = vL + (i + 1/2)(vR - vL)/n_v
 * (vL = 1 km/s, vR = 1000 km/s) the ENA intensity jena[a*n_v + i] (before
 * the reference's output factor 6.75186e7); samples[a] = samples taken.
 * cell_center (3 per cell) and cell_par (Cell_par, 16 per cell) as in
 * setmhd_upload_mesh; tri (3 per triangle, 0-based cell numbers, counter-
 * clockwise) a triangulation of the cell centres -- the reference builds a
 * CGAL Delaunay triangulation (:242-257) -- and tri_nb[3 t + k] the triangle
 * across the edge opposite vertex k (-1 on the hull).  Synchronous on
 * `device` (the calling thread's current device is left as it was);
 * allocates its own device memory.  SETMHD_ERR_ARG also when a
 * walk between triangles does not end (inconsistent neighbours). */
int setmhd_ena_spectra(int32_t device, int32_t n_cell, const double *cell_center, const double *cell_par,
                       int32_t n_tri, const int32_t *tri, const int32_t *tri_nb, int32_t n_angle,
                       const double *angle, int32_t n_v, double *jena, int32_t *samples);

/* Message of the last failure of this context ("" if none). */
int setmhd_last_error(const setmhd_ctx *ctx, char *buf, size_t n);

#ifdef __cplusplus
}
#endif
#endif
