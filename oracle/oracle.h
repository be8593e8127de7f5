/*
 * oracle.h -- CPU FP64 restatement of SET's ideal-MHD finite-volume step.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP
 * product (set_amd/libsetmhd.so) is compared against, and the CPU baseline
 * timed beside it.  Nothing in the product links, loads or calls it.
 *
 * It restates, line by line and in the same evaluation order, the
 * reference's CUDA-Fortran hot path:
 *   chlld            SET_Fortran_Cuda_MHD/Solvers.f90:1143-1859
 *   signum/minmod/linear                 Solvers.f90:108-150
 *   face pass  CUF_MGD_grans / _RK2      cuf_solvers.cuf:62-241, 243-440
 *   cell pass  CUF_MGD_cells / _cell_RK2 cuf_solvers.cuf:443-558, 561-794
 *   Source_MK                            cuf_kernel.cuf:215-288
 *   step loop (lagged dt)                cuf_kernel.cuf:758-885
 * on the reference's own array layout (Storage_module.f90:16-37):
 * column-major (component fastest), 1-based cell/face numbers, -1 = none.
 *
 * Fortran default-REAL literals that the reference mixes into REAL(8)
 * expressions (0.2, 0.02, 0.05, -0.1, 1e-7, ...) are promoted from their
 * single-precision value, as the Fortran standard prescribes; F4() below.
 */
#ifndef SET_ORACLE_H
#define SET_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_CPI4 12.56637061435917295384 /* Storage_module.f90:8  */
#define ORC_GGG (5.0 / 3.0)              /* Storage_module.f90:9  */
#define F4(x) ((double)(float)(x))       /* default-REAL literal  */

/* physics / semantics switches (bitmask), same meaning as SETMHD_F_* */
enum {
  ORC_F_ZERO_BW = 1,     /* w = Bx = By = Bz = 0 after update (cuf_solvers.cuf:725-733) */
  ORC_F_AXISYM = 2,      /* cylindrical y-sources /rr(2)   (cuf_solvers.cuf:706-736)   */
  ORC_F_SOURCE_MK = 4,   /* charge exchange source          (cuf_kernel.cuf:215-288)   */
  ORC_F_HELIO_HACKS = 8, /* LAX wedge, u<=-0.1 outer clamp, v=0.01 tail clamp          */
};

typedef struct orc_config {
  int32_t riemann;     /* chlld n_state: 0 LAX, 1 HLL, 2 HLLC, 3 HLLD              */
  int32_t recon;       /* 0 first order, 1 minmod TVD (cuf_solvers.cuf:363-412)     */
  int32_t time_scheme; /* 0 forward Euler (CUF_MGD_grans/cells), 1 Heun RK2         */
  int32_t flags;       /* ORC_F_*                                                   */
  double cfl;          /* dt_face = cfl*dist/(max|SL|,|SR|); reference F4(0.2)      */
  double dt_cap;       /* 'time = 100000.0' start value of the face min             */
  double r_skip_face;  /* faces with |centre(s1)| < r skipped; reference F4(0.02)  */
  double r_hold_cell;  /* cells with |centre| < r not updated; reference F4(0.05)  */
} orc_config;

typedef struct orc_mesh {
  int32_t n_cell, n_gran;
  const double *cell_center;        /* [3*n_cell]               */
  const double *cell_volume;        /* [n_cell]                 */
  const int32_t *cell_gran;         /* [6*n_cell] 1-based, 0 = none */
  const double *gran_normal;        /* [3*n_gran]               */
  const double *gran_square;        /* [n_gran]                 */
  const double *gran_center;        /* [3*n_gran]               */
  const int32_t *gran_neighbour;    /* [2*n_gran] 1-based, -1   */
  const int32_t *gran_neighbour_tvd;/* [2*n_gran] 1-based, -1   */
  const int32_t *gran_type;         /* [n_gran]                 */
} orc_mesh;

/* status codes, shared with the product's error protocol */
enum {
  ORC_OK = 0,
  ORC_ERR_RHO = -1,     /* density <= 0 after update (reference STOPs)            */
  ORC_ERR_REGION = -2,  /* chlld HLLD found no wave region (Solvers.f90:1816-1825) */
  ORC_ERR_FACE_RHO = -3 /* negative face density after TVD (cuf_solvers.cuf:414) */
};

/* point physics */
int orc_chlld(int n_state, double al, double be, double ge, double w,
              const double *qqq1, const double *qqq2, double *dsl, double *dsp,
              double *dsc, double *qqq, int n_disc);
int orc_signum(double x);
double orc_minmod(double x, double y);
double orc_linear(double x1, double t1, double x2, double t2, double x3,
                  double t3, double y);
void orc_source_mk(const double *cell_par16, double *out3);

/* The two states face gr (0-based) hands to chlld in stage `stage`:
 * boundary ghosts and the minmod TVD reconstruction with its positivity
 * fallback (cuf_solvers.cuf:300-420).  Returns 1 if the face is skipped,
 * 0, or ORC_ERR_FACE_RHO.                                                */
int orc_face_states(const orc_config *cfg, const orc_mesh *m, int gr, int stage,
                    const double *W, const double *W2, double *q1, double *q2);

/* passes over the face list / cell list.
 * stage: 0 = Euler kernels, 1/2 = Heun stage.  W = Cell_par (16*Nc),
 * W2 = Cell_par2 (16*Nc).  potok = Gran_POTOK (9*Nf).  dt_min is the
 * device scalar time_step2 (min-reduced in Euler and stage 1).           */
int orc_face_pass(const orc_config *cfg, const orc_mesh *m, int stage,
                  const double *W, const double *W2, double *potok,
                  double *dt_min, int nthreads);
int orc_cell_pass(const orc_config *cfg, const orc_mesh *m, int stage,
                  double dt, double *W, double *W2, const double *potok,
                  int nthreads);

/* min over faces of the CFL candidate of state W without writing fluxes  */
double orc_dt_scan(const orc_config *cfg, const orc_mesh *m, const double *W,
                   int nthreads);

/* CUDA_START_MGD's loop (cuf_kernel.cuf:774-885) for nsteps:
 * time_step = *dt_next; *dt_next = 1e15; *t += time_step; passes.
 * W2 must be a copy of W on the first call (Set_CUDA, cuf_kernel.cuf:361). */
int orc_run(const orc_config *cfg, const orc_mesh *m, double *W, double *W2,
            double *potok, double *t, double *dt_next, int nsteps,
            int nthreads);

/* ENA line of sight (SET_ENA/SET_ENA.cpp:341-956; ena.c).  A triangulation
 * of the cell centres: tri[3*t..3*t+2] are 0-based cell numbers.           */
typedef struct orc_ena_mesh {
  int n_cell, n_tri;
  const double *cell_center; /* [3*n_cell] */
  const double *cell_par;    /* [16*n_cell] */
  const int *tri;            /* [3*n_tri]  */
} orc_ena_mesh;

double orc_ena_velocity1(double u, double cp);
double orc_ena_maxwell(double rho, double cp, double u, double v, double w, double vx, double vy,
                       double vz);
double orc_ena_f(int which, double x);
double orc_ena_int1(double x, double cp);
int orc_ena_los(const orc_ena_mesh *m, double the, int n_v, double *jena);

#ifdef __cplusplus
}
#endif
#endif
