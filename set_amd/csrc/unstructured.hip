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
#define SETMHD_DROPIN_NS exact
#include "dropin_body.inc"
