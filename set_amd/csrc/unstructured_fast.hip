// unstructured_fast.hip -- the drop-in kernels of unstructured.hip compiled
// for speed: the Makefile builds this file with -freciprocal-math and
// -ffp-contract=fast, so divisions become reciprocal estimates refined to
// about 1 ulp and products fuse into FMAs; square roots stay correctly
// rounded.  Selected by SETMHD_F_FAST; agrees with the exact kernels to
// rounding (tests/test_gpu_dropin.py) rather than bit for bit.  Faces with
// zero field on both sides and the LAX or HLLC solver take hllc_b0, the same
// flux formed in the global frame.
#define SETMHD_DROPIN_NS fast
#define SETMHD_DROPIN_FAST 1
#include "dropin_body.inc"
