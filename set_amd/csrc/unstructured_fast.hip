// unstructured_fast.hip -- the drop-in kernels of unstructured.hip compiled
// for speed: the Makefile builds this file with -freciprocal-math and
// -ffp-contract=fast, so divisions become reciprocal estimates refined to
// about 1 ulp and products fuse into FMAs; square roots stay correctly
// rounded.  Selected by SETMHD_F_FAST; agrees with the exact kernels to
// rounding (tests/test_gpu_dropin.py) rather than bit for bit.
#define SETMHD_DROPIN_NS fast
#include "dropin_body.inc"
