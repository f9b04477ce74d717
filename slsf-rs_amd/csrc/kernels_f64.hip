// FP64 replay kernels.  Built with -ffp-contract=off: every expression is
// evaluated as written, which the bit-exact replay parity relies on.
#include "launch.h"
#include "global_kernels.inc"
#include "slab.inc"
#include "replay_resident.inc"
#define SLSF_REAL double
#include "kernels_body.inc"
#undef SLSF_REAL
