// The damped half of the mixed-mesh kernels (artificial viscosity on): fp_mixed.hip compiled with
// FP_MIXED_PART 1, so that the two halves build in parallel.
#define FP_MIXED_PART 1
#include "fp_mixed.hip"
