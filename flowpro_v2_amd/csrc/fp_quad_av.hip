// The damped half of the quadrilateral kernels (artificial viscosity on): fp_quad.hip compiled with
// FP_QUAD_PART 1, so that the two halves build in parallel.
#define FP_QUAD_PART 1
#include "fp_quad.hip"
