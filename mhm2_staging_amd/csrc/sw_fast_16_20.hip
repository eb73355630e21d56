// Instantiation of sw_fast_kernel<16, 20> (see sw_kernels.hip).
#define MHMSW_FAST_PART 1
#define MHMSW_FAST_G 16
#define MHMSW_FAST_P0 20
#define MHMSW_FAST_PS 20
#include "sw_kernels.hip"
