// Instantiation of sw_fast_kernel<16, 12> (see sw_kernels.hip).
#define MHMSW_FAST_PART 1
#define MHMSW_FAST_G 16
#define MHMSW_FAST_P0 12
#define MHMSW_FAST_PS 12
#include "sw_kernels.hip"
