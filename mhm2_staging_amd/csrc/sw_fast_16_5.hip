// Instantiations of sw_fast_kernel<16, P> for P = 5, 6, 7, 8 (see sw_kernels.hip).
#define MHMSW_FAST_PART 1
#define MHMSW_FAST_G 16
#define MHMSW_FAST_P0 5
#define MHMSW_FAST_PS 5, 6, 7, 8
#include "sw_kernels.hip"
