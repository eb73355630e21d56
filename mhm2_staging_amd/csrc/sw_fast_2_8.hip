// Instantiations of sw_fast_kernel<2, P> for P = 8, 16 (see sw_kernels.hip).
#define MHMSW_FAST_PART 1
#define MHMSW_FAST_G 2
#define MHMSW_FAST_P0 8
#define MHMSW_FAST_PS 8, 16
#include "sw_kernels.hip"
