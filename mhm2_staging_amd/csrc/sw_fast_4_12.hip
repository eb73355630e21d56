// Instantiations of sw_fast_kernel<4, P> for P = 12, 13 (see sw_kernels.hip).
#define MHMSW_FAST_PART 1
#define MHMSW_FAST_G 4
#define MHMSW_FAST_P0 12
#define MHMSW_FAST_PS 12, 13
#include "sw_kernels.hip"
