// Instantiations of sw_fast_kernel<8, P> for P = 14, 16 (see sw_kernels.hip).
#define MHMSW_FAST_PART 1
#define MHMSW_FAST_G 8
#define MHMSW_FAST_P0 14
#define MHMSW_FAST_PS 14, 16
#include "sw_kernels.hip"
