// Instantiations of sw_fast_kernel<64, P> for P = 1, 2, 3, 4 (see sw_kernels.hip).
#define MHMSW_FAST_PART 1
#define MHMSW_FAST_G 64
#define MHMSW_FAST_P0 1
#define MHMSW_FAST_PS 1, 2, 3, 4
#include "sw_kernels.hip"
