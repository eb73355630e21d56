// klign includes the driver as "adept-sw/driver.hpp" (src/klign.cpp:62-64,
// include/klign.hpp:49-51); this forwards to the drop-in declarations.
#pragma once
#include "../adept_sw/driver.hpp"
