// Instantiations of the FOD and fraction kernels with 10 column slot(s) per lane and 12 or 16 row
// slots (the plain FP64 fits of 513 to 1,024 measurements).
#include "kernels.hpp"

namespace lsd {
#ifndef LSD_FAST_BUILD
LSD_LAUNCHER_DEF(10, 12)
LSD_LAUNCHER_DEF(10, 16)
#endif
} // namespace lsd
