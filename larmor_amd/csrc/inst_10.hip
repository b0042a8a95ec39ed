// Instantiations of the FOD and fraction kernels with 10 column slot(s) per lane.
#include "kernels.hpp"

namespace lsd {
#ifndef LSD_FAST_BUILD
LSD_LAUNCHER_DEF(10, 4)
LSD_LAUNCHER_DEF(10, 6)
LSD_LAUNCHER_DEF(10, 8)
#endif
} // namespace lsd
