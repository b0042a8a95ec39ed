// Instantiations of the FOD and fraction kernels with 1 column slot(s) per lane.
#include "kernels.hpp"

namespace lsd {
LSD_LAUNCHER_DEF(1, 4)
#ifndef LSD_FAST_BUILD
LSD_LAUNCHER_DEF(1, 6)
LSD_LAUNCHER_DEF(1, 8)
#endif
} // namespace lsd
