// Instantiations of the FOD and fraction kernels with 3 column slot(s) per lane.
#include "kernels.hpp"

namespace lsd {
#ifndef LSD_FAST_BUILD
LSD_LAUNCHER_DEF(3, 4)
LSD_LAUNCHER_DEF(3, 6)
LSD_LAUNCHER_DEF(3, 8)
#endif
} // namespace lsd
