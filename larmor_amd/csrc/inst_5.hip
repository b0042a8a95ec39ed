// Instantiations of the FOD and fraction kernels with 5 column slot(s) per lane.
#include "kernels.hpp"

namespace lsd {
LSD_LAUNCHER_DEF(5, 4)
#ifndef LSD_FAST_BUILD
LSD_LAUNCHER_DEF(5, 6)
LSD_LAUNCHER_DEF(5, 8)
#endif
} // namespace lsd
