// Instantiations of the final kernels with 1 column slot(s) per lane.
#include "kernels.hpp"

namespace lsd {
#ifndef LSD_FAST_BUILD
LSD_LAUNCHER_DEF(final, FinalArgs, 1, 4)
LSD_LAUNCHER_DEF(final, FinalArgs, 1, 6)
LSD_LAUNCHER_DEF(final, FinalArgs, 1, 8)
#endif
} // namespace lsd
