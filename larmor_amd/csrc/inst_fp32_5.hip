// Instantiations of the plain FP32 form's FOD and fraction kernels with 5 column slot(s) per lane.
#define LSD_FP32 1
#include "kernels.hpp"

namespace lsd32 {
LSD_LAUNCHER_DEF(5, 4)
LSD_LAUNCHER_DEF(5, 6)
LSD_LAUNCHER_DEF(5, 8)
} // namespace lsd32
