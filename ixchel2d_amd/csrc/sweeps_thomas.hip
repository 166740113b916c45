// Sweep kernels and their launchers for the bit-faithful path (Thomas, upstream-order assembly).
#define IXH_SWEEP_LAUNCH_DEFS
#include "sweep_launch.hpp"

namespace ixh {
IXH_INSTANTIATE_SWEEPS(SOLVE_THOMAS, double)
}  // namespace ixh
