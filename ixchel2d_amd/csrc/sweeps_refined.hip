// Sweep kernels and their launchers for the FP32-refined path (increment sweeps).
#define IXH_SWEEP_LAUNCH_DEFS
#include "sweep_launch.hpp"

namespace ixh {
IXH_INSTANTIATE_SWEEPS(SOLVE_DELTA, float)
}  // namespace ixh
