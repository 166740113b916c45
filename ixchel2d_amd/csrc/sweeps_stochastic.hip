// Sweep kernels and their launchers for the FP32-stochastic path (increment
// sweeps over FP32 fields, stochastically rounded stores).
#define IXH_SWEEP_LAUNCH_DEFS
#include "sweep_launch.hpp"

namespace ixh {
IXH_INSTANTIATE_SWEEPS(SOLVE_DELTA_SR, float)
}  // namespace ixh
