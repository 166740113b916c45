// Sweep kernels and their launchers for the FP64 path (PCR).
#define IXH_SWEEP_LAUNCH_DEFS
#include "sweep_launch.hpp"

namespace ixh {
IXH_INSTANTIATE_SWEEPS(SOLVE_PCR, double)
}  // namespace ixh
