// Sweep kernels and their launchers for the PCR with upstream-order assembly.
#define IXH_SWEEP_LAUNCH_DEFS
#include "sweep_launch.hpp"

namespace ixh {
IXH_INSTANTIATE_SWEEPS(SOLVE_PCR_EXACT, double)
}  // namespace ixh
