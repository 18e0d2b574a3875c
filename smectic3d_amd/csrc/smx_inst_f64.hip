// Instantiates the FP64 pass kernels (FP64 path and all diagnostics).
#include "smx_launch.h"
template struct smx::PeriodicOps<double>;
template struct smx::NeumannOps<double>;
template struct smx::AdvectOps<double>;
template struct smx::NeuAdvOps<double>;
template struct smx::AdjointOps<double>;
