// Instantiates the FP32 pass kernels (FP32-mixed path).
#include "smx_launch.h"
template struct smx::PeriodicOps<float>;
template struct smx::NeumannOps<float>;
template struct smx::AdvectOps<float>;
template struct smx::NeuAdvOps<float>;
template struct smx::AdjointOps<float>;
