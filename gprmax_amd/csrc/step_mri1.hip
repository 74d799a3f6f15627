// Fused time-step kernels for models with MRIPML, one CFS term (one translation unit
// per PML form, compiled in parallel).
#include "gpx_step.h"

namespace gpx {
StepKernels step_kernel_mri1(int id_bytes, int P, bool C)
{
    return step_kernel_table<1, 1>(id_bytes, P, C);
}
}  // namespace gpx
