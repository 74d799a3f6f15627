// Fused time-step kernels for models with HORIPML, one CFS term (one translation unit
// per PML form, compiled in parallel).
#include "gpx_step.h"

namespace gpx {
StepKernels step_kernel_hori1(int id_bytes, int P, bool C)
{
    return step_kernel_table<0, 1>(id_bytes, P, C);
}
}  // namespace gpx
