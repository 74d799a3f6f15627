// Fused time-step kernels for models with no PML slabs (one translation unit
// per PML form, compiled in parallel).
#include "gpx_step.h"

namespace gpx {
StepKernels step_kernel_plain(int id_bytes, int P, bool C)
{
    return step_kernel_table<0, 0>(id_bytes, P, C);
}
}  // namespace gpx
