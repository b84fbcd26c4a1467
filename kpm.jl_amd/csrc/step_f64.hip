// step_f64.hip -- instantiations of the fused recurrence pass for
// ComplexF64 recurrence blocks (kernel template in step_kernel.h).
#include "step_kernel.h"

namespace kpm {

StepLauncher pick_step_f64(int lpr, int unr, int kb, bool realh, int mode) {
  return pick_step<KPM_VEC_F64>(lpr, unr, kb, realh, mode);
}

}  // namespace kpm
