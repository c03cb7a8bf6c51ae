// Single-lattice step kernels with FP32 storage.
#include "step_launch.h"

namespace strh {

hipError_t launch_step_fp32(const StepLaunch& L, hipStream_t s) { return pick_parity<ConvF32>(L, s); }

}  // namespace strh
