// Single-lattice step kernels with FP16 storage.
#include "step_launch.h"

namespace strh {

hipError_t launch_step_fp16(const StepLaunch& L, hipStream_t s) { return pick_parity<ConvF16>(L, s); }

}  // namespace strh
