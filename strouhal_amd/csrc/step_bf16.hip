// Single-lattice step kernels with BF16 storage.
#include "step_launch.h"

namespace strh {

hipError_t launch_step_bf16(const StepLaunch& L, hipStream_t s) { return pick_parity<ConvBF16>(L, s); }

}  // namespace strh
