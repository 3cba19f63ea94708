// FP32 stage-state conversion; the FP32 residual kernels are instantiated in ty_inst_*.hip.
#include "ty_kernels_impl.h"

namespace ty {

template hipError_t launch_to_stage<float>(const double*, float*, int, int, double, hipStream_t);

}  // namespace ty
