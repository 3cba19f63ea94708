// Instantiations of the residual kernels (ty_kernels_impl.h) for FP32 arithmetic.
#include "ty_kernels_impl.h"

namespace ty {

#define TY_INST(DIM, R)                                                          \
    template hipError_t launch_gradient<DIM, R>(const KArgs<R>&, hipStream_t); \
    template hipError_t launch_flux<DIM, R>(const KArgs<R>&, hipStream_t);
TY_INST(2, float)
TY_INST(3, float)
template hipError_t launch_to_stage<float>(const double*, float*, int, int, double, hipStream_t);

}  // namespace ty
