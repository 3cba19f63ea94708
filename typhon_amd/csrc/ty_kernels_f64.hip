// Instantiations of the residual kernels (ty_kernels_impl.h) for FP64 arithmetic,
// plus the precision-independent time-step kernels.
#include "ty_kernels_impl.h"

namespace ty {

int grid_blocks(int n)
{
    int nb = (n + BLOCK - 1) / BLOCK;
    return (nb + 7) & ~7;
}

#define TY_INST(DIM, R)                                                          \
    template hipError_t launch_gradient<DIM, R>(const KArgs<R>&, hipStream_t); \
    template hipError_t launch_flux<DIM, R>(const KArgs<R>&, hipStream_t);
TY_INST(2, double)
TY_INST(3, double)
template hipError_t launch_timestep<2>(const DtArgs&, double, double*, int*, hipStream_t);
template hipError_t launch_timestep<3>(const DtArgs&, double, double*, int*, hipStream_t);
template hipError_t launch_to_stage<double>(const double*, double*, int, int, double, hipStream_t);

}  // namespace ty
