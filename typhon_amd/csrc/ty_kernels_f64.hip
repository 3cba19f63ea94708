// The precision-independent time-step and steady-cycle kernels; the residual kernels of
// ty_kernels_impl.h are instantiated per precision, dimension and kernel in ty_inst_*.hip
// (separate translation units so that they compile in parallel).
#include "ty_kernels_impl.h"

namespace ty {

int grid_blocks(int n)
{
    int nb = (n + BLOCK - 1) / BLOCK;
    return (nb + 7) & ~7;
}

template hipError_t launch_timestep<2>(const DtArgs&, double, double*, int*, hipStream_t);
template hipError_t launch_timestep<3>(const DtArgs&, double, double*, int*, hipStream_t);
template hipError_t launch_to_stage<double>(const double*, double*, int, int, double, hipStream_t);

// steady cycle: per-block sums of the per-cell residual terms in a fixed order (wavefront
// butterflies, then the block's four wavefronts in order), so the residual is bitwise repeatable
__device__ __forceinline__ double wave_sumd(double v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

static __global__ __launch_bounds__(BLOCK) void k_norm_partial(const double* rnorm, int n, const int* stop,
                                                               double* part)
{
    __shared__ double sw[BLOCK / 64];
    if (*stop) return;
    const int c = blockIdx.x * BLOCK + threadIdx.x;
    double v = c < n ? rnorm[c] : 0.0;
    v = wave_sumd(v);
    if ((threadIdx.x & 63) == 0) sw[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = sw[0];
        for (int i = 1; i < BLOCK / 64; ++i) t += sw[i];
        part[blockIdx.x] = t;
    }
}

// update_field's residual (update_field.f90:42-60) summed over the blocks, then check_end_cycle
// (check_end_cycle.f90:32-33: residu_ref = max(residu_ref, cur_res), end at cur_res / residu_ref <=
// maxres or at maxit iterations) and the next iteration's CFL number (calc_zonetimestep.f90:69-73)
static __global__ __launch_bounds__(1024) void k_steady_update(SteadyArgs a)
{
    __shared__ double sw[16];
    if (a.st->stop) return;
    double v = 0.0;
    for (int i = threadIdx.x; i < a.nblk; i += blockDim.x) v += a.part[i];
    v = wave_sumd(v);
    if ((threadIdx.x & 63) == 0) sw[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x != 0) return;
    double cur = sw[0];
    for (int i = 1; i < (int)(blockDim.x >> 6); ++i) cur += sw[i];
    SteadyDev& s = *a.st;
    s.cur_res = cur;
    a.hist[s.iter] = cur;
    s.iter += 1;
    s.residu_ref = fmax_(s.residu_ref, cur);
    if (cur / s.residu_ref <= a.maxres || s.iter >= a.maxit) s.stop = 1;
    double cfl = a.stabnb * fmax_(1.0, 1.0 - log10(cur / s.residu_ref) - a.wlog);
    s.cfl = fmin_(cfl, a.cfl_max);
}

hipError_t launch_steady_update(const double* rnorm, int n, const SteadyArgs& a, double* part, hipStream_t s)
{
    const int nb = (n + BLOCK - 1) / BLOCK;
    hipLaunchKernelGGL(k_norm_partial, dim3(nb), dim3(BLOCK), 0, s, rnorm, n, (const int*)&a.st->stop, part);
    hipLaunchKernelGGL(k_steady_update, dim3(1), dim3(1024), 0, s, a);
    return hipGetLastError();
}


}  // namespace ty
