// Vector kernels of the matrix-free implicit solve (restarted GMRES with
// classical Gram-Schmidt applied twice, block-Jacobi preconditioning, FD
// Jacobian-vector products).  Vectors are [nv][n] arrays of length len = nv*n.
// Reductions are two-pass (per-block partials in a fixed order, then one block per
// output), so every dot product is bitwise repeatable.
#include "ty_internal.h"
#include "ty_krylov.h"

namespace ty {

constexpr int KB = 256;  // threads per block for vector kernels

__device__ __forceinline__ double wsum(double v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// dst[c] = the stage-state (primitive) record of u[.][c] + a * v[.][c]: a perturbed state,
// pressure recovered in FP64 (g1 = gamma - 1)
// (ap, when given, holds a on the device)
template <typename P, typename T>
__global__ void k_fill_state(P* __restrict__ dst, const double* __restrict__ u, double a, const double* __restrict__ ap,
                             const T* __restrict__ v, int nv, int n, double g1)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    if (ap) a = *ap;
    double w[NVMAX];
    for (int k = 0; k < nv; ++k) {
        const size_t i = (size_t)k * n + c;
        w[k] = u[i] + a * (double)v[i];
    }
    store_prim<P>(w, nv, g1, dst + (size_t)c * nv);
}

// w = z - (r1 - r0) * s     (A z = z - J z with J z from a one-sided difference)
// central form: w = z - (r1 - r0) * s with r0 the residual at u - eps z
// (sp, when given, holds s on the device)
template <typename T>
__global__ void k_apply_a(T* __restrict__ w, const T* __restrict__ z, const double* __restrict__ r1,
                          const double* __restrict__ r0, double s, const double* __restrict__ sp, long len)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (sp) s = *sp;
    if (i < len) w[i] = (T)((double)z[i] - (r1[i] - r0[i]) * s);
}

// the one-sided JVP's step from a norm on the device: e[0] = eps = sq1u / sqrt(zz / len) (sq1u =
// sqrt(machine eps) (1 + ||u||)), e[1] = 1 / eps; both 0 for zz = 0, where the difference then
// vanishes and A z = z
__global__ void k_jvp_step(const double* __restrict__ zz, double len, double sq1u, double* __restrict__ e)
{
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    const double zn = sqrt(*zz / len);
    const double eps = zn == 0.0 ? 0.0 : sq1u / zn;
    e[0] = eps;
    e[1] = zn == 0.0 ? 0.0 : 1.0 / eps;
}

// z = Dinv v per cell (nv x nv blocks, [nv*nv][n] row-major entries)
template <int NV, typename T>
__global__ void k_block_apply(const T* __restrict__ dinv, const T* __restrict__ v, T* __restrict__ z, int n)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    T x[NV];
#pragma unroll
    for (int j = 0; j < NV; ++j) x[j] = v[(size_t)j * n + c];
#pragma unroll
    for (int i = 0; i < NV; ++i) {
        T s = T(0);
#pragma unroll
        for (int j = 0; j < NV; ++j) s += dinv[(size_t)(i * NV + j) * n + c] * x[j];
        z[(size_t)i * n + c] = s;
    }
}

// partial[b][i] = sum over block b of w . V_i, i < m (m <= MAXK): vectors in groups of MG held
// in registers (compile-time indexed), each group one grid-stride sweep that re-reads w
constexpr int MG = 8;
constexpr int CH = 8;  // basis vectors loaded together by k_cgs_pass and k_lincomb
// the fixed grid of the multi-dot and the fused CGS dot pass (the same grid, so their partial sums
// agree bit for bit).  With the dot pass's registers bounded by the basis size, 4096 blocks keep more
// waves in flight than 1024 (implicit 4.1M-hex step 117.9 -> 113.2 ms); 2048, 8192 and 16384 measured
// 117.4, 114.7 and 119.1 ms
constexpr long MD_BLOCKS = 4096;
template <typename T>
__global__ __launch_bounds__(KB) void k_multidot(const T* __restrict__ w, const T* __restrict__ V, long stride,
                                                 int m, long len, double* __restrict__ partial)
{
    __shared__ double red[KB / 64][MG];
    const int wid = threadIdx.x >> 6, lane = threadIdx.x & 63;
    for (int g0 = 0; g0 < m; g0 += MG) {
        const int mg = m - g0 < MG ? m - g0 : MG;
        const T* Vg = V + (long)g0 * stride;
        double acc[MG];
#pragma unroll
        for (int i = 0; i < MG; ++i) acc[i] = 0.0;
        for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < len; e += (long)gridDim.x * blockDim.x) {
            const double we = (double)w[e];
            T vc[MG];  // loads issued together, indices clamped to the group's last vector
#pragma unroll
            for (int i = 0; i < MG; ++i) vc[i] = Vg[(long)min(i, mg - 1) * stride + e];
#pragma unroll
            for (int i = 0; i < MG; ++i)
                if (i < mg) acc[i] += we * (double)vc[i];
        }
#pragma unroll
        for (int i = 0; i < MG; ++i) {
            const double sv = wsum(acc[i]);
            if (lane == 0) red[wid][i] = sv;
        }
        __syncthreads();
        if (threadIdx.x < mg) {
            double sv = 0.0;
            for (int k = 0; k < KB / 64; ++k) sv += red[k][threadIdx.x];
            partial[(long)blockIdx.x * MAXK + g0 + threadIdx.x] = sv;
        }
        __syncthreads();
    }
}

// out[i] = sum over blocks of partial[b][i]: one block per output, strided per-thread sums, a
// butterfly within each wave and the four wave sums in order (fixed summation order)
__global__ __launch_bounds__(KB) void k_reduce_partials(const double* __restrict__ partial, int nblk, int m,
                                                        double* __restrict__ out)
{
    __shared__ double red[KB / 64];
    const int i = blockIdx.x;
    if (i >= m) return;
    double sv = 0.0;
    for (int b = threadIdx.x; b < nblk; b += KB) sv += partial[(long)b * MAXK + i];
    sv = wsum(sv);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = sv;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = red[0];
        for (int k = 1; k < KB / 64; ++k) t += red[k];
        out[i] = t;
    }
}

// One fused classical Gram-Schmidt pass after the first multi-dot: w -= sum_i h_i V_i (as
// a CGS update), then, from the updated w as stored, either the partial dots V_i . w for the
// second pass (DOTS) or the partial ||w||^2.  The DOTS form runs on k_multidot's grid, so each
// element is owned by the thread k_multidot gives it and the partials equal what k_multidot
// would return on the stored w, bit for bit; the basis is read once instead of twice.  The norm
// form runs on a larger grid (fixed, so its sum is repeatable too).
// NM >= m bounds the basis held in registers: the launcher picks 8, 16, 24 or MAXK, so that short
// bases (the first steps of every cycle) run with fewer registers and more waves in flight.
template <typename T, bool DOTS, int NM>
__global__ __launch_bounds__(KB) void k_cgs_pass(T* __restrict__ w, const T* __restrict__ V, long stride,
                                                 const double* __restrict__ h, int m, long len,
                                                 double* __restrict__ partial)
{
    constexpr int NA = DOTS ? NM : 1;
    __shared__ double red[KB / 64][NA];
    const int wid = threadIdx.x >> 6, lane = threadIdx.x & 63;
    double hr[NM], acc[NA];
#pragma unroll
    for (int i = 0; i < NM; ++i) hr[i] = i < m ? h[i] : 0.0;
#pragma unroll
    for (int i = 0; i < NA; ++i) acc[i] = 0.0;
    for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < len; e += (long)gridDim.x * blockDim.x) {
        T vi[NM];
        double s = (double)w[e];
        // the basis in chunks of CH vectors: each chunk's loads issued together (indices clamped to
        // the last vector, so no load waits behind a branch), applied in vector order
#pragma unroll
        for (int c0 = 0; c0 < NM; c0 += CH) {
            if (c0 >= m) break;
            T vc[CH];
#pragma unroll
            for (int j = 0; j < CH; ++j) vc[j] = V[(long)min(c0 + j, m - 1) * stride + e];
#pragma unroll
            for (int j = 0; j < CH; ++j) {
                if (c0 + j < m) {
                    vi[c0 + j] = vc[j];
                    s -= hr[c0 + j] * (double)vc[j];
                }
            }
        }
        const T ws = (T)s;
        w[e] = ws;
        const double wd = (double)ws;
        if (DOTS) {
#pragma unroll
            for (int i = 0; i < NA; ++i)
                if (i < m) acc[i] += wd * (double)vi[i];
        } else {
            acc[0] += wd * wd;
        }
    }
    const int no = DOTS ? m : 1;
#pragma unroll
    for (int i = 0; i < NA; ++i) {
        if (i < no) {
            const double sv = wsum(acc[i]);
            if (lane == 0) red[wid][i] = sv;
        }
    }
    __syncthreads();
    if (threadIdx.x < no) {
        double sv = 0.0;
        for (int k = 0; k < KB / 64; ++k) sv += red[k][threadIdx.x];
        partial[(long)blockIdx.x * MAXK + threadIdx.x] = sv;
    }
}

// out = sum_i y_i V_i
template <typename T>
__global__ void k_lincomb(T* __restrict__ out, const T* __restrict__ V, long stride, const double* __restrict__ y,
                          int m, long len)
{
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= len) return;
    double s = 0.0;
    // in chunks of CH vectors whose loads are issued together (indices clamped), summed in order
    for (int c0 = 0; c0 < m; c0 += CH) {
        T vc[CH];
#pragma unroll
        for (int j = 0; j < CH; ++j) vc[j] = V[(long)min(c0 + j, m - 1) * stride + e];
#pragma unroll
        for (int j = 0; j < CH; ++j)
            if (c0 + j < m) s += y[c0 + j] * (double)vc[j];
    }
    out[e] = (T)s;
}

// dst = a * src (type conversion allowed)
template <typename T, typename U>
__global__ void k_scale(T* __restrict__ dst, const U* __restrict__ src, double a, long len)
{
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < len) dst[e] = (T)(a * (double)src[e]);
}

// dst += src (double accumulator)
template <typename U>
__global__ void k_accumulate(double* __restrict__ dst, const U* __restrict__ src, long len)
{
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < len) dst[e] += (double)src[e];
}

// r = b - (x - Jx) with Jx = (r1 - r0) s   (outer refinement residual, FP64)
__global__ void k_outer_residual(double* __restrict__ r, const double* __restrict__ b, const double* __restrict__ x,
                                 const double* __restrict__ r1, const double* __restrict__ r0, double s, long len)
{
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < len) r[e] = b[e] - (x[e] - (r1[e] - r0[e]) * s);
}

static int nb_of(long len) { return (int)((len + KB - 1) / KB); }

int multidot_blocks(long len)
{
    long nb = (len + KB - 1) / KB;
    return (int)(nb < MD_BLOCKS ? nb : MD_BLOCKS);
}

// the update + norm sweep is light on registers: a grid large enough for full occupancy
static int norm_pass_blocks(long len)
{
    long nb = (len + KB - 1) / KB;
    return (int)(nb < 8192 ? nb : 8192);
}

long krylov_partials(long len)
{
    const long a = multidot_blocks(len), b = norm_pass_blocks(len);
    return (a > b ? a : b) * MAXK;
}

template <typename P, typename T>
hipError_t fill_state(P* dst, const double* u, double a, const T* v, int nv, int n, double gamma, hipStream_t s,
                      const double* ap)
{
    hipLaunchKernelGGL((k_fill_state<P, T>), dim3(nb_of(n)), dim3(KB), 0, s, dst, u, a, ap, v, nv, n, gamma - 1.0);
    return hipGetLastError();
}
template <typename T>
hipError_t apply_a(T* w, const T* z, const double* r1, const double* r0, double sc, long len, hipStream_t s,
                   const double* sp)
{
    hipLaunchKernelGGL((k_apply_a<T>), dim3(nb_of(len)), dim3(KB), 0, s, w, z, r1, r0, sc, sp, len);
    return hipGetLastError();
}
hipError_t jvp_step(const double* zz, long len, double sq1u, double* e, hipStream_t s)
{
    hipLaunchKernelGGL(k_jvp_step, dim3(1), dim3(64), 0, s, zz, (double)len, sq1u, e);
    return hipGetLastError();
}
template <typename T>
hipError_t block_apply(int nv, const T* dinv, const T* v, T* z, int n, hipStream_t s)
{
    const int nb = (n + KB - 1) / KB;
    if (nv == 4) hipLaunchKernelGGL((k_block_apply<4, T>), dim3(nb), dim3(KB), 0, s, dinv, v, z, n);
    else hipLaunchKernelGGL((k_block_apply<5, T>), dim3(nb), dim3(KB), 0, s, dinv, v, z, n);
    return hipGetLastError();
}
template <typename T>
hipError_t multidot(const T* w, const T* V, long stride, int m, long len, double* partial, double* out, hipStream_t s)
{
    const int nb = multidot_blocks(len);
    hipLaunchKernelGGL((k_multidot<T>), dim3(nb), dim3(KB), 0, s, w, V, stride, m, len, partial);
    hipLaunchKernelGGL(k_reduce_partials, dim3(m), dim3(KB), 0, s, (const double*)partial, nb, m, out);
    return hipGetLastError();
}
template <typename T>
hipError_t cgs_pass(T* w, const T* V, long stride, const double* h, int m, bool dots, long len, double* partial,
                    double* out, hipStream_t s)
{
    const int nb = dots ? multidot_blocks(len) : norm_pass_blocks(len);
    if (dots && m <= 8)
        hipLaunchKernelGGL((k_cgs_pass<T, true, 8>), dim3(nb), dim3(KB), 0, s, w, V, stride, h, m, len, partial);
    else if (dots && m <= 16)
        hipLaunchKernelGGL((k_cgs_pass<T, true, 16>), dim3(nb), dim3(KB), 0, s, w, V, stride, h, m, len, partial);
    else if (dots && m <= 24)
        hipLaunchKernelGGL((k_cgs_pass<T, true, 24>), dim3(nb), dim3(KB), 0, s, w, V, stride, h, m, len, partial);
    else if (dots)
        hipLaunchKernelGGL((k_cgs_pass<T, true, MAXK>), dim3(nb), dim3(KB), 0, s, w, V, stride, h, m, len, partial);
    else hipLaunchKernelGGL((k_cgs_pass<T, false, MAXK>), dim3(nb), dim3(KB), 0, s, w, V, stride, h, m, len, partial);
    hipLaunchKernelGGL(k_reduce_partials, dim3(dots ? m : 1), dim3(KB), 0, s, (const double*)partial, nb,
                       dots ? m : 1, out);
    return hipGetLastError();
}
template <typename T>
hipError_t lincomb(T* out, const T* V, long stride, const double* y, int m, long len, hipStream_t s)
{
    hipLaunchKernelGGL((k_lincomb<T>), dim3(nb_of(len)), dim3(KB), 0, s, out, V, stride, y, m, len);
    return hipGetLastError();
}
template <typename T, typename U>
hipError_t scale(T* dst, const U* src, double a, long len, hipStream_t s)
{
    hipLaunchKernelGGL((k_scale<T, U>), dim3(nb_of(len)), dim3(KB), 0, s, dst, src, a, len);
    return hipGetLastError();
}
template <typename U>
hipError_t accumulate(double* dst, const U* src, long len, hipStream_t s)
{
    hipLaunchKernelGGL((k_accumulate<U>), dim3(nb_of(len)), dim3(KB), 0, s, dst, src, len);
    return hipGetLastError();
}
hipError_t outer_residual(double* r, const double* b, const double* x, const double* r1, const double* r0, double sc,
                          long len, hipStream_t s)
{
    hipLaunchKernelGGL(k_outer_residual, dim3(nb_of(len)), dim3(KB), 0, s, r, b, x, r1, r0, sc, len);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// the reference's assembled implicit matrix (bdlu_yeqax, SPMAT_BDLU.f90:139-163) and the vector
// operations of its GMRES (bdlu_gmres.f90); vectors [nv][N], N = n interior + ghost unknowns

// y = A x: interior rows D_c x_c + sum over slots B_k x_tgt(k); ghost rows are the identity
template <int NV>
__global__ __launch_bounds__(KB) void k_bdlu_apply(const double* __restrict__ D, const double* __restrict__ B,
                                                   const int* __restrict__ tgt, int ns, int n, long N,
                                                   const double* __restrict__ x, double* __restrict__ y)
{
    const long c = (long)blockIdx.x * KB + threadIdx.x;
    if (c >= N) return;
    if (c >= n) {
#pragma unroll
        for (int v = 0; v < NV; ++v) y[(size_t)v * N + c] = x[(size_t)v * N + c];
        return;
    }
    double xc[NV], acc[NV];
#pragma unroll
    for (int j = 0; j < NV; ++j) xc[j] = x[(size_t)j * N + c];
#pragma unroll
    for (int r = 0; r < NV; ++r) {
        double sv = 0.0;
#pragma unroll
        for (int j = 0; j < NV; ++j) sv += D[(size_t)(r * NV + j) * n + c] * xc[j];
        acc[r] = sv;
    }
    for (int k = 0; k < ns; ++k) {
        const int t = tgt[(size_t)k * n + c];
        if (t < 0) continue;
        double xt[NV];
#pragma unroll
        for (int j = 0; j < NV; ++j) xt[j] = x[(size_t)j * N + t];
        const double* Bk = B + (size_t)k * NV * NV * n + c;
#pragma unroll
        for (int r = 0; r < NV; ++r) {
            double sv = 0.0;
#pragma unroll
            for (int j = 0; j < NV; ++j) sv += Bk[(size_t)(r * NV + j) * n] * xt[j];
            acc[r] += sv;
        }
    }
#pragma unroll
    for (int v = 0; v < NV; ++v) y[(size_t)v * N + c] = acc[v];
}

// x0 = rhs / diag(ib, ib) (bdlu_gmres.f90:75-81); ghost rows have a unit diagonal
__global__ void k_bdlu_x0(const double* __restrict__ D, const double* __restrict__ p, double* __restrict__ x,
                          int nv, int n, long N)
{
    const long i = (long)blockIdx.x * KB + threadIdx.x;
    if (i >= (long)nv * N) return;
    const int v = (int)(i / N);
    const long c = i - (long)v * N;
    x[i] = c < n ? p[i] / D[(size_t)(v * nv + v) * n + c] : p[i];
}

__global__ void k_axpy(double* __restrict__ y, const double* __restrict__ x, double a, long len)
{
    const long i = (long)blockIdx.x * KB + threadIdx.x;
    if (i < len) y[i] = y[i] + a * x[i];
}

// r = p - r
__global__ void k_pmr(double* __restrict__ r, const double* __restrict__ p, long len)
{
    const long i = (long)blockIdx.x * KB + threadIdx.x;
    if (i < len) r[i] = p[i] - r[i];
}

// per-block partial sums of |x| (fixed order), reduced by k_reduce_partials
__global__ __launch_bounds__(KB) void k_asum(const double* __restrict__ x, long len, double* __restrict__ partial)
{
    __shared__ double red[KB / 64];
    double sv = 0.0;
    for (long e = (long)blockIdx.x * KB + threadIdx.x; e < len; e += (long)gridDim.x * KB) sv += fabs(x[e]);
    sv = wsum(sv);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = sv;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = red[0];
        for (int k = 1; k < KB / 64; ++k) t += red[k];
        partial[(long)blockIdx.x * MAXK] = t;
    }
}

hipError_t bdlu_apply(int nv, const double* D, const double* B, const int* tgt, int ns, int n, long N, const double* x,
                      double* y, hipStream_t s)
{
    const dim3 grid((unsigned)((N + KB - 1) / KB));
    if (nv == 4) hipLaunchKernelGGL((k_bdlu_apply<4>), grid, dim3(KB), 0, s, D, B, tgt, ns, n, N, x, y);
    else hipLaunchKernelGGL((k_bdlu_apply<5>), grid, dim3(KB), 0, s, D, B, tgt, ns, n, N, x, y);
    return hipGetLastError();
}
hipError_t bdlu_x0(const double* D, const double* p, double* x, int nv, int n, long N, hipStream_t s)
{
    hipLaunchKernelGGL(k_bdlu_x0, dim3(nb_of((long)nv * N)), dim3(KB), 0, s, D, p, x, nv, n, N);
    return hipGetLastError();
}
hipError_t axpy(double* y, const double* x, double a, long len, hipStream_t s)
{
    hipLaunchKernelGGL(k_axpy, dim3(nb_of(len)), dim3(KB), 0, s, y, x, a, len);
    return hipGetLastError();
}
hipError_t pmr(double* r, const double* p, long len, hipStream_t s)
{
    hipLaunchKernelGGL(k_pmr, dim3(nb_of(len)), dim3(KB), 0, s, r, p, len);
    return hipGetLastError();
}
hipError_t asum(const double* x, long len, double* partial, double* out, hipStream_t s)
{
    const int nb = multidot_blocks(len);
    hipLaunchKernelGGL(k_asum, dim3(nb), dim3(KB), 0, s, x, len, partial);
    hipLaunchKernelGGL(k_reduce_partials, dim3(1), dim3(KB), 0, s, (const double*)partial, nb, 1, out);
    return hipGetLastError();
}

#define TY_KINST(T)                                                                                       \
    template hipError_t apply_a<T>(T*, const T*, const double*, const double*, double, long, hipStream_t, const double*); \
    template hipError_t block_apply<T>(int, const T*, const T*, T*, int, hipStream_t);                     \
    template hipError_t multidot<T>(const T*, const T*, long, int, long, double*, double*, hipStream_t);   \
    template hipError_t cgs_pass<T>(T*, const T*, long, const double*, int, bool, long, double*, double*,   \
                                    hipStream_t);                                                          \
    template hipError_t lincomb<T>(T*, const T*, long, const double*, int, long, hipStream_t);             \
    template hipError_t accumulate<T>(double*, const T*, long, hipStream_t);
TY_KINST(float)
TY_KINST(double)
template hipError_t fill_state<float, float>(float*, const double*, double, const float*, int, int, double, hipStream_t, const double*);
template hipError_t fill_state<float, double>(float*, const double*, double, const double*, int, int, double, hipStream_t, const double*);
template hipError_t fill_state<double, double>(double*, const double*, double, const double*, int, int, double, hipStream_t, const double*);
template hipError_t fill_state<double, float>(double*, const double*, double, const float*, int, int, double, hipStream_t, const double*);
template hipError_t scale<float, float>(float*, const float*, double, long, hipStream_t);
template hipError_t scale<float, double>(float*, const double*, double, long, hipStream_t);
template hipError_t scale<double, double>(double*, const double*, double, long, hipStream_t);
template hipError_t scale<double, float>(double*, const float*, double, long, hipStream_t);

}  // namespace ty
