// Mesh preprocessing on the device: the per-slot geometry streams the residual
// kernels read (canonical area vector, LSQ weights, cell-to-face and cell-to-cell
// vectors, Venkatakrishnan epsilon), built from the per-face and per-cell arrays of
// the uploaded mesh.  One work-item per cell; the slot outputs are [slot][3][cell], so
// every store of a wavefront is coalesced.  The arithmetic restates the LSQ setup of
// precalc_grad_lsq.f90:60-155 (A = sum d d^T, or d d^T / |d|^4 for weighted LSQ,
// boundary faces once, 2D A33 += 1) with cholesky_decomp / cholesky_solve
// (SOURCE/MATH/MATRIX.f90:42-111) folded into per-slot weights w_k = A^-1 d_k.
#include "ty_internal.h"

namespace ty {

template <typename R>
__global__ __launch_bounds__(BLOCK) void k_geo_slots(GeoSrc s, R* __restrict__ out)
{
    const int i = blockIdx.x * BLOCK + threadIdx.x;
    const int n = s.n;
    if (i >= n) return;
    const size_t S3 = (size_t)3 * s.ns * n;
    R* ss = out;
    R* sw = out + S3;
    R* sr = out + 2 * S3;
    R* sd = out + 3 * S3;
    R* eps = out + 4 * S3 + (size_t)4 * s.nf;
    R* sh = eps + n;
    auto at = [&](R* base, int k, int e) -> R& { return base[((size_t)k * 3 + e) * n + i]; };
    const double x[3] = {s.cc[i], s.cc[(size_t)n + i], s.cc[(size_t)2 * n + i]};
    double A[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    double dl[MAX_SLOTS][3];
#pragma unroll
    for (int k = 0; k < MAX_SLOTS; ++k) {
        if (k >= s.ns) break;
        const int wd = s.fw[(size_t)k * n + i];
        if (wd == SLOT_UNUSED) continue;
        const int nbv = s.nbr[(size_t)k * n + i];
        const int f = (int)((unsigned)wd & FACE_MASK);
        const bool other = wd < 0;
        double T[3] = {0.0, 0.0, 0.0};
        if ((unsigned)wd & PERIODIC_BIT) {
            const int j = f - s.per_base;
            for (int e = 0; e < 3; ++e) T[e] = s.ps[(size_t)e * s.nper + j];
        }
        double d[3], r[3];
        for (int e = 0; e < 3; ++e) {
            const double xf = s.fc[(size_t)e * s.nf + f];
            if (nbv >= 0) {
                const double xo = s.cc[(size_t)e * n + nbv];
                d[e] = other ? (x[e] + T[e]) - xo : (xo + T[e]) - x[e];
                r[e] = other ? (xf - T[e]) - x[e] : xf - x[e];
            } else {
                d[e] = s.gc[(size_t)e * s.nb + (-nbv - 1)] - x[e];
                r[e] = xf - x[e];
            }
            at(ss, k, e) = (R)(s.fn[(size_t)e * s.nf + f] * s.fs[f]);
            at(sr, k, e) = (R)r[e];
            at(sd, k, e) = (R)d[e];
        }
        {
            // viscous distance weight from the face's canonical sides, the owner's frame
            // (calc_flux_viscous_cell.f90:85-95): both cells of a face store the same value
            double rl2 = 0.0, rr2 = 0.0;
            for (int e = 0; e < 3; ++e) {
                const double xf = s.fc[(size_t)e * s.nf + f];
                double xl, xr;
                if (nbv < 0) {
                    xl = x[e];
                    xr = s.gc[(size_t)e * s.nb + (-nbv - 1)];
                } else if (other) {
                    xl = s.cc[(size_t)e * n + nbv];
                    xr = x[e] + T[e];
                } else {
                    xl = x[e];
                    xr = s.cc[(size_t)e * n + nbv] + T[e];
                }
                const double rl = xf - xl, rr = xf - xr;
                rl2 += rl * rl;
                rr2 += rr * rr;
            }
            const double hl = sqrt(rl2), hr = sqrt(rr2);
            sh[(size_t)k * n + i] = (R)(hl / (hl + hr));
        }
        const double s2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
        double m[3];
        for (int e = 0; e < 3; ++e) m[e] = s.wlsq ? d[e] / s2 : d[e];
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b) A[a + 3 * b] += m[a] * m[b];
        // right-hand-side direction x_nb - x_c in this cell's frame
        for (int e = 0; e < 3; ++e) dl[k][e] = (other ? -d[e] : d[e]) / (s.wlsq ? s2 * s2 : 1.0);
    }
    if (s.dim == 2) A[8] += 1.0;
    // cholesky_decomp, lower factor in place (column-major)
    for (int a = 0; a < 3; ++a)
        for (int b = a; b < 3; ++b) {
            double sum = A[a + 3 * b];
            for (int k = a - 1; k >= 0; --k) sum = sum - A[a + 3 * k] * A[b + 3 * k];
            if (a == b) A[a + 3 * a] = sqrt(sum);
            else A[b + 3 * a] = sum / A[a + 3 * a];
        }
    const double l11 = A[0], l21 = A[1], l31 = A[2], l22 = A[4], l32 = A[5], l33 = A[8];
#pragma unroll
    for (int k = 0; k < MAX_SLOTS; ++k) {
        if (k >= s.ns) break;
        if (s.fw[(size_t)k * n + i] == SLOT_UNUSED) continue;
        const double* b = dl[k];
        const double y1 = b[0] / l11;
        const double y2 = (b[1] - l21 * y1) / l22;
        const double y3 = ((b[2] - l31 * y1) - l32 * y2) / l33;
        const double x3 = y3 / l33;
        const double x2 = (y2 - l32 * x3) / l22;
        const double x1 = ((y1 - l21 * x2) - l31 * x3) / l11;
        at(sw, k, 0) = (R)x1;
        at(sw, k, 1) = (R)x2;
        at(sw, k, 2) = (R)x3;
    }
    const double v = s.vol[i];
    const double h3 = s.dim == 3 ? v : pow(v, 1.5);
    eps[i] = (R)(s.K * s.K * s.K * h3);
}

// per-face normal and area in the arithmetic precision ([3][nf] then [nf])
template <typename R>
__global__ __launch_bounds__(BLOCK) void k_geo_faces(GeoSrc s, R* __restrict__ out)
{
    const size_t S3 = (size_t)3 * s.ns * s.n;
    R* fn = out + 4 * S3;
    R* fs = fn + (size_t)3 * s.nf;
    for (size_t f = (size_t)blockIdx.x * BLOCK + threadIdx.x; f < (size_t)s.nf; f += (size_t)gridDim.x * BLOCK) {
        for (int e = 0; e < 3; ++e) fn[(size_t)e * s.nf + f] = (R)s.fn[(size_t)e * s.nf + f];
        fs[f] = (R)s.fs[f];
    }
}

template <typename R>
hipError_t launch_geo(const GeoSrc& s, R* out, hipStream_t st)
{
    // a slot-array layout padded to whole slots: zero the unused slots first
    const size_t S3 = (size_t)3 * s.ns * s.n;
    hipError_t e = hipMemsetAsync(out, 0, 4 * S3 * sizeof(R), st);
    if (e != hipSuccess) return e;
    e = hipMemsetAsync(out + 4 * S3 + (size_t)4 * s.nf + s.n, 0, (size_t)s.ns * s.n * sizeof(R), st);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL((k_geo_slots<R>), dim3((s.n + BLOCK - 1) / BLOCK), dim3(BLOCK), 0, st, s, out);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    const int nblk = (int)std::min<size_t>(((size_t)s.nf + BLOCK - 1) / BLOCK, 65536);
    hipLaunchKernelGGL((k_geo_faces<R>), dim3(std::max(1, nblk)), dim3(BLOCK), 0, st, s, out);
    return hipGetLastError();
}

// link word bits in the fourth element of a packed record
__device__ __forceinline__ float link_as(float, int L) { return __int_as_float(L); }
__device__ __forceinline__ double link_as(double, int L) { return __longlong_as_double((long long)(unsigned)L); }

__device__ __forceinline__ void st_pack(float4* p, size_t i, size_t, float x, float y, float z, float w)
{
    p[i] = make_float4(x, y, z, w);
}
__device__ __forceinline__ void st_pack(double2* p, size_t i, size_t plane, double x, double y, double z, double w)
{
    p[i] = make_double2(x, y);
    p[plane + i] = make_double2(z, w);
}

template <typename R>
__global__ __launch_bounds__(BLOCK) void k_geo_pack(int n, int ns, Geo<R> g, const int* __restrict__ link,
                                                   typename Pack4<R>::S* __restrict__ out)
{
    using P = typename Pack4<R>::S;
    const int c = blockIdx.x * BLOCK + threadIdx.x;
    if (c >= n) return;
    // FP32: pw | pa | pb; FP64: pw only
    const size_t plane = (size_t)ns * n, arr = Pack4<R>::planes * plane;
    P* pw = out;
    P* pa = out + arr;
    P* pb = out + 2 * arr;
    for (int k = 0; k < ns; ++k) {
        const size_t i = (size_t)k * n + c;
        const size_t v = (size_t)k * 3 * n + c;
        const R lb = link_as(R(0), link[i]);
        st_pack(pw, i, plane, g.sw[v], g.sw[v + n], g.sw[v + 2 * (size_t)n], lb);
        if (sizeof(R) == 4) {
            st_pack(pa, i, plane, g.ss[v], g.ss[v + n], g.ss[v + 2 * (size_t)n], g.sh[i]);
            st_pack(pb, i, plane, g.sd[v], g.sd[v + n], g.sd[v + 2 * (size_t)n], lb);
        }
    }
}

template <typename R>
hipError_t launch_geo_pack(int n, int ns, const Geo<R>& g, const int* link, typename Pack4<R>::S* out, hipStream_t st)
{
    hipLaunchKernelGGL(k_geo_pack<R>, dim3((n + BLOCK - 1) / BLOCK), dim3(BLOCK), 0, st, n, ns, g, link, out);
    return hipGetLastError();
}

template hipError_t launch_geo_pack<float>(int, int, const Geo<float>&, const int*, float4*, hipStream_t);
template hipError_t launch_geo_pack<double>(int, int, const Geo<double>&, const int*, double2*, hipStream_t);

template hipError_t launch_geo<float>(const GeoSrc&, float*, hipStream_t);
template hipError_t launch_geo<double>(const GeoSrc&, double*, hipStream_t);

}  // namespace ty

// ---------------------------------------------------------------------------
// Stable ascending sort of n 64-bit keys on the device (LSD radix sort over bits
// [0, end_bit)); idx_out[r] receives the input position of the r-th smallest key, ties
// in input order -- the order std::stable_sort gives on (key, position).
#include <hipcub/hipcub.hpp>

namespace ty {

namespace {
__global__ void k_iota(int* v, int n)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) v[i] = i;
}
}  // namespace

hipError_t sort_keys_device(const uint64_t* keys, int n, int end_bit, int* idx_out, hipStream_t st)
{
    if (n <= 0) return hipSuccess;
    uint64_t *k0 = nullptr, *k1 = nullptr;
    int *v0 = nullptr, *v1 = nullptr;
    void* tmp = nullptr;
    size_t tb = 0;
    hipError_t e = hipSuccess;
    auto done = [&](hipError_t r) {
        // hipFree waits for the device; a failing free must not hide the call's own result
        (void)hipFree(k0);
        (void)hipFree(k1);
        (void)hipFree(v0);
        (void)hipFree(v1);
        (void)hipFree(tmp);
        return r;
    };
    if ((e = hipMalloc(&k0, (size_t)n * 8)) != hipSuccess || (e = hipMalloc(&k1, (size_t)n * 8)) != hipSuccess ||
        (e = hipMalloc(&v0, (size_t)n * 4)) != hipSuccess || (e = hipMalloc(&v1, (size_t)n * 4)) != hipSuccess)
        return done(e);
    if ((e = hipMemcpyAsync(k0, keys, (size_t)n * 8, hipMemcpyHostToDevice, st)) != hipSuccess) return done(e);
    hipLaunchKernelGGL(k_iota, dim3(std::min(65535, (n + 255) / 256)), dim3(256), 0, st, v0, n);
    if ((e = hipGetLastError()) != hipSuccess) return done(e);
    if ((e = hipcub::DeviceRadixSort::SortPairs(nullptr, tb, k0, k1, v0, v1, n, 0, end_bit, st)) != hipSuccess)
        return done(e);
    if ((e = hipMalloc(&tmp, std::max<size_t>(tb, 1))) != hipSuccess) return done(e);
    if ((e = hipcub::DeviceRadixSort::SortPairs(tmp, tb, k0, k1, v0, v1, n, 0, end_bit, st)) != hipSuccess)
        return done(e);
    if ((e = hipMemcpyAsync(idx_out, v1, (size_t)n * 4, hipMemcpyDeviceToHost, st)) != hipSuccess) return done(e);
    return done(hipStreamSynchronize(st));
}

}  // namespace ty
