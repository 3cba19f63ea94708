// Launchers of the Krylov vector kernels (ty_krylov.hip).
#pragma once
#include <hip/hip_runtime.h>

namespace ty {

constexpr int MAXK = 32;  // Krylov basis size limit (GMRES restart length + 1)

int multidot_blocks(long len);
long krylov_partials(long len);  // doubles of partial-sum space the reductions below need
// stage-state records dst[c][v] = (P)(u[v][c] + a v[v][c]) for the residual kernels (ap: a on the device)
template <typename P, typename T>
hipError_t fill_state(P* dst, const double* u, double a, const T* v, int nv, int n, double gamma, hipStream_t s,
                      const double* ap = nullptr);
// w = z - (r1 - r0) sc (sp: sc on the device)
template <typename T>
hipError_t apply_a(T* w, const T* z, const double* r1, const double* r0, double sc, long len, hipStream_t s,
                   const double* sp = nullptr);
// e[0] = sq1u / sqrt(*zz / len) and e[1] = 1 / e[0] on the device (0 and 0 for *zz = 0)
hipError_t jvp_step(const double* zz, long len, double sq1u, double* e, hipStream_t s);
template <typename T>
hipError_t block_apply(int nv, const T* dinv, const T* v, T* z, int n, hipStream_t s);
template <typename T>
hipError_t multidot(const T* w, const T* V, long stride, int m, long len, double* partial, double* out, hipStream_t s);
// w -= V h, then (dots) out[i] = V_i . w for i < m, or out[0] = w . w: one sweep over the basis
template <typename T>
hipError_t cgs_pass(T* w, const T* V, long stride, const double* h, int m, bool dots, long len, double* partial,
                    double* out, hipStream_t s);
template <typename T>
hipError_t lincomb(T* out, const T* V, long stride, const double* y, int m, long len, hipStream_t s);
template <typename T, typename U>
hipError_t scale(T* dst, const U* src, double a, long len, hipStream_t s);
template <typename U>
hipError_t accumulate(double* dst, const U* src, long len, hipStream_t s);
hipError_t outer_residual(double* r, const double* b, const double* x, const double* r1, const double* r0, double sc,
                          long len, hipStream_t s);
// the reference's assembled matrix: y = A x over [nv][N] vectors (N = n interior + ghost unknowns),
// diagonal blocks D [nv*nv][n], per-slot blocks B [ns][nv*nv][n] and their unknowns tgt [ns][n]
hipError_t bdlu_apply(int nv, const double* D, const double* B, const int* tgt, int ns, int n, long N, const double* x,
                      double* y, hipStream_t s);
hipError_t bdlu_x0(const double* D, const double* p, double* x, int nv, int n, long N, hipStream_t s);  // x = p/diag
hipError_t axpy(double* y, const double* x, double a, long len, hipStream_t s);                          // y += a x
hipError_t pmr(double* r, const double* p, long len, hipStream_t s);                                     // r = p - r
hipError_t asum(const double* x, long len, double* partial, double* out, hipStream_t s);                 // sum |x|

}  // namespace ty
