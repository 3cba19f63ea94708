// Dual numbers (value, one directional derivative) for the tangent-linear and
// adjoint residual kernels.  The flux, boundary-state and primitive-variable
// routines are templates over the arithmetic type, so instantiating them with
// Dual propagates an exact derivative through the very operations the primal
// kernels perform: forward-mode differentiation by operator overloading.
// Comparisons act on values, so branches (wave-speed selection, upwinding,
// limiter cases) follow the primal and the derivative is that of the active
// branch.
#pragma once
#include <hip/hip_runtime.h>

namespace ty {

struct Dual {
    double v, d;
    Dual() = default;
    __host__ __device__ constexpr Dual(double x) : v(x), d(0.0) {}
    __host__ __device__ constexpr Dual(double x, double dx) : v(x), d(dx) {}
};

__device__ __forceinline__ Dual operator+(Dual a, Dual b) { return Dual(a.v + b.v, a.d + b.d); }
__device__ __forceinline__ Dual operator-(Dual a, Dual b) { return Dual(a.v - b.v, a.d - b.d); }
__device__ __forceinline__ Dual operator-(Dual a) { return Dual(-a.v, -a.d); }
__device__ __forceinline__ Dual operator*(Dual a, Dual b) { return Dual(a.v * b.v, a.d * b.v + a.v * b.d); }
__device__ __forceinline__ Dual operator/(Dual a, Dual b)
{
    const double q = a.v / b.v;
    return Dual(q, (a.d - q * b.d) / b.v);
}
__device__ __forceinline__ Dual& operator+=(Dual& a, Dual b) { return a = a + b; }
__device__ __forceinline__ Dual& operator-=(Dual& a, Dual b) { return a = a - b; }
__device__ __forceinline__ Dual& operator*=(Dual& a, Dual b) { return a = a * b; }
__device__ __forceinline__ bool operator<(Dual a, Dual b) { return a.v < b.v; }
__device__ __forceinline__ bool operator>(Dual a, Dual b) { return a.v > b.v; }
__device__ __forceinline__ bool operator<=(Dual a, Dual b) { return a.v <= b.v; }
__device__ __forceinline__ bool operator>=(Dual a, Dual b) { return a.v >= b.v; }
__device__ __forceinline__ bool operator==(Dual a, Dual b) { return a.v == b.v; }

__device__ __forceinline__ Dual sqrt(Dual a)
{
    const double s = ::sqrt(a.v);
    return Dual(s, s > 0.0 ? a.d / (2.0 * s) : 0.0);
}
__device__ __forceinline__ Dual fabs(Dual a) { return a.v < 0.0 ? -a : a; }
__device__ __forceinline__ Dual copysign(Dual a, Dual b)
{
    const bool flip = (a.v < 0.0) != (b.v < 0.0);
    return flip ? -a : a;
}
__device__ __forceinline__ Dual exp(Dual a)
{
    const double e = ::exp(a.v);
    return Dual(e, e * a.d);
}
// d erf(x) / dx = 2 / sqrt(pi) exp(-x^2)
__device__ __forceinline__ Dual erf(Dual a)
{
    return Dual(::erf(a.v), 1.1283791670955125739 * ::exp(-a.v * a.v) * a.d);
}
// a^b with both operands carrying derivatives
__device__ __forceinline__ Dual dpow(Dual a, Dual b)
{
    const double p = ::pow(a.v, b.v);
    double d = b.v * ::pow(a.v, b.v - 1.0) * a.d;
    if (b.d != 0.0) d += p * ::log(a.v) * b.d;
    return Dual(p, d);
}

}  // namespace ty
