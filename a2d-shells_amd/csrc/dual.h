// dual.h -- forward-mode dual numbers (value, one directional derivative)
// for the coordinate-derivative kernels.  Arithmetic follows the product,
// quotient and chain rules; comparisons are made on values by the callers.
#pragma once

#include <hip/hip_runtime.h>

namespace shellgpu {

using ::sqrt;  // keep the double overloads visible beside sqrt(Dual)

struct Dual {
  double v, d;
  __device__ Dual() = default;
  __device__ constexpr Dual(double x) : v(x), d(0.0) {}
  __device__ constexpr Dual(double x, double dx) : v(x), d(dx) {}
};

__device__ __forceinline__ Dual operator+(Dual a, Dual b) { return Dual(a.v + b.v, a.d + b.d); }
__device__ __forceinline__ Dual operator-(Dual a, Dual b) { return Dual(a.v - b.v, a.d - b.d); }
__device__ __forceinline__ Dual operator-(Dual a) { return Dual(-a.v, -a.d); }
__device__ __forceinline__ Dual operator*(Dual a, Dual b) { return Dual(a.v * b.v, a.d * b.v + a.v * b.d); }
__device__ __forceinline__ Dual operator/(Dual a, Dual b) {
  const double r = 1.0 / b.v;
  const double q = a.v * r;
  return Dual(q, (a.d - q * b.d) * r);
}
__device__ __forceinline__ Dual operator+(Dual a, double b) { return Dual(a.v + b, a.d); }
__device__ __forceinline__ Dual operator+(double a, Dual b) { return Dual(a + b.v, b.d); }
__device__ __forceinline__ Dual operator-(Dual a, double b) { return Dual(a.v - b, a.d); }
__device__ __forceinline__ Dual operator-(double a, Dual b) { return Dual(a - b.v, -b.d); }
__device__ __forceinline__ Dual operator*(Dual a, double b) { return Dual(a.v * b, a.d * b); }
__device__ __forceinline__ Dual operator*(double a, Dual b) { return Dual(a * b.v, a * b.d); }
__device__ __forceinline__ Dual operator/(double a, Dual b) {
  const double r = 1.0 / b.v;
  const double q = a * r;
  return Dual(q, -q * b.d * r);
}
__device__ __forceinline__ Dual &operator+=(Dual &a, Dual b) {
  a.v += b.v;
  a.d += b.d;
  return a;
}
__device__ __forceinline__ Dual sqrt(Dual a) {
  const double s = ::sqrt(a.v);
  return Dual(s, 0.5 * a.d / s);
}
__device__ __forceinline__ double value(Dual a) { return a.v; }
__device__ __forceinline__ double value(double a) { return a; }

}  // namespace shellgpu
