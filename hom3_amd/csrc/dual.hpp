// First-order forward-mode number for the tangent and for local Jacobians
// of the adjoint.  Value and one directional derivative; comparisons and
// branches follow the value, so a piecewise formula is differentiated on the
// branch the value takes.
#pragma once
#include "physics.hpp"

namespace hom3g {

struct Dual {
  double v, d;
  Dual() = default;
  __host__ __device__ constexpr explicit Dual(double x) : v(x), d(0) {}
  __host__ __device__ constexpr Dual(double x, double y) : v(x), d(y) {}
  __host__ __device__ Dual& operator+=(const Dual& b) { v += b.v; d += b.d; return *this; }
  __host__ __device__ Dual& operator-=(const Dual& b) { v -= b.v; d -= b.d; return *this; }
  __host__ __device__ Dual& operator*=(const Dual& b) { d = d * b.v + v * b.d; v *= b.v; return *this; }
};

__host__ __device__ inline Dual operator+(Dual a, Dual b) { return Dual(a.v + b.v, a.d + b.d); }
__host__ __device__ inline Dual operator-(Dual a, Dual b) { return Dual(a.v - b.v, a.d - b.d); }
__host__ __device__ inline Dual operator*(Dual a, Dual b) { return Dual(a.v * b.v, a.d * b.v + a.v * b.d); }
__host__ __device__ inline Dual operator/(Dual a, Dual b) {
  const double q = a.v / b.v;
  return Dual(q, (a.d - q * b.d) / b.v);
}
__host__ __device__ inline Dual operator-(Dual a) { return Dual(-a.v, -a.d); }
__host__ __device__ inline Dual operator+(Dual a, double b) { return Dual(a.v + b, a.d); }
__host__ __device__ inline Dual operator+(double a, Dual b) { return Dual(a + b.v, b.d); }
__host__ __device__ inline Dual operator-(Dual a, double b) { return Dual(a.v - b, a.d); }
__host__ __device__ inline Dual operator-(double a, Dual b) { return Dual(a - b.v, -b.d); }
__host__ __device__ inline Dual operator*(Dual a, double b) { return Dual(a.v * b, a.d * b); }
__host__ __device__ inline Dual operator*(double a, Dual b) { return Dual(a * b.v, a * b.d); }
__host__ __device__ inline Dual operator/(Dual a, double b) { return Dual(a.v / b, a.d / b); }
__host__ __device__ inline Dual operator/(double a, Dual b) {
  const double q = a / b.v;
  return Dual(q, -q * b.d / b.v);
}
__host__ __device__ inline bool operator<(Dual a, Dual b) { return a.v < b.v; }
__host__ __device__ inline bool operator>(Dual a, Dual b) { return a.v > b.v; }
__host__ __device__ inline bool operator<=(Dual a, Dual b) { return a.v <= b.v; }
__host__ __device__ inline bool operator>=(Dual a, Dual b) { return a.v >= b.v; }
__host__ __device__ inline bool operator<(Dual a, double b) { return a.v < b; }
__host__ __device__ inline bool operator>(Dual a, double b) { return a.v > b; }
__host__ __device__ inline bool operator<=(Dual a, double b) { return a.v <= b; }
__host__ __device__ inline bool operator>=(Dual a, double b) { return a.v >= b; }

template <> __device__ __forceinline__ Dual dev_sqrt<Dual>(Dual x) {
  const double s = __builtin_sqrt(x.v);
  return Dual(s, x.d / (2.0 * s));
}

__device__ __forceinline__ double val(double x) { return x; }
__device__ __forceinline__ double val(Dual x) { return x.v; }

}  // namespace hom3g
