// dual.hpp -- forward-mode dual numbers for the adjoint and tangent kernels.
//
// Dual<N> carries a value and N directional derivatives.  The leaf physics
// (physics.hpp) is templated on its scalar type, so evaluating it on
// Dual<4> seeded with the identity yields the 4x4 Jacobian of a flux with
// respect to the q~ it starts from, and on Dual<1> a tangent.  Each
// operation applies its textbook derivative rule; exp, log, sqrt and erf
// call the device library for the value.
#pragma once

#include <hip/hip_runtime.h>

namespace lskum {

// the double overloads stay visible beside the Dual ones declared here
using ::erf;
using ::exp;
using ::fabs;
using ::log;
using ::sqrt;

template <int N>
struct Dual {
  double v;
  double d[N];
  __device__ Dual() : v(0.0) {
    for (int j = 0; j < N; j++) d[j] = 0.0;
  }
  __device__ Dual(double x) : v(x) {  // a constant
    for (int j = 0; j < N; j++) d[j] = 0.0;
  }
};

template <int N>
__device__ __forceinline__ Dual<N> operator+(const Dual<N> &a, const Dual<N> &b) {
  Dual<N> r;
  r.v = a.v + b.v;
  for (int j = 0; j < N; j++) r.d[j] = a.d[j] + b.d[j];
  return r;
}
template <int N>
__device__ __forceinline__ Dual<N> operator-(const Dual<N> &a, const Dual<N> &b) {
  Dual<N> r;
  r.v = a.v - b.v;
  for (int j = 0; j < N; j++) r.d[j] = a.d[j] - b.d[j];
  return r;
}
template <int N>
__device__ __forceinline__ Dual<N> operator-(const Dual<N> &a) {
  Dual<N> r;
  r.v = -a.v;
  for (int j = 0; j < N; j++) r.d[j] = -a.d[j];
  return r;
}
template <int N>
__device__ __forceinline__ Dual<N> operator*(const Dual<N> &a, const Dual<N> &b) {
  Dual<N> r;
  r.v = a.v * b.v;
  for (int j = 0; j < N; j++) r.d[j] = a.d[j] * b.v + a.v * b.d[j];
  return r;
}
template <int N>
__device__ __forceinline__ Dual<N> operator*(const Dual<N> &a, double b) {
  Dual<N> r;
  r.v = a.v * b;
  for (int j = 0; j < N; j++) r.d[j] = a.d[j] * b;
  return r;
}
template <int N>
__device__ __forceinline__ Dual<N> operator/(const Dual<N> &a, const Dual<N> &b) {
  Dual<N> r;
  const double inv = 1.0 / b.v;
  r.v = a.v * inv;
  for (int j = 0; j < N; j++) r.d[j] = (a.d[j] - r.v * b.d[j]) * inv;
  return r;
}
template <int N>
__device__ __forceinline__ Dual<N> exp(const Dual<N> &a) {
  Dual<N> r;
  r.v = ::exp(a.v);
  for (int j = 0; j < N; j++) r.d[j] = r.v * a.d[j];
  return r;
}
template <int N>
__device__ __forceinline__ Dual<N> log(const Dual<N> &a) {
  Dual<N> r;
  r.v = ::log(a.v);
  const double inv = 1.0 / a.v;
  for (int j = 0; j < N; j++) r.d[j] = a.d[j] * inv;
  return r;
}
template <int N>
__device__ __forceinline__ Dual<N> sqrt(const Dual<N> &a) {
  Dual<N> r;
  r.v = ::sqrt(a.v);
  const double h = 0.5 / r.v;
  for (int j = 0; j < N; j++) r.d[j] = a.d[j] * h;
  return r;
}
template <int N>
__device__ __forceinline__ Dual<N> erf(const Dual<N> &a) {
  // d erf(x)/dx = 2/sqrt(pi) exp(-x^2)
  Dual<N> r;
  r.v = ::erf(a.v);
  const double g = 1.1283791670955126 * ::exp(-a.v * a.v);
  for (int j = 0; j < N; j++) r.d[j] = g * a.d[j];
  return r;
}

// Comparisons look at the values: a branch (a limiter switch, a min or max
// changing hands) is taken as the state takes it, and the derivative is that
// of the branch taken.
template <int N>
__device__ __forceinline__ bool operator<(const Dual<N> &a, const Dual<N> &b) {
  return a.v < b.v;
}
template <int N>
__device__ __forceinline__ bool operator>(const Dual<N> &a, const Dual<N> &b) {
  return a.v > b.v;
}
template <int N>
__device__ __forceinline__ bool operator<=(const Dual<N> &a, const Dual<N> &b) {
  return a.v <= b.v;
}
template <int N>
__device__ __forceinline__ bool operator!=(const Dual<N> &a, const Dual<N> &b) {
  return a.v != b.v;
}
template <int N>
__device__ __forceinline__ Dual<N> fabs(const Dual<N> &a) {
  return a.v < 0.0 ? -a : a;
}

// the value of a scalar of either kind
__device__ __forceinline__ double vpart(double a) { return a; }
template <int N>
__device__ __forceinline__ double vpart(const Dual<N> &a) {
  return a.v;
}

}  // namespace lskum
