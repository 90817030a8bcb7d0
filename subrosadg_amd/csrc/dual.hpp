// dual.hpp -- forward-mode dual numbers with N tangent directions, for the pointwise Jacobians of
// the discrete adjoint. The pointwise physics (physics.hpp) is templated on its scalar type, so
// evaluating it on Dual<N> yields the derivatives of every output with respect to N seeded inputs
// in one pass, with the same branches (Riemann-solver cases, entropy fix, max) as the FP64 value.
#ifndef SDG_DUAL_HPP_
#define SDG_DUAL_HPP_

#include <hip/hip_runtime.h>

namespace sdg {

// keep the double overloads visible next to the Dual ones declared in this namespace
using ::exp;
using ::fabs;
using ::fmax;
using ::log;
using ::sqrt;

template <int N>
struct Dual {
  double v;
  double d[N];
  Dual() = default;
  __host__ __device__ Dual(double x) : v(x) {  // NOLINT: constants promote implicitly
#pragma unroll
    for (int i = 0; i < N; i++) d[i] = 0.0;
  }
};

template <int N>
__device__ __forceinline__ Dual<N> seed(double x, int k) {
  Dual<N> r(x);
  r.d[k] = 1.0;
  return r;
}

#define SDG_DUAL_BINARY(OP, VAL, DER)                                                   \
  template <int N>                                                                      \
  __device__ __forceinline__ Dual<N> operator OP(const Dual<N>& a, const Dual<N>& b) { \
    Dual<N> r;                                                                          \
    r.v = VAL;                                                                          \
    _Pragma("unroll") for (int i = 0; i < N; i++) r.d[i] = DER;                        \
    return r;                                                                           \
  }
SDG_DUAL_BINARY(+, a.v + b.v, a.d[i] + b.d[i])
SDG_DUAL_BINARY(-, a.v - b.v, a.d[i] - b.d[i])
SDG_DUAL_BINARY(*, a.v * b.v, a.d[i] * b.v + a.v * b.d[i])
#undef SDG_DUAL_BINARY

template <int N>
__device__ __forceinline__ Dual<N> operator/(const Dual<N>& a, const Dual<N>& b) {
  Dual<N> r;
  r.v = a.v / b.v;
  const double ib = 1.0 / b.v;
#pragma unroll
  for (int i = 0; i < N; i++) r.d[i] = (a.d[i] - r.v * b.d[i]) * ib;
  return r;
}
template <int N>
__device__ __forceinline__ Dual<N> operator-(const Dual<N>& a) {
  Dual<N> r;
  r.v = -a.v;
#pragma unroll
  for (int i = 0; i < N; i++) r.d[i] = -a.d[i];
  return r;
}
template <int N>
__device__ __forceinline__ Dual<N>& operator+=(Dual<N>& a, const Dual<N>& b) {
  a = a + b;
  return a;
}
template <int N>
__device__ __forceinline__ Dual<N>& operator-=(Dual<N>& a, const Dual<N>& b) {
  a = a - b;
  return a;
}
template <int N>
__device__ __forceinline__ Dual<N>& operator*=(Dual<N>& a, const Dual<N>& b) {
  a = a * b;
  return a;
}

#define SDG_DUAL_CMP(OP) \
  template <int N>       \
  __device__ __forceinline__ bool operator OP(const Dual<N>& a, const Dual<N>& b) { return a.v OP b.v; }
SDG_DUAL_CMP(<)
SDG_DUAL_CMP(<=)
SDG_DUAL_CMP(>)
SDG_DUAL_CMP(>=)
#undef SDG_DUAL_CMP

template <int N>
__device__ __forceinline__ Dual<N> sqrt(const Dual<N>& a) {
  Dual<N> r;
  r.v = ::sqrt(a.v);
  const double h = 0.5 / r.v;
#pragma unroll
  for (int i = 0; i < N; i++) r.d[i] = a.d[i] * h;
  return r;
}
template <int N>
__device__ __forceinline__ Dual<N> fabs(const Dual<N>& a) {
  return a.v < 0.0 ? -a : a;
}
template <int N>
__device__ __forceinline__ Dual<N> fmax(const Dual<N>& a, const Dual<N>& b) {
  return a.v >= b.v ? a : b;
}
template <int N>
__device__ __forceinline__ Dual<N> exp(const Dual<N>& a) {
  Dual<N> r;
  r.v = ::exp(a.v);
#pragma unroll
  for (int i = 0; i < N; i++) r.d[i] = a.d[i] * r.v;
  return r;
}
template <int N>
__device__ __forceinline__ Dual<N> log(const Dual<N>& a) {
  Dual<N> r;
  r.v = ::log(a.v);
  const double ia = 1.0 / a.v;
#pragma unroll
  for (int i = 0; i < N; i++) r.d[i] = a.d[i] * ia;
  return r;
}

}  // namespace sdg

#endif  // SDG_DUAL_HPP_
