// Forward-mode dual numbers v + d eps (eps^2 = 0) for the shape sensitivities: the element routines
// are templates on their arithmetic type, so running them on Dual<R> with one coordinate seeded
// gives the exact directional derivative of the element residual with respect to that coordinate.
#pragma once
#include "femhot_element.h"

namespace femhot {

template <class T> struct Dual {
  T v, d;
  __host__ __device__ constexpr Dual() : v(T(0)), d(T(0)) {}
  __host__ __device__ constexpr Dual(double x) : v(T(x)), d(T(0)) {}  // constants
  __host__ __device__ constexpr Dual(T x, T dx, int) : v(x), d(dx) {}
  __host__ __device__ Dual &operator+=(const Dual &o) {
    v += o.v;
    d += o.d;
    return *this;
  }
  __host__ __device__ Dual &operator-=(const Dual &o) {
    v -= o.v;
    d -= o.d;
    return *this;
  }
  __host__ __device__ Dual &operator*=(const Dual &o) {
    d = d * o.v + v * o.d;
    v *= o.v;
    return *this;
  }
  __host__ __device__ Dual &operator/=(const Dual &o) {
    const T r = T(1) / o.v;
    d = (d - v * r * o.d) * r;
    v *= r;
    return *this;
  }
  // hidden friends, so mixed Dual / constant expressions convert implicitly
  __host__ __device__ friend Dual operator-(const Dual &a) { return Dual(-a.v, -a.d, 0); }
  __host__ __device__ friend Dual operator+(Dual a, const Dual &b) { return a += b; }
  __host__ __device__ friend Dual operator-(Dual a, const Dual &b) { return a -= b; }
  __host__ __device__ friend Dual operator*(Dual a, const Dual &b) { return a *= b; }
  __host__ __device__ friend Dual operator/(Dual a, const Dual &b) { return a /= b; }
  __host__ __device__ friend bool operator<(const Dual &a, const Dual &b) { return a.v < b.v; }
  __host__ __device__ friend bool operator>(const Dual &a, const Dual &b) { return a.v > b.v; }
  __host__ __device__ friend bool operator<=(const Dual &a, const Dual &b) { return a.v <= b.v; }
  __host__ __device__ friend bool operator>=(const Dual &a, const Dual &b) { return a.v >= b.v; }
  __host__ __device__ friend bool operator==(const Dual &a, const Dual &b) { return a.v == b.v; }
  __host__ __device__ friend bool operator!=(const Dual &a, const Dual &b) { return a.v != b.v; }
};

template <class U> struct BaseType<Dual<U>> { using T = U; };

template <class T> __device__ __forceinline__ Dual<T> fabs(const Dual<T> &x) { return x.v < T(0) ? -x : x; }
template <class T> __device__ __forceinline__ Dual<T> recip(const Dual<T> &x) {
  const T r = recip(x.v);
  return Dual<T>(r, -x.d * r * r, 0);
}
template <class T> __device__ __forceinline__ Dual<T> dlog1p(const Dual<T> &x) {
  return Dual<T>(dlog1p(x.v), x.d / (T(1) + x.v), 0);
}
template <class T> __device__ __forceinline__ Dual<T> dexpm1(const Dual<T> &x) {
  return Dual<T>(dexpm1(x.v), x.d * dexp(x.v), 0);
}
template <class T> __device__ __forceinline__ Dual<T> dexp(const Dual<T> &x) {
  const T e = dexp(x.v);
  return Dual<T>(e, x.d * e, 0);
}
// x^y; the exponents of these losses are constants (their derivative part is zero)
template <class T> __device__ __forceinline__ Dual<T> dpow(const Dual<T> &x, const Dual<T> &y) {
  const T p = dpow(x.v, y.v);
  return Dual<T>(p, x.d * y.v * dpow(x.v, y.v - T(1)), 0);
}

}  // namespace femhot
