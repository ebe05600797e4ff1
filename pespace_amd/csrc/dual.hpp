// Forward-mode dual numbers for device code.
//
// Dual<T, N> carries a value of type T and N tangents of type T.  T is double
// (first derivatives) or Dual<double, 1> (so Dual<Dual<double,1>, 4> carries the
// value, d/du, d/dtheta_i and the mixed d2/dtheta_i du that the node tables need).
#pragma once
#include <hip/hip_runtime.h>

#define PES_HD __host__ __device__ inline

namespace pes {

template <class T, int N>
struct Dual {
  T v;
  T d[N];
};

PES_HD double value_of(double x) { return x; }
template <class T, int N>
PES_HD double value_of(const Dual<T, N>& x) { return value_of(x.v); }

template <class T>
struct Lift {
  PES_HD static T from(double c) { return c; }
};
template <class T, int N>
struct Lift<Dual<T, N>> {
  PES_HD static Dual<T, N> from(double c) {
    Dual<T, N> r;
    r.v = Lift<T>::from(c);
    for (int i = 0; i < N; ++i) r.d[i] = Lift<T>::from(0.0);
    return r;
  }
};
template <class T>
PES_HD T lift(double c) { return Lift<T>::from(c); }

// ---- dual (op) dual ------------------------------------------------------------------
template <class T, int N>
PES_HD Dual<T, N> operator+(const Dual<T, N>& a, const Dual<T, N>& b) {
  Dual<T, N> r;
  r.v = a.v + b.v;
  for (int i = 0; i < N; ++i) r.d[i] = a.d[i] + b.d[i];
  return r;
}
template <class T, int N>
PES_HD Dual<T, N> operator-(const Dual<T, N>& a, const Dual<T, N>& b) {
  Dual<T, N> r;
  r.v = a.v - b.v;
  for (int i = 0; i < N; ++i) r.d[i] = a.d[i] - b.d[i];
  return r;
}
template <class T, int N>
PES_HD Dual<T, N> operator-(const Dual<T, N>& a) {
  Dual<T, N> r;
  r.v = -a.v;
  for (int i = 0; i < N; ++i) r.d[i] = -a.d[i];
  return r;
}
template <class T, int N>
PES_HD Dual<T, N> operator*(const Dual<T, N>& a, const Dual<T, N>& b) {
  Dual<T, N> r;
  r.v = a.v * b.v;
  for (int i = 0; i < N; ++i) r.d[i] = a.d[i] * b.v + a.v * b.d[i];
  return r;
}
template <class T, int N>
PES_HD Dual<T, N> operator/(const Dual<T, N>& a, const Dual<T, N>& b) {
  Dual<T, N> r;
  T inv = 1.0 / b.v;
  r.v = a.v * inv;
  for (int i = 0; i < N; ++i) r.d[i] = (a.d[i] - r.v * b.d[i]) * inv;
  return r;
}

// ---- dual (op) double -----------------------------------------------------------------
template <class T, int N>
PES_HD Dual<T, N> operator+(const Dual<T, N>& a, double b) {
  Dual<T, N> r = a;
  r.v = a.v + b;
  return r;
}
template <class T, int N>
PES_HD Dual<T, N> operator+(double b, const Dual<T, N>& a) { return a + b; }
template <class T, int N>
PES_HD Dual<T, N> operator-(const Dual<T, N>& a, double b) {
  Dual<T, N> r = a;
  r.v = a.v - b;
  return r;
}
template <class T, int N>
PES_HD Dual<T, N> operator-(double b, const Dual<T, N>& a) {
  Dual<T, N> r;
  r.v = b - a.v;
  for (int i = 0; i < N; ++i) r.d[i] = -a.d[i];
  return r;
}
template <class T, int N>
PES_HD Dual<T, N> operator*(const Dual<T, N>& a, double b) {
  Dual<T, N> r;
  r.v = a.v * b;
  for (int i = 0; i < N; ++i) r.d[i] = a.d[i] * b;
  return r;
}
template <class T, int N>
PES_HD Dual<T, N> operator*(double b, const Dual<T, N>& a) { return a * b; }
template <class T, int N>
PES_HD Dual<T, N> operator/(const Dual<T, N>& a, double b) { return a * (1.0 / b); }
template <class T, int N>
PES_HD Dual<T, N> operator/(double b, const Dual<T, N>& a) {
  Dual<T, N> r;
  T inv = 1.0 / a.v;
  r.v = b * inv;
  T k = -(r.v * inv);
  for (int i = 0; i < N; ++i) r.d[i] = k * a.d[i];
  return r;
}

// ---- elementary functions (chain rule) ----------------------------------------------
PES_HD double dsqrt(double x) { return sqrt(x); }
PES_HD double dlog(double x) { return log(x); }
PES_HD double dexp(double x) { return exp(x); }
PES_HD double datan(double x) { return atan(x); }
PES_HD double dpow(double x, double p) { return pow(x, p); }
PES_HD double dcbrt(double x) { return cbrt(x); }

template <class T, int N>
PES_HD Dual<T, N> dsqrt(const Dual<T, N>& x) {
  Dual<T, N> r;
  r.v = dsqrt(x.v);
  T k = 0.5 / r.v;
  for (int i = 0; i < N; ++i) r.d[i] = k * x.d[i];
  return r;
}
template <class T, int N>
PES_HD Dual<T, N> dcbrt(const Dual<T, N>& x) {
  Dual<T, N> r;
  r.v = dcbrt(x.v);
  T k = (1.0 / 3.0) * r.v / x.v;
  for (int i = 0; i < N; ++i) r.d[i] = k * x.d[i];
  return r;
}
template <class T, int N>
PES_HD Dual<T, N> dlog(const Dual<T, N>& x) {
  Dual<T, N> r;
  r.v = dlog(x.v);
  T k = 1.0 / x.v;
  for (int i = 0; i < N; ++i) r.d[i] = k * x.d[i];
  return r;
}
template <class T, int N>
PES_HD Dual<T, N> dexp(const Dual<T, N>& x) {
  Dual<T, N> r;
  r.v = dexp(x.v);
  for (int i = 0; i < N; ++i) r.d[i] = r.v * x.d[i];
  return r;
}
template <class T, int N>
PES_HD Dual<T, N> datan(const Dual<T, N>& x) {
  Dual<T, N> r;
  r.v = datan(x.v);
  T k = 1.0 / (1.0 + x.v * x.v);
  for (int i = 0; i < N; ++i) r.d[i] = k * x.d[i];
  return r;
}
template <class T, int N>
PES_HD Dual<T, N> dpow(const Dual<T, N>& x, double p) {
  Dual<T, N> r;
  r.v = dpow(x.v, p);
  T k = p * r.v / x.v;
  for (int i = 0; i < N; ++i) r.d[i] = k * x.d[i];
  return r;
}

PES_HD double dsin(double x) { return sin(x); }
PES_HD double dcos(double x) { return cos(x); }
template <class T, int N>
PES_HD Dual<T, N> dsin(const Dual<T, N>& x) {
  Dual<T, N> r;
  r.v = dsin(x.v);
  T k = dcos(x.v);
  for (int i = 0; i < N; ++i) r.d[i] = k * x.d[i];
  return r;
}
template <class T, int N>
PES_HD Dual<T, N> dcos(const Dual<T, N>& x) {
  Dual<T, N> r;
  r.v = dcos(x.v);
  T k = -dsin(x.v);
  for (int i = 0; i < N; ++i) r.d[i] = k * x.d[i];
  return r;
}

template <class T>
PES_HD T select(bool c, const T& a, const T& b) { return c ? a : b; }

}  // namespace pes
