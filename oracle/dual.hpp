// dual.hpp -- forward-mode dual numbers for the CPU restatement.
//
// TEST INFRASTRUCTURE ONLY.  Instantiating the restatement (templated on its
// scalar type T) on Dual<N> carries N directional derivatives through every
// operation of a step, the algorithm of ForwardDiff.jl's Dual{Tag,V,N}
// (value and partials; each elementary function applies its derivative to
// the partials): the ForwardDiff gradient of the CPU path that the reverse
// mode of the HIP library is checked against.  The value part of every
// operation is the same double operation the T = double restatement does,
// so values are bit-identical to it; branches compare values only.
#pragma once

#include <cmath>

namespace orc {

template <int N> struct Dual {
  double v;
  double d[N];
  Dual() : v(0.0) {
    for (int k = 0; k < N; k++) d[k] = 0.0;
  }
  Dual(double x) : v(x) {  // NOLINT: implicit, as the restatement mixes literals and T
    for (int k = 0; k < N; k++) d[k] = 0.0;
  }
  Dual& operator+=(const Dual& o) { return *this = *this + o; }
  Dual& operator-=(const Dual& o) { return *this = *this - o; }
  Dual& operator*=(const Dual& o) { return *this = *this * o; }
  Dual& operator/=(const Dual& o) { return *this = *this / o; }

  friend Dual operator+(const Dual& a, const Dual& b) {
    Dual r(a.v + b.v);
    for (int k = 0; k < N; k++) r.d[k] = a.d[k] + b.d[k];
    return r;
  }
  friend Dual operator-(const Dual& a, const Dual& b) {
    Dual r(a.v - b.v);
    for (int k = 0; k < N; k++) r.d[k] = a.d[k] - b.d[k];
    return r;
  }
  friend Dual operator-(const Dual& a) {
    Dual r(-a.v);
    for (int k = 0; k < N; k++) r.d[k] = -a.d[k];
    return r;
  }
  friend Dual operator*(const Dual& a, const Dual& b) {
    Dual r(a.v * b.v);
    for (int k = 0; k < N; k++) r.d[k] = a.d[k] * b.v + a.v * b.d[k];
    return r;
  }
  friend Dual operator/(const Dual& a, const Dual& b) {
    Dual r(a.v / b.v);
    for (int k = 0; k < N; k++) r.d[k] = (a.d[k] * b.v - a.v * b.d[k]) / (b.v * b.v);
    return r;
  }
  friend Dual operator+(const Dual& a, double b) { return a + Dual(b); }
  friend Dual operator+(double a, const Dual& b) { return Dual(a) + b; }
  friend Dual operator-(const Dual& a, double b) { return a - Dual(b); }
  friend Dual operator-(double a, const Dual& b) { return Dual(a) - b; }
  friend Dual operator*(const Dual& a, double b) {
    Dual r(a.v * b);
    for (int k = 0; k < N; k++) r.d[k] = a.d[k] * b;
    return r;
  }
  friend Dual operator*(double a, const Dual& b) {
    Dual r(a * b.v);
    for (int k = 0; k < N; k++) r.d[k] = a * b.d[k];
    return r;
  }
  friend Dual operator/(const Dual& a, double b) {
    Dual r(a.v / b);
    for (int k = 0; k < N; k++) r.d[k] = a.d[k] / b;
    return r;
  }
  friend Dual operator/(double a, const Dual& b) { return Dual(a) / b; }

  // comparisons look at the value (the restatement's branches)
  friend bool operator<(const Dual& a, const Dual& b) { return a.v < b.v; }
  friend bool operator>(const Dual& a, const Dual& b) { return a.v > b.v; }
  friend bool operator<=(const Dual& a, const Dual& b) { return a.v <= b.v; }
  friend bool operator>=(const Dual& a, const Dual& b) { return a.v >= b.v; }
  friend bool operator<(const Dual& a, double b) { return a.v < b; }
  friend bool operator>(const Dual& a, double b) { return a.v > b; }
};

template <int N> inline double val(const Dual<N>& x) { return x.v; }

template <int N> inline Dual<N> sqrt(const Dual<N>& a) {
  Dual<N> r(std::sqrt(a.v));
  for (int k = 0; k < N; k++) r.d[k] = a.d[k] / (2.0 * r.v);
  return r;
}
template <int N> inline Dual<N> log(const Dual<N>& a) {
  Dual<N> r(std::log(a.v));
  for (int k = 0; k < N; k++) r.d[k] = a.d[k] / a.v;
  return r;
}
// a^b = exp(b log a): d = a^b (b' log a + b a'/a)
template <int N> inline Dual<N> pow(const Dual<N>& a, const Dual<N>& b) {
  Dual<N> r(std::pow(a.v, b.v));
  const double la = a.v > 0.0 ? std::log(a.v) : 0.0;
  for (int k = 0; k < N; k++) r.d[k] = r.v * (b.d[k] * la + b.v * a.d[k] / a.v);
  return r;
}

}  // namespace orc
