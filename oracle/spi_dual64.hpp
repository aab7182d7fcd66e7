// ORACLE — test infrastructure only.  FP64 forward-mode (tangent) scalar for the oracle
// engine: instantiating Engine<DualD<N>> differentiates the FP64 rollout exactly, giving the
// reference derivatives that the GPU sensitivities are checked against.
#pragma once
#include <cmath>

namespace spio_ad {

template <int N> struct DualD {
  double v = 0.0;
  double d[N] = {};
  DualD() = default;
  DualD(double x) : v(x) {}
  DualD(float x) : v(x) {}
  DualD(int x) : v(x) {}
  explicit operator double() const { return v; }
  DualD operator-() const { DualD r; r.v = -v; for (int i = 0; i < N; ++i) r.d[i] = -d[i]; return r; }
  DualD& operator+=(const DualD& b) { v += b.v; for (int i = 0; i < N; ++i) d[i] += b.d[i]; return *this; }
  DualD& operator-=(const DualD& b) { v -= b.v; for (int i = 0; i < N; ++i) d[i] -= b.d[i]; return *this; }
  DualD& operator*=(const DualD& b) { return *this = *this * b; }
  friend DualD operator+(DualD a, const DualD& b) { return a += b; }
  friend DualD operator-(DualD a, const DualD& b) { return a -= b; }
  friend DualD operator*(const DualD& a, const DualD& b) {
    DualD r;
    r.v = a.v * b.v;
    for (int i = 0; i < N; ++i) r.d[i] = a.v * b.d[i] + b.v * a.d[i];
    return r;
  }
  friend DualD operator/(const DualD& a, const DualD& b) {
    DualD r;
    r.v = a.v / b.v;
    for (int i = 0; i < N; ++i) r.d[i] = (a.d[i] - r.v * b.d[i]) / b.v;
    return r;
  }
  friend bool operator<(const DualD& a, const DualD& b) { return a.v < b.v; }
  friend bool operator>(const DualD& a, const DualD& b) { return a.v > b.v; }
  friend bool operator<=(const DualD& a, const DualD& b) { return a.v <= b.v; }
  friend bool operator>=(const DualD& a, const DualD& b) { return a.v >= b.v; }
};

template <int N> DualD<N> unary(double fv, double slope, const DualD<N>& a) {
  DualD<N> r;
  r.v = fv;
  for (int i = 0; i < N; ++i) r.d[i] = slope * a.d[i];
  return r;
}
template <int N> DualD<N> sqrt(const DualD<N>& a) {
  double s = std::sqrt(a.v);
  return unary(s, s > 0.0 ? 0.5 / s : 0.0, a);
}
template <int N> DualD<N> sin(const DualD<N>& a) { return unary(std::sin(a.v), std::cos(a.v), a); }
template <int N> DualD<N> cos(const DualD<N>& a) { return unary(std::cos(a.v), -std::sin(a.v), a); }
template <int N> DualD<N> tanh(const DualD<N>& a) {
  double t = std::tanh(a.v);
  return unary(t, 1.0 - t * t, a);
}
template <int N> DualD<N> abs(const DualD<N>& a) { return a.v < 0.0 ? -a : a; }

}  // namespace spio_ad
