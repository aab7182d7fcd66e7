// Forward-mode (tangent) arithmetic for the per-lane dynamics: a value and N directional
// derivatives carried through every operation of spi_dyn.h.  Instantiating the rollout
// with this scalar gives the exact derivative of the discrete rollout (the same branches
// as the value: contact on/off, cone saturation, clipping) with respect to whichever inputs
// are seeded with a unit tangent.
//
// Include before spi_dyn.h with
//   #define SPI_REAL_T spi_ad::Dual<N>   #define SPI_REAL_HAS_MATH   #define SPI_DYN_NS <ns>
// and declare the four lane math functions of spi_dyn.h (spi_tanh, spi_sincos, spi_rcp,
// spi_rsqrt) for the type in <ns> (SPI_DUAL_MATH(ns, N) below does that).
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>

#include "spi_quad.h"

#define SPI_AD_HD __host__ __device__ __forceinline__

namespace spi_ad {

template <int N> struct Dual {
  float v;
  float d[N];
  SPI_AD_HD Dual() : v(0.0f) {
#pragma unroll
    for (int i = 0; i < N; ++i) d[i] = 0.0f;
  }
  SPI_AD_HD Dual(float x) : v(x) {  // constants and data: zero tangent
#pragma unroll
    for (int i = 0; i < N; ++i) d[i] = 0.0f;
  }
  SPI_AD_HD Dual(double x) : Dual(static_cast<float>(x)) {}
  SPI_AD_HD Dual(int x) : Dual(static_cast<float>(x)) {}
  SPI_AD_HD Dual operator-() const {
    Dual r;
    r.v = -v;
#pragma unroll
    for (int i = 0; i < N; ++i) r.d[i] = -d[i];
    return r;
  }
  SPI_AD_HD Dual& operator+=(const Dual& b);
  SPI_AD_HD Dual& operator-=(const Dual& b);
  SPI_AD_HD Dual& operator*=(const Dual& b);
};

// value with tangent t * (derivative scale) of a unary function: r.v = fv, r.d = s * a.d
template <int N> SPI_AD_HD Dual<N> chain(float fv, float s, const Dual<N>& a) {
  Dual<N> r;
  r.v = fv;
#pragma unroll
  for (int i = 0; i < N; ++i) r.d[i] = s * a.d[i];
  return r;
}

template <int N> SPI_AD_HD Dual<N> operator+(const Dual<N>& a, const Dual<N>& b) {
  Dual<N> r;
  r.v = a.v + b.v;
#pragma unroll
  for (int i = 0; i < N; ++i) r.d[i] = a.d[i] + b.d[i];
  return r;
}
template <int N> SPI_AD_HD Dual<N> operator-(const Dual<N>& a, const Dual<N>& b) {
  Dual<N> r;
  r.v = a.v - b.v;
#pragma unroll
  for (int i = 0; i < N; ++i) r.d[i] = a.d[i] - b.d[i];
  return r;
}
template <int N> SPI_AD_HD Dual<N> operator*(const Dual<N>& a, const Dual<N>& b) {
  Dual<N> r;
  r.v = a.v * b.v;
#pragma unroll
  for (int i = 0; i < N; ++i) r.d[i] = a.v * b.d[i] + b.v * a.d[i];
  return r;
}
template <int N> SPI_AD_HD Dual<N> operator+(const Dual<N>& a, float b) {
  Dual<N> r = a;
  r.v = a.v + b;
  return r;
}
template <int N> SPI_AD_HD Dual<N> operator+(float a, const Dual<N>& b) { return b + a; }
template <int N> SPI_AD_HD Dual<N> operator-(const Dual<N>& a, float b) {
  Dual<N> r = a;
  r.v = a.v - b;
  return r;
}
template <int N> SPI_AD_HD Dual<N> operator-(float a, const Dual<N>& b) {
  Dual<N> r = -b;
  r.v = a - b.v;
  return r;
}
template <int N> SPI_AD_HD Dual<N> operator*(const Dual<N>& a, float b) { return chain(a.v * b, b, a); }
template <int N> SPI_AD_HD Dual<N> operator*(float a, const Dual<N>& b) { return chain(a * b.v, a, b); }

SPI_AD_HD float rcp_f(float x) {
#ifdef __HIP_DEVICE_COMPILE__
  return __builtin_amdgcn_rcpf(x);
#else
  return 1.0f / x;
#endif
}
template <int N> SPI_AD_HD Dual<N> operator/(const Dual<N>& a, const Dual<N>& b) {
  const float ib = rcp_f(b.v);
  Dual<N> r;
  r.v = a.v * ib;
#pragma unroll
  for (int i = 0; i < N; ++i) r.d[i] = (a.d[i] - r.v * b.d[i]) * ib;
  return r;
}
template <int N> SPI_AD_HD Dual<N> operator/(const Dual<N>& a, float b) { return a * rcp_f(b); }
template <int N> SPI_AD_HD Dual<N> operator/(float a, const Dual<N>& b) { return Dual<N>(a) / b; }

template <int N> SPI_AD_HD Dual<N>& Dual<N>::operator+=(const Dual<N>& b) { return *this = *this + b; }
template <int N> SPI_AD_HD Dual<N>& Dual<N>::operator-=(const Dual<N>& b) { return *this = *this - b; }
template <int N> SPI_AD_HD Dual<N>& Dual<N>::operator*=(const Dual<N>& b) { return *this = *this * b; }

// comparisons look at values only: the tangent follows the branch the value takes
template <int N> SPI_AD_HD bool operator<(const Dual<N>& a, const Dual<N>& b) { return a.v < b.v; }
template <int N> SPI_AD_HD bool operator>(const Dual<N>& a, const Dual<N>& b) { return a.v > b.v; }
template <int N> SPI_AD_HD bool operator<=(const Dual<N>& a, const Dual<N>& b) { return a.v <= b.v; }
template <int N> SPI_AD_HD bool operator>=(const Dual<N>& a, const Dual<N>& b) { return a.v >= b.v; }
template <int N> SPI_AD_HD bool operator<(const Dual<N>& a, float b) { return a.v < b; }
template <int N> SPI_AD_HD bool operator>(const Dual<N>& a, float b) { return a.v > b; }
template <int N> SPI_AD_HD bool operator<=(const Dual<N>& a, float b) { return a.v <= b; }
template <int N> SPI_AD_HD bool operator>=(const Dual<N>& a, float b) { return a.v >= b; }
template <int N> SPI_AD_HD bool operator==(const Dual<N>& a, float b) { return a.v == b; }
template <int N> SPI_AD_HD bool operator!=(const Dual<N>& a, float b) { return a.v != b; }

template <int N> SPI_AD_HD Dual<N> fminf(const Dual<N>& a, const Dual<N>& b) { return b.v < a.v ? b : a; }
template <int N> SPI_AD_HD Dual<N> fmaxf(const Dual<N>& a, const Dual<N>& b) { return b.v > a.v ? b : a; }
template <int N> SPI_AD_HD Dual<N> fminf(const Dual<N>& a, float b) { return b < a.v ? Dual<N>(b) : a; }
template <int N> SPI_AD_HD Dual<N> fmaxf(const Dual<N>& a, float b) { return b > a.v ? Dual<N>(b) : a; }
template <int N> SPI_AD_HD Dual<N> fabsf(const Dual<N>& a) { return a.v < 0.0f ? -a : a; }
// copysign(a, b) = |a| sgn(b): the tangent of a, negated when the sign changes
template <int N> SPI_AD_HD Dual<N> copysignf(const Dual<N>& a, const Dual<N>& b) {
  const bool flip = __builtin_signbit(a.v) != __builtin_signbit(b.v);
  return flip ? -a : a;
}
template <int N> SPI_AD_HD Dual<N> sqrtf(const Dual<N>& a) {
  const float s = ::sqrtf(a.v);
  return chain(s, s > 0.0f ? 0.5f * rcp_f(s) : 0.0f, a);
}

// lane math of spi_dyn.h for Dual<N>
SPI_AD_HD float tanh_f(float x) {
  // odd Taylor series near 0, 1 - 2 / (1 + e^{2|x|}) elsewhere, sign restored
  const float ax = ::fabsf(x);
#ifdef __HIP_DEVICE_COMPILE__
  const float big = 1.0f - 2.0f * __builtin_amdgcn_rcpf(1.0f + __expf(2.0f * ax));
#else
  const float big = 1.0f - 2.0f / (1.0f + ::expf(2.0f * ax));
#endif
  const float x2 = ax * ax;
  const float small = ax * (1.0f + x2 * (-1.0f / 3.0f + x2 * (2.0f / 15.0f + x2 * (-17.0f / 315.0f))));
  return ::copysignf(ax < 0.0625f ? small : big, x);
}
template <int N> SPI_AD_HD Dual<N> tanh_d(const Dual<N>& x) {
  const float t = tanh_f(x.v);
  return chain(t, 1.0f - t * t, x);
}
template <int N> SPI_AD_HD void sincos_d(const Dual<N>& x, Dual<N>* s, Dual<N>* c) {
#ifdef __HIP_DEVICE_COMPILE__
  const float sv = __sinf(x.v), cv = __cosf(x.v);
#else
  const float sv = ::sinf(x.v), cv = ::cosf(x.v);
#endif
  *s = chain(sv, cv, x);
  *c = chain(cv, -sv, x);
}
template <int N> SPI_AD_HD Dual<N> rcp_d(const Dual<N>& x) {
  const float r = rcp_f(x.v);
  return chain(r, -r * r, x);
}
template <int N> SPI_AD_HD Dual<N> rsqrt_d(const Dual<N>& x) {
#ifdef __HIP_DEVICE_COMPILE__
  const float r = __builtin_amdgcn_rsqf(x.v);
#else
  const float r = 1.0f / ::sqrtf(x.v);
#endif
  return chain(r, -0.5f * r * r * r, x);
}

// exchange of a whole dual number (value and tangents) with lane (this lane xor M) of its quad
template <int M, int N> __device__ __forceinline__ Dual<N> shfl_xor_lane(const Dual<N>& x) {
  Dual<N> r;
  r.v = spi_quad_xor<M>(x.v);
#pragma unroll
  for (int i = 0; i < N; ++i) r.d[i] = spi_quad_xor<M>(x.d[i]);
  return r;
}
// every component as a leaf for reassociation (spi_quad_pin)
template <int N> __device__ __forceinline__ Dual<N> quad_pin(const Dual<N>& x) {
  Dual<N> r;
  r.v = spi_quad_pin(x.v);
#pragma unroll
  for (int i = 0; i < N; ++i) r.d[i] = spi_quad_pin(x.d[i]);
  return r;
}

}  // namespace spi_ad

// the four lane math functions spi_dyn.h calls, for Dual<N>, declared in namespace NS
#define SPI_DUAL_MATH(NS, NN)                                                                        \
  namespace NS {                                                                                    \
  SPI_AD_HD spi_ad::Dual<NN> spi_tanh(spi_ad::Dual<NN> x) { return spi_ad::tanh_d(x); }              \
  SPI_AD_HD void spi_sincos(spi_ad::Dual<NN> x, spi_ad::Dual<NN>* s, spi_ad::Dual<NN>* c) {          \
    spi_ad::sincos_d(x, s, c);                                                                      \
  }                                                                                                 \
  SPI_AD_HD spi_ad::Dual<NN> spi_rcp(spi_ad::Dual<NN> x) { return spi_ad::rcp_d(x); }                \
  SPI_AD_HD spi_ad::Dual<NN> spi_rsqrt(spi_ad::Dual<NN> x) { return spi_ad::rsqrt_d(x); }            \
  }
