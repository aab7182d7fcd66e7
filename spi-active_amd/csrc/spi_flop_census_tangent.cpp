// FLOP census of the tangent (forward-mode) rollout behind the Fisher-information sensitivities
// (host-only measurement program, the companion of spi_flop_census.cpp).
//
// Instantiates spi_dyn.h with a counting dual number that follows the arithmetic of
// spi_dual.h's Dual<N> component by component (sums and products of the value and of each
// tangent, the chain rule of tanh / sincos / reciprocal / reciprocal square root), with every
// component a counting scalar: an operation is counted when it acts on per-lane runtime values,
// and a product with a compile-time zero (a tangent that is still structurally zero, a zero
// of the model) is not.  It runs the excitation of the Fisher workload (the Go2 settled from a
// standing pose, then per-joint Fourier PD targets, 8 terms per joint with coefficients
// U[-0.3, 0.3] rad over a horizon of 50 control steps) with two parameter rows seeded (mass
// and the front-left Kp scale: the pair kernel carries two tangents on one value path) and
// prints the FLOPs per control step of one pair unit and per tangent.
//
// Build and run:  hipcc --offload-host-only -O0 -std=c++17 -I include -I spi-active_amd/csrc \
//                 spi-active_amd/csrc/spi_flop_census_tangent.cpp -o /tmp/census_t && /tmp/census_t
#include <cmath>
#include <cstdio>
#include <random>

namespace census {
long long g_add = 0, g_mul = 0, g_div = 0, g_trans = 0;

struct CountF {
  float v;
  bool zero;  // compile-time zero constant (folded on the device)
  bool rt;    // per-lane runtime value
  CountF() : v(0.f), zero(true), rt(false) {}
  CountF(float x) : v(x), zero(x == 0.0f), rt(false) {}
  static CountF input(float x) { CountF c(x); c.zero = false; c.rt = true; return c; }
  CountF operator-() const { CountF r = *this; r.v = -v; return r; }
};
inline CountF runtime(float x) { return CountF::input(x); }
inline CountF result(float x, const CountF& a, const CountF& b, long long& counter) {
  if (a.rt || b.rt) { ++counter; return runtime(x); }
  return CountF(x);
}
inline CountF operator+(CountF a, CountF b) {
  if (a.zero) return b;
  if (b.zero) return a;
  return result(a.v + b.v, a, b, g_add);
}
inline CountF operator-(CountF a, CountF b) {
  if (b.zero) return a;
  if (a.zero) return -b;
  return result(a.v - b.v, a, b, g_add);
}
inline CountF operator*(CountF a, CountF b) {
  if (a.zero || b.zero) return CountF(0.0f);
  return result(a.v * b.v, a, b, g_mul);
}
inline CountF operator/(CountF a, CountF b) {
  if (a.zero) return CountF(0.0f);
  return result(a.v / b.v, a, b, g_div);
}
inline CountF trans(float x, const CountF& a) {
  if (!a.rt) return CountF(x);
  ++g_trans;
  return runtime(x);
}

// dual number with counting components, spi_dual.h's rules
template <int N> struct CD {
  CountF v;
  CountF d[N];
  CD() {}
  CD(float x) : v(x) {}
  CD(double x) : v(float(x)) {}
  CD(int x) : v(float(x)) {}
  CD(const CountF& x) : v(x) {}
  CD operator-() const { CD r; r.v = -v; for (int i = 0; i < N; ++i) r.d[i] = -d[i]; return r; }
  CD& operator+=(const CD& b);
  CD& operator-=(const CD& b);
  CD& operator*=(const CD& b);
};
template <int N> CD<N> chain(const CountF& fv, const CountF& s, const CD<N>& a) {
  CD<N> r;
  r.v = fv;
  for (int i = 0; i < N; ++i) r.d[i] = s * a.d[i];
  return r;
}
template <int N> CD<N> operator+(const CD<N>& a, const CD<N>& b) {
  CD<N> r; r.v = a.v + b.v; for (int i = 0; i < N; ++i) r.d[i] = a.d[i] + b.d[i]; return r;
}
template <int N> CD<N> operator-(const CD<N>& a, const CD<N>& b) {
  CD<N> r; r.v = a.v - b.v; for (int i = 0; i < N; ++i) r.d[i] = a.d[i] - b.d[i]; return r;
}
template <int N> CD<N> operator*(const CD<N>& a, const CD<N>& b) {
  CD<N> r; r.v = a.v * b.v; for (int i = 0; i < N; ++i) r.d[i] = a.v * b.d[i] + b.v * a.d[i]; return r;
}
template <int N> CD<N> operator+(const CD<N>& a, float b) { CD<N> r = a; r.v = a.v + CountF(b); return r; }
template <int N> CD<N> operator+(float a, const CD<N>& b) { return b + a; }
template <int N> CD<N> operator-(const CD<N>& a, float b) { CD<N> r = a; r.v = a.v - CountF(b); return r; }
template <int N> CD<N> operator-(float a, const CD<N>& b) { CD<N> r = -b; r.v = CountF(a) - b.v; return r; }
template <int N> CD<N> operator*(const CD<N>& a, float b) { return chain(a.v * CountF(b), CountF(b), a); }
template <int N> CD<N> operator*(float a, const CD<N>& b) { return chain(CountF(a) * b.v, CountF(a), b); }
inline CountF rcp_c(const CountF& x) { return CountF(1.0f) / x; }
template <int N> CD<N> operator/(const CD<N>& a, const CD<N>& b) {
  const CountF ib = rcp_c(b.v);
  CD<N> r;
  r.v = a.v * ib;
  for (int i = 0; i < N; ++i) r.d[i] = (a.d[i] - r.v * b.d[i]) * ib;
  return r;
}
template <int N> CD<N> operator/(const CD<N>& a, float b) { return a * (1.0f / b); }
template <int N> CD<N> operator/(float a, const CD<N>& b) { return CD<N>(a) / b; }
template <int N> CD<N>& CD<N>::operator+=(const CD<N>& b) { return *this = *this + b; }
template <int N> CD<N>& CD<N>::operator-=(const CD<N>& b) { return *this = *this - b; }
template <int N> CD<N>& CD<N>::operator*=(const CD<N>& b) { return *this = *this * b; }
template <int N> bool operator<(const CD<N>& a, const CD<N>& b) { return a.v.v < b.v.v; }
template <int N> bool operator>(const CD<N>& a, const CD<N>& b) { return a.v.v > b.v.v; }
template <int N> bool operator<=(const CD<N>& a, const CD<N>& b) { return a.v.v <= b.v.v; }
template <int N> bool operator>=(const CD<N>& a, const CD<N>& b) { return a.v.v >= b.v.v; }
template <int N> bool operator<(const CD<N>& a, float b) { return a.v.v < b; }
template <int N> bool operator>(const CD<N>& a, float b) { return a.v.v > b; }
template <int N> bool operator<=(const CD<N>& a, float b) { return a.v.v <= b; }
template <int N> bool operator>=(const CD<N>& a, float b) { return a.v.v >= b; }
template <int N> bool operator==(const CD<N>& a, float b) { return a.v.v == b; }
template <int N> bool operator!=(const CD<N>& a, float b) { return a.v.v != b; }
template <int N> CD<N> fminf(const CD<N>& a, const CD<N>& b) { return b.v.v < a.v.v ? b : a; }
template <int N> CD<N> fmaxf(const CD<N>& a, const CD<N>& b) { return b.v.v > a.v.v ? b : a; }
template <int N> CD<N> fminf(const CD<N>& a, float b) { return b < a.v.v ? CD<N>(b) : a; }
template <int N> CD<N> fmaxf(const CD<N>& a, float b) { return b > a.v.v ? CD<N>(b) : a; }
template <int N> CD<N> fabsf(const CD<N>& a) { return a.v.v < 0.0f ? -a : a; }
template <int N> CD<N> copysignf(const CD<N>& a, const CD<N>& b) {
  return std::signbit(a.v.v) != std::signbit(b.v.v) ? -a : a;
}
template <int N> CD<N> sqrtf(const CD<N>& a) {
  const CountF s = trans(std::sqrt(a.v.v), a.v);
  return chain(s, s.v > 0.0f ? CountF(0.5f) * rcp_c(s) : CountF(0.0f), a);
}
}  // namespace census

constexpr int kTan = 2;   // two tangents per unit, as spi_traj_sensitivity_pairs
using TD = census::CD<kTan>;

namespace spi {
inline TD spi_tanh(TD x) {
  const census::CountF t = census::trans(std::tanh(x.v.v), x.v);
  return census::chain(t, census::CountF(1.0f) - t * t, x);
}
inline void spi_sincos(TD x, TD* s, TD* c) {
  const census::CountF sv = census::trans(std::sin(x.v.v), x.v), cv = census::trans(std::cos(x.v.v), x.v);
  *s = census::chain(sv, cv, x);
  *c = census::chain(cv, -sv, x);
}
inline TD spi_rcp(TD x) {
  const census::CountF r = census::rcp_c(x.v);
  return census::chain(r, -(r * r), x);
}
inline TD spi_rsqrt(TD x) {
  const census::CountF r = census::trans(1.0f / std::sqrt(x.v.v), x.v);
  return census::chain(r, census::CountF(-0.5f) * r * r * r, x);
}
}  // namespace spi

#define SPI_REAL_T TD
#define SPI_REAL_HAS_MATH
#include "spi_dyn.h"

int main() {
  using namespace spi;
  using census::CountF;
  SpiSimArgs a{};
  a.dt = 0.005f; a.decimation = 4; a.substeps = 4; a.gravity = 9.81f;
  const float q0[12] = {0.1f, 0.8f, -1.5f, -0.1f, 0.8f, -1.5f, 0.1f, 1.0f, -1.5f, -0.1f, 1.0f, -1.5f};
  const float eff[12] = {23.7f, 23.7f, 35.55f, 23.7f, 23.7f, 35.55f, 23.7f, 23.7f, 35.55f, 23.7f, 23.7f, 35.55f};
  for (int j = 0; j < 12; ++j) { a.kp[j] = 20.f; a.kd[j] = 0.5f; a.q0[j] = q0[j]; a.effort[j] = eff[j]; }
  a.action_scale = 0.25f; a.action_clip = 20.f; a.hip_scale = 1.f; a.motor_model = SPI_MOTOR_NONE;
  a.k_contact = 5e4f; a.c_contact = 100.f; a.mu = 0.8f; a.c_tangent = 100.f; a.fric_vel_inv = 5.f;
  a.state_limit = 1e4f;
  auto par = [](float x, int slot) {           // a parameter, seeded with a unit tangent in `slot`
    TD r(CountF::input(x));
    if (slot >= 0) r.d[slot] = CountF::input(1.0f);
    return r;
  };
  LaneParams lp;
  lp.base = make_base_inertia(par(6.921f, 0), par(0.021112f, -1), par(0.f, -1), par(-0.005366f, -1),
                              par(0.02448f, -1), par(0.098077f, -1), par(0.107f, -1), par(0.00012166f, -1),
                              par(0.0014849f, -1), par(-3.12e-05f, -1));
  for (int l = 0; l < 4; ++l) {
    lp.kp_s[l] = par(1.f, l == 0 ? 1 : -1); lp.kd_s[l] = par(1.f, -1); lp.fric[l] = par(0.1f, -1);
  }
  for (int g = 0; g < 3; ++g) lp.motor[g] = par(20.f, -1);

  const int n_traj = 8, settle = 50, H = 50, n_terms = 8;
  std::mt19937 rng(11);
  std::uniform_real_distribution<float> uc(-0.3f, 0.3f);
  long long flops = 0;
  for (int tr = 0; tr < n_traj; ++tr) {
    State st;
    st.p = V3{TD(CountF::input(0.f)), TD(CountF::input(0.f)), TD(CountF::input(0.30f))};
    st.qx = TD(CountF::input(0.f)); st.qy = TD(CountF::input(0.f)); st.qz = TD(CountF::input(0.f));
    st.qw = TD(CountF::input(1.f));
    st.w = V3{TD(CountF::input(0.f)), TD(CountF::input(0.f)), TD(CountF::input(0.f))};
    st.v = st.w;
    for (int j = 0; j < 12; ++j) { st.q[j] = TD(CountF::input(q0[j])); st.qd[j] = TD(CountF::input(0.f)); }
    float c[12][8];
    for (int j = 0; j < 12; ++j)
      for (int k = 0; k < n_terms; ++k) c[j][k] = uc(rng);
    for (int t = 0; t < settle + H; ++t) {
      TD act[12];
      for (int j = 0; j < 12; ++j) {
        float off = 0.f;
        if (t >= settle)
          for (int k = 0; k < n_terms; ++k) off += c[j][k] * std::sin(6.2831853f * (k + 1) * (t - settle) / H);
        act[j] = TD(CountF::input(off / 0.25f));
      }
      const long long before = census::g_add + census::g_mul + census::g_div + census::g_trans;
      for (int d = 0; d < a.decimation; ++d) {
        TD tau[12];
        control_torques(a, lp, act, st, tau);
        physics_step(a, lp, st, tau, nullptr);
      }
      const long long after = census::g_add + census::g_mul + census::g_div + census::g_trans;
      if (t >= settle) flops += after - before;
    }
  }
  const double per_unit = double(flops) / (n_traj * H);
  std::printf("{\"tangent_pair_flops_per_control_step\": %.1f, \"flops_per_tangent_control_step\": %.1f, "
              "\"tangents_per_unit\": %d, \"control_steps\": %d}\n",
              per_unit, per_unit / kTan, kTan, n_traj * H);
  return 0;
}
