// FLOP census of the rollout kernel's algorithm (host-only measurement program).
//
// Instantiates the device dynamics of spi_dyn.h with an operation-counting scalar and runs
// a representative rollout on the host: the Go2 settles from a standing pose, then follows
// per-joint sinusoidal PD targets (amplitude U[0.1, 0.4] rad, frequency U[0.5, 3] Hz).  It
// prints the mean number of floating-point operations per control step of the fused
// rollout (PD + motor model + ABA + contact + integration per sub-interval, plus the
// tracking cost), which is the algorithmic work unit behind the roofline figure.
//
// Counting rule: +, -, *, / one FLOP each (a fused multiply-add therefore counts two);
// sqrt, exp, sin, cos one each; min/max/abs/copysign and comparisons none.  Only operations
// on per-lane runtime values count: an operation whose operands are all compile-time
// constants of the model, or that has a compile-time zero operand, is folded away by the
// device compiler and is not counted.  So the census counts what the algorithm needs per
// rollout, not what a generic dense implementation would do.
//
// It then replays segment records cut from the same excitation through the reverse-mode cost
// + gradient of one rollout (spi_adj.h rollout_grad on the same counting scalar: forward pass
// with checkpoints, block recompute, physics-step and sub-interval recompute, sub-interval
// adjoints) and prints its FLOPs per control step (H 20, checkpoints every 10 control steps).
//
// Build and run:  hipcc --offload-host-only -O0 -std=c++17 -I include -I spi-active_amd/csrc \
//                 spi-active_amd/csrc/spi_flop_census.cpp -o /tmp/census && /tmp/census
// (-O0: optimising the counting-scalar instantiation of the adjoint takes the compiler many minutes)
#include <cmath>
#include <cstdio>
#include <random>

namespace census {
long long g_add = 0, g_mul = 0, g_div = 0, g_trans = 0;

struct CountF {
  float v;
  bool zero;  // compile-time zero constant (folded on the device)
  bool rt;    // per-lane runtime value (state, parameters, or anything computed from them)
  CountF() : v(0.f), zero(true), rt(false) {}
  CountF(float x) : v(x), zero(x == 0.0f), rt(false) {}
  CountF(double x) : v(float(x)), zero(x == 0.0), rt(false) {}
  CountF(int x) : v(float(x)), zero(x == 0), rt(false) {}
  static CountF input(float x) { CountF c(x); c.zero = false; c.rt = true; return c; }
  static CountF konst(float x) { return CountF(x); }
  CountF operator-() const { CountF r = *this; r.v = -v; return r; }
  CountF& operator+=(const CountF& b);
  CountF& operator-=(const CountF& b);
  CountF& operator*=(const CountF& b);
};
inline CountF runtime(float x) { return CountF::input(x); }
// result of an operation: runtime if any operand is, else a (folded) constant
inline CountF result(float x, const CountF& a, const CountF& b, long long& counter) {
  if (a.rt || b.rt) { ++counter; return runtime(x); }
  return CountF::konst(x);
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
inline CountF& CountF::operator+=(const CountF& b) { return *this = *this + b; }
inline CountF& CountF::operator-=(const CountF& b) { return *this = *this - b; }
inline CountF& CountF::operator*=(const CountF& b) { return *this = *this * b; }
inline bool operator<(CountF a, CountF b) { return a.v < b.v; }
inline bool operator>(CountF a, CountF b) { return a.v > b.v; }
inline bool operator<=(CountF a, CountF b) { return a.v <= b.v; }
inline bool operator>=(CountF a, CountF b) { return a.v >= b.v; }
inline bool operator==(CountF a, CountF b) { return a.v == b.v; }
inline bool operator!=(CountF a, CountF b) { return a.v != b.v; }
inline CountF fminf(CountF a, CountF b) { CountF r = runtime(std::fmin(a.v, b.v)); r.rt = a.rt || b.rt; return r; }
inline CountF fmaxf(CountF a, CountF b) { CountF r = runtime(std::fmax(a.v, b.v)); r.rt = a.rt || b.rt; return r; }
inline CountF fabsf(CountF a) { CountF r = a; r.v = std::fabs(a.v); return r; }
inline CountF copysignf(CountF a, CountF b) { CountF r = a; r.v = std::copysign(a.v, b.v); return r; }
inline CountF trans(float x, const CountF& a) {
  if (!a.rt) return CountF::konst(x);
  ++g_trans;
  return runtime(x);
}
inline CountF sqrtf(CountF a) { return trans(std::sqrt(a.v), a); }
inline CountF expf(CountF a) { return trans(std::exp(a.v), a); }
inline CountF sinf(CountF a) { return trans(std::sin(a.v), a); }
inline CountF cosf(CountF a) { return trans(std::cos(a.v), a); }
inline bool isfinite(CountF a) { return std::isfinite(a.v); }
}  // namespace census

#define SPI_REAL_T census::CountF
#include "spi_dyn.h"
#include "spi_adj.h"

#include <vector>

namespace census {
struct CountStore {              // the adjoint's checkpoint store (device: a strided HBM buffer)
  std::vector<spi::State> s;
  void put(int k, const spi::State& x) { s[k] = x; }
  void get(int k, spi::State& x) { x = s[k]; }
};
}  // namespace census

using census::CountF;

int main() {
  using namespace spi;
  SpiSimArgs a{};
  a.dt = 0.005f; a.decimation = 4; a.substeps = 4; a.gravity = 9.81f;
  const float q0[12] = {0.1f, 0.8f, -1.5f, -0.1f, 0.8f, -1.5f, 0.1f, 1.0f, -1.5f, -0.1f, 1.0f, -1.5f};
  const float eff[12] = {23.7f, 23.7f, 35.55f, 23.7f, 23.7f, 35.55f, 23.7f, 23.7f, 35.55f, 23.7f, 23.7f, 35.55f};
  for (int j = 0; j < 12; ++j) { a.kp[j] = 20.f; a.kd[j] = 0.5f; a.q0[j] = q0[j]; a.effort[j] = eff[j]; }
  a.action_scale = 0.25f; a.action_clip = 20.f; a.hip_scale = 1.f; a.motor_model = SPI_MOTOR_NONE;
  a.k_contact = 5e4f; a.c_contact = 100.f; a.mu = 0.8f; a.c_tangent = 100.f; a.fric_vel_inv = 5.f;
  a.state_limit = 1e4f;
  SpiCostArgs cost{};
  cost.mode = SPI_COST_TRACKING;
  const float w[6] = {10.f, 5.f, 1.f, 1.f, 1.f, 0.05f};
  for (int g = 0; g < 6; ++g) cost.w[g] = w[g];

  LaneParams lp;
  lp.base = make_base_inertia(CountF::input(6.921f), CountF::input(0.021112f), CountF::input(0.f),
                              CountF::input(-0.005366f), CountF::input(0.02448f), CountF::input(0.098077f),
                              CountF::input(0.107f), CountF::input(0.00012166f), CountF::input(0.0014849f),
                              CountF::input(-3.12e-05f));
  for (int l = 0; l < 4; ++l) {
    lp.kp_s[l] = CountF::input(1.f); lp.kd_s[l] = CountF::input(1.f); lp.fric[l] = CountF::input(0.1f);
  }
  for (int g = 0; g < 3; ++g) lp.motor[g] = CountF::input(20.f);

  const int n_traj = 8, settle = 50, steps = 50;   // 1 s of excitation per trajectory
  // segment records of the rollout kernels' layout cut from the excitation (runtime values):
  // the adjoint census replays them (the 22-parameter gradient workload: H 20, checkpoints
  // every 10 control steps)
  const int H = 20, every = 10;
  std::vector<CountF> recs;
  long long ctrl_flops = 0, contact_feet = 0, physics_steps = 0;
  float min_height = 1e9f;
  CountF tgt[37];
  for (int e = 0; e < 37; ++e) tgt[e] = CountF::input(0.01f * e);
  std::mt19937 rng(7);
  std::uniform_real_distribution<float> ua(0.1f, 0.4f), uf(0.5f, 3.0f), up(0.f, 6.2831853f);
  for (int tr = 0; tr < n_traj; ++tr) {
    State st;
    st.p = V3{CountF::input(0.f), CountF::input(0.f), CountF::input(0.30f)};
    st.qx = CountF::input(0.f); st.qy = CountF::input(0.f); st.qz = CountF::input(0.f);
    st.qw = CountF::input(1.f);
    st.w = V3{CountF::input(0.f), CountF::input(0.f), CountF::input(0.f)};
    st.v = st.w;
    for (int j = 0; j < 12; ++j) { st.q[j] = CountF::input(q0[j]); st.qd[j] = CountF::input(0.f); }
    float A[12], F[12], P[12];
    for (int j = 0; j < 12; ++j) { A[j] = ua(rng); F[j] = uf(rng); P[j] = up(rng); }
    for (int t = 0; t < settle + steps; ++t) {
      CountF act[12];
      for (int j = 0; j < 12; ++j)
        act[j] = CountF::input(t < settle ? 0.f
                                          : A[j] * std::sin(6.2831853f * F[j] * 0.02f * (t - settle) + P[j]) / 0.25f);
      long long before = census::g_add + census::g_mul + census::g_div + census::g_trans;
      for (int d = 0; d < a.decimation; ++d) {
        CountF tau[12];
        control_torques(a, lp, act, st, tau);
        CountF ff[12];
        physics_step(a, lp, st, tau, ff);
        if (t >= settle) {
          for (int f = 0; f < 4; ++f) contact_feet += (ff[3 * f + 2].v > 0.f);
          ++physics_steps;
        }
      }
      CountF c = step_cost(cost, st, tgt);
      (void)c;
      (void)state_norm(st);
      if (t == settle + 10) {                  // a record starts here: init state, then H x (action, target)
        CountF r[13], q[12], qd[12];
        state_to_root(st, r, q, qd, 1);
        for (int i = 0; i < 13; ++i) recs.push_back(CountF::input(r[i].v));
        for (int j = 0; j < 12; ++j) recs.push_back(CountF::input(q[j].v));
        for (int j = 0; j < 12; ++j) recs.push_back(CountF::input(qd[j].v));
        for (int k = 0; k < H; ++k) {
          const float tt = 0.02f * (t + 1 + k - settle);
          for (int j = 0; j < 12; ++j) recs.push_back(CountF::input(A[j] * std::sin(6.2831853f * F[j] * tt + P[j]) / 0.25f));
          for (int e = 0; e < 37; ++e) recs.push_back(CountF::input(0.01f * e));
        }
      }
      long long after = census::g_add + census::g_mul + census::g_div + census::g_trans;
      if (t >= settle) {
        ctrl_flops += after - before;
        min_height = std::fmin(min_height, st.p.z.v);
      }
    }
  }
  double per_ctrl = double(ctrl_flops) / (n_traj * steps);
  std::printf("{\"flops_per_control_step\": %.1f, \"flops_per_physics_step\": %.1f, "
              "\"flops_per_forward_dynamics\": %.1f, \"mean_feet_in_contact\": %.3f, "
              "\"add\": %lld, \"mul\": %lld, \"div\": %lld, \"trans\": %lld, \"control_steps\": %d}\n",
              per_ctrl, per_ctrl / a.decimation, per_ctrl / (a.decimation * a.substeps),
              double(contact_feet) / physics_steps, census::g_add, census::g_mul, census::g_div, census::g_trans,
              n_traj * steps);
  std::printf("lowest base height during excitation %.4f m\n", min_height);

  // reverse mode: cost + gradient with respect to all 25 parameter rows of one rollout
  // (spi_adj.h rollout_grad: forward pass with checkpoints, block recompute, physics-step and
  // sub-interval recompute, sub-interval adjoints), per control step
  const int stride = SPI_STATE_DIM + H * (12 + SPI_STATE_DIM);
  const int n_rec = int(recs.size()) / stride;
  const CountF raw[4] = {lp.base.m, CountF::input(0.021112f), CountF::input(0.f), CountF::input(-0.005366f)};
  census::g_add = census::g_mul = census::g_div = census::g_trans = 0;
  for (int r = 0; r < n_rec; ++r) {
    census::CountStore store;
    store.s.resize((H + every - 1) / every + every);
    State ps[8], ss[8];
    ParBar pb;
    CountF cst(0.f);
    rollout_grad(a, cost, lp, raw, recs.data() + size_t(r) * stride, H, every, store, ps, ss, cst, pb);
  }
  const double adj = double(census::g_add + census::g_mul + census::g_div + census::g_trans) / (n_rec * H);
  std::printf("{\"adjoint_flops_per_control_step\": %.1f, \"adjoint_over_primal\": %.2f, \"records\": %d, "
              "\"horizon\": %d, \"checkpoint_every\": %d}\n", adj, adj / per_ctrl, n_rec, H, every);
  return 0;
}
