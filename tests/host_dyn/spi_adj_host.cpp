// Test-only host build of the rollout adjoint (spi_adj.h) so the reverse pass can be checked
// against the FP64 oracle's forward-mode gradient and against its own finite differences on
// machines without a GPU.  Built in float (the device arithmetic) or, with -DSPI_HOST_DOUBLE,
// in double with the lane math routed to double precision.
#include <cmath>
#include <vector>

#ifdef SPI_HOST_DOUBLE
#define SPI_REAL_T double
namespace spi {
inline double fminf(double a, double b) { return std::fmin(a, b); }
inline double fmaxf(double a, double b) { return std::fmax(a, b); }
inline double fabsf(double a) { return std::fabs(a); }
inline double copysignf(double a, double b) { return std::copysign(a, b); }
inline double sqrtf(double a) { return std::sqrt(a); }
inline double expf(double a) { return std::exp(a); }
inline double sinf(double a) { return std::sin(a); }
inline double cosf(double a) { return std::cos(a); }
}  // namespace spi
#endif
#include "spi_adj.h"

using namespace spi;

namespace {
struct HostStore {
  std::vector<State> s;
  void put(int k, const State& x) { s[k] = x; }
  void get(int k, State& x) { x = s[k]; }
};
LaneParams make_lp(const double* par) {
  LaneParams lp;
  lp.base = make_base_inertia(Real(par[0]), Real(par[1]), Real(par[2]), Real(par[3]), Real(par[4]), Real(par[5]),
                              Real(par[6]), Real(par[7]), Real(par[8]), Real(par[9]));
  for (int l = 0; l < 4; ++l) {
    lp.kp_s[l] = Real(par[10 + l]); lp.kd_s[l] = Real(par[14 + l]); lp.fric[l] = Real(par[18 + l]);
  }
  for (int g = 0; g < 3; ++g) lp.motor[g] = Real(par[22 + g]);
  return lp;
}
}  // namespace

// cost of one rollout over the segment record rec0 (layout of the rollout kernels) and its
// gradient with respect to the 25 parameter rows; returns 0, or 1 if the rollout diverged.
extern "C" int spi_host_rollout_grad(const SpiSimArgs* a, const SpiCostArgs* c, const double* par,
                                     const float* rec0, int H, int every, double* cost, double* grad25) {
  const LaneParams lp = make_lp(par);
  const Real raw[4] = {Real(par[0]), Real(par[1]), Real(par[2]), Real(par[3])};
  const int nb = (H + every - 1) / every;
  HostStore st;
  st.s.resize(nb + every);
  State ps[8], ss[8];
  ParBar pb;
  Real cst(0.f);
  const bool ok = rollout_grad(*a, *c, lp, raw, rec0, H, every, st, ps, ss, cst, pb);
  *cost = double(cst);
  for (int k = 0; k < SPI_P_COUNT; ++k) grad25[k] = ok ? double(pb.g[k]) : 0.0;
  return ok ? 0 : 1;
}

// primal cost only (for finite differences of this build)
extern "C" double spi_host_rollout_cost(const SpiSimArgs* a, const SpiCostArgs* c, const double* par,
                                        const float* rec0, int H) {
  const LaneParams lp = make_lp(par);
  State st;
  Real r[13], q[12], qd[12];
  for (int i = 0; i < 13; ++i) r[i] = Real(rec0[i]);
  for (int j = 0; j < 12; ++j) { q[j] = Real(rec0[13 + j]); qd[j] = Real(rec0[25 + j]); }
  state_from_root(st, r, q, qd, 1);
  Real acc(0.f);
  for (int t = 0; t < H; ++t) {
    const float* rec = rec0 + SPI_STATE_DIM + t * (12 + SPI_STATE_DIM);
    Real act[12];
    for (int j = 0; j < 12; ++j) act[j] = Real(rec[j]);
    for (int d = 0; d < a->decimation; ++d) {
      Real tau[12];
      control_torques(*a, lp, act, st, tau);
      physics_step(*a, lp, st, tau, nullptr);
    }
    if (c->mode == SPI_COST_TRACKING || t == H - 1) acc += step_cost(*c, st, rec + 12);
  }
  return double(acc);
}
