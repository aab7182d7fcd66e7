// Test-only host build of the device dynamics (spi_dyn.h with float) so the specialised
// algorithm can be checked against the oracle on machines without a GPU.  Not part of the
// product: the library never runs this code on the CPU.
#include "spi_dyn.h"

extern "C" int spi_host_forward_dynamics(const SpiSimArgs* a, const float* par /* [25] */, const float* s37,
                                         const float* tau_motor, float* acc18, float* foot_force12) {
  using namespace spi;
  LaneParams lp;
  lp.base = make_base_inertia(par[0], par[1], par[2], par[3], par[4], par[5], par[6], par[7], par[8], par[9]);
  for (int l = 0; l < 4; ++l) { lp.kp_s[l] = par[10 + l]; lp.kd_s[l] = par[14 + l]; lp.fric[l] = par[18 + l]; }
  for (int g = 0; g < 3; ++g) lp.motor[g] = par[22 + g];
  State st;
  state_from_root(st, s37, s37 + 13, s37 + 25, 1);
  M6 a0;
  float qdd[12];
  forward_dynamics(*a, lp, st, tau_motor, a0, qdd, foot_force12);
  acc18[0] = a0.w.x; acc18[1] = a0.w.y; acc18[2] = a0.w.z;
  acc18[3] = a0.v.x; acc18[4] = a0.v.y; acc18[5] = a0.v.z;
  for (int j = 0; j < 12; ++j) acc18[6 + j] = qdd[j];
  return 0;
}

extern "C" int spi_host_control_step(const SpiSimArgs* a, const float* par, float* s37, const float* act12) {
  using namespace spi;
  LaneParams lp;
  lp.base = make_base_inertia(par[0], par[1], par[2], par[3], par[4], par[5], par[6], par[7], par[8], par[9]);
  for (int l = 0; l < 4; ++l) { lp.kp_s[l] = par[10 + l]; lp.kd_s[l] = par[14 + l]; lp.fric[l] = par[18 + l]; }
  for (int g = 0; g < 3; ++g) lp.motor[g] = par[22 + g];
  State st;
  state_from_root(st, s37, s37 + 13, s37 + 25, 1);
  for (int d = 0; d < a->decimation; ++d) {
    float tau[12];
    control_torques(*a, lp, act12, st, tau);
    physics_step(*a, lp, st, tau, nullptr);
  }
  state_to_root(st, s37, s37 + 13, s37 + 25, 1);
  return 0;
}
