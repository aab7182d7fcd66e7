// Launch-shape and bit-level helpers shared by the kernel translation units.
#pragma once
#include <hip/hip_runtime.h>

#include "spi_hip.h"

namespace spi_launch {

// Recorded PD gains of segment s: when the cost settings carry a per-segment gain table, the
// kernel's copy of the simulation settings runs the PD law with that segment's (kp, kd) on
// all 12 joints (scripts/eval.py:192-201 sets env.p_gains/d_gains from the recorded
// pd_gain_kp/kd).  s is uniform over the workgroup, so the gains stay in scalar registers.
__device__ __forceinline__ void apply_segment_gains(SpiSimArgs& a, const SpiCostArgs& c, int s) {
  if (c.seg_gains != nullptr) {
    const float kp = c.seg_gains[2 * s], kd = c.seg_gains[2 * s + 1];
#pragma unroll
    for (int j = 0; j < 12; ++j) {
      a.kp[j] = kp;
      a.kd[j] = kd;
    }
  }
}

// Workgroups are dispatched round-robin over the 8 XCDs of an MI355X.  Remap the hardware
// block id so that logically consecutive blocks (same segment or excitation, neighbouring
// candidate blocks) run on the same XCD and share its L2.
__device__ __forceinline__ int xcd_remap(int b, int nb) {
  constexpr int NX = 8;
  int q = nb / NX, r = nb % NX;
  int x = b % NX, i = b / NX;
  int base = x < r ? x * (q + 1) : r * (q + 1) + (x - r) * q;
  return base + i;
}

// Finiteness from the exponent bits, for files compiled with -ffinite-math-only (which lets
// the compiler fold the model's many x*0 terms).  Under that flag the compiler may assume that
// no float is NaN or infinite and fold a plain bit test of one to "finite"; the bits therefore
// pass through an empty asm first, whose output the compiler knows nothing about.
__device__ __forceinline__ bool finite_bits(float x) {
  unsigned int u = __float_as_uint(x);
  asm volatile("" : "+v"(u));
  return (u & 0x7f800000u) != 0x7f800000u;
}

// A candidate's parameters describe a robot the dynamics can integrate: every one of the 25
// entries finite, base mass > 0, the base link's rotational inertia about its CoM positive
// definite (leading minors of [ixx ixy ixz; ixy iyy iyz; ixz iyz izz]), and positive motor
// terms where the motor model divides by them (A tanh(tau / A)).  A rollout of an invalid
// candidate is not run: its cost is +inf (gradient 0), and no other lane is affected.
__device__ __forceinline__ bool lane_params_valid(const float* __restrict__ par, int ld, int i, int motor_model) {
  bool ok = true;
#pragma unroll
  for (int p = 0; p < SPI_P_COUNT; ++p) ok = ok && finite_bits(par[p * ld + i]);
  const float m = par[0 * ld + i], ixx = par[4 * ld + i], iyy = par[5 * ld + i], izz = par[6 * ld + i];
  const float ixy = par[7 * ld + i], ixz = par[8 * ld + i], iyz = par[9 * ld + i];
  const float m2 = ixx * iyy - ixy * ixy;
  const float det = ixx * (iyy * izz - iyz * iyz) - ixy * (ixy * izz - iyz * ixz) + ixz * (ixy * iyz - iyy * ixz);
  ok = ok && m > 0.0f && ixx > 0.0f && m2 > 0.0f && det > 0.0f;
  if (motor_model == SPI_MOTOR_VEC3_TANH)
    ok = ok && par[22 * ld + i] > 0.0f && par[23 * ld + i] > 0.0f && par[24 * ld + i] > 0.0f;
  return ok;
}

}  // namespace spi_launch
