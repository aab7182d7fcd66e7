// Differentiable rollouts for active exploration and Gauss-Newton refinement (gfx950), four
// lanes per tangent unit (lane % 4 = leg, spi_dyn_leg.h): trajectory sensitivities
// (spi_traj_sensitivity_legs) and weighted tracking residuals with their Jacobian
// (spi_residual_jacobian_legs).  The dynamics of spi_dyn.h instantiated with forward-mode
// arithmetic (spi_dual.h), one tangent direction per unit.  Built with -ffinite-math-only
// (divergence is caught on exponent bits) and with reassociation; the one-lane-per-tangent
// forms are in spi_sens_lane.hip.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>

#include "spi_dual.h"
#include "spi_types.h"

using DualT = spi_ad::Dual<1>;
#define SPI_REAL_T DualT
#define SPI_REAL_HAS_MATH
#define SPI_DYN_NS spi_tan
SPI_DUAL_MATH(spi_tan, 1)
#include "spi_dyn.h"
#include "spi_dyn_leg.h"
#include "spi_launch.h"

using namespace spi_tan;
using spi_launch::finite_bits;
using spi_launch::apply_segment_gains;
using spi_launch::xcd_remap;

#include "spi_sens_common.h"

namespace {

// traj_sens_kernel with four lanes per (group, tangent), one per leg (spi_dyn_leg.h): the
// tangent state of a lane is a leg and the base instead of the whole robot, so the dual
// arithmetic fits the registers.  Block b of excitation e covers tangent units
// [64 b, 64 b + 64) of that excitation, unit u = sample (u / n_tan), tangent (u % n_tan).
template <class T> __device__ __forceinline__ T pick4(int l, const T& x0, const T& x1, const T& x2, const T& x3) {
  return l == 0 ? x0 : (l == 1 ? x1 : (l == 2 ? x2 : x3));
}
__global__ __launch_bounds__(kSensThreads) void traj_sens_leg_kernel(SpiSimArgs a, SpiTangentSpec tan,
                                                                     const float* __restrict__ par, int ld,
                                                                     int n_samples, const float* __restrict__ init,
                                                                     const float* __restrict__ actions, int H,
                                                                     int blocks_per_exc, float* __restrict__ J,
                                                                     float* __restrict__ Y, int* __restrict__ status) {
  extern __shared__ float lds[];   // [37 init][H * 12 actions]
  const int logical = xcd_remap(blockIdx.x, gridDim.x);
  const int e = logical / blocks_per_exc, sb = logical % blocks_per_exc;
  const float* src = actions + static_cast<size_t>(e) * H * 12;
  for (int i = threadIdx.x; i < 37; i += kSensThreads) lds[i] = init[i];
  for (int i = threadIdx.x; i < H * 12; i += kSensThreads) lds[37 + i] = src[i];
  __syncthreads();
  const int P = tan.n_tan;
  const long long u = static_cast<long long>(sb) * (kSensThreads / 4) + (threadIdx.x >> 2);
  const int l = threadIdx.x & 3;
  if (u >= static_cast<long long>(n_samples) * P) return;  // whole 4-lane units leave together
  const int c = static_cast<int>(u / P), k = static_cast<int>(u % P);
  const size_t g = static_cast<size_t>(e) * n_samples + c;
  const LaneParams full = load_params_tan(par, ld, c, tan.param[k]);
  LegParams lp;
  lp.base = full.base;
  lp.kp_s = pick4(l, full.kp_s[0], full.kp_s[1], full.kp_s[2], full.kp_s[3]);
  lp.kd_s = pick4(l, full.kd_s[0], full.kd_s[1], full.kd_s[2], full.kd_s[3]);
  lp.fric = pick4(l, full.fric[0], full.fric[1], full.fric[2], full.fric[3]);
#pragma unroll
  for (int q = 0; q < 3; ++q) lp.motor[q] = full.motor[q];
  const LegK K = leg_constants(l);
  float kp[3], kd[3], q0[3], eff[3];
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    kp[j] = pick4(l, a.kp[j], a.kp[3 + j], a.kp[6 + j], a.kp[9 + j]);
    kd[j] = pick4(l, a.kd[j], a.kd[3 + j], a.kd[6 + j], a.kd[9 + j]);
    q0[j] = pick4(l, a.q0[j], a.q0[3 + j], a.q0[6 + j], a.q0[9 + j]);
    eff[j] = pick4(l, a.effort[j], a.effort[3 + j], a.effort[6 + j], a.effort[9 + j]);
  }
  LegState st;
  leg_state_from_root(st, lds, l);
  float* Jg = J + g * static_cast<size_t>(H) * kFimDim * P;
  bool ok = true;
  for (int t = 0; t < H; ++t) {
    float* row = Jg + static_cast<size_t>(t) * kFimDim * P + k;
    if (ok) {
      DualT act[3];
#pragma unroll
      for (int j = 0; j < 3; ++j) act[j] = DualT(lds[37 + t * 12 + 3 * l + j]);
      for (int d = 0; d < a.decimation; ++d) {
        DualT tau[3];
        control_torques_leg(a, lp, kp, kd, q0, eff, act, st.q, st.qd, tau);
        physics_step_leg(a, lp, K, st, tau);
      }
      const float n = state_norm_leg(st).v;
      ok = finite_bits(n) && n < a.state_limit;
    }
    float* y = (Y && k == 0) ? Y + (g * H + t) * kFimDim : nullptr;
    if (ok) {
      if (l == 0) {
        DualT r[13];
        leg_root(st, r);
#pragma unroll
        for (int i = 0; i < 13; ++i) row[i * P] = r[i].d[0];
        if (y)
#pragma unroll
          for (int i = 0; i < 13; ++i) y[i] = r[i].v;
      }
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        row[(13 + 3 * l + j) * P] = st.q[j].d[0];
        if (y) y[13 + 3 * l + j] = st.q[j].v;
      }
    } else {
      if (l == 0) {
#pragma unroll
        for (int i = 0; i < 13; ++i) row[i * P] = 0.0f;
        if (y)
#pragma unroll
          for (int i = 0; i < 13; ++i) y[i] = 0.0f;
      }
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        row[(13 + 3 * l + j) * P] = 0.0f;
        if (y) y[13 + 3 * l + j] = 0.0f;
      }
    }
  }
  if (k == 0 && l == 0) status[g] = ok ? 0 : 1;
}

// residual_sens_kernel with four lanes per (group, tangent) unit, one per leg (spi_dyn_leg.h):
// unit u = g * (n_tan + 1) + k, lanes 4u .. 4u + 3.  Lane 0 of a unit writes the 13 root rows,
// lane l its leg's joint rows 13 + 3l .. and 25 + 3l ..; the same matrix as the one-lane kernel
// up to the order of the legs' sums.  For the few hundred groups of a refinement step, whose
// one-lane form leaves most SIMDs idle.
__global__ __launch_bounds__(kSensThreads) void residual_sens_leg_kernel(SpiSimArgs a, SpiCostArgs cost,
                                                                         SpiTangentSpec tan,
                                                                         const float* __restrict__ par, int ld,
                                                                         int n_cand, const float* __restrict__ seg,
                                                                         const float* __restrict__ seg_w, int n_seg,
                                                                         int H, int stride, float* __restrict__ Jr,
                                                                         int* __restrict__ status) {
  const int P = tan.n_tan, L = P + 1;
  const long long u = static_cast<long long>(blockIdx.x) * (kSensThreads / 4) + (threadIdx.x >> 2);
  const int l = threadIdx.x & 3;
  if (u >= static_cast<long long>(n_cand) * n_seg * L) return;  // whole 4-lane units leave together
  const long long g = u / L;
  const int k = static_cast<int>(u % L);
  const int c = static_cast<int>(g / n_seg), s = static_cast<int>(g % n_seg);
  apply_segment_gains(a, cost, s);
  const float* rec0 = seg + static_cast<size_t>(s) * stride;
  const LaneParams full = load_params_tan(par, ld, c, k < P ? tan.param[k] : -1);
  LegParams lp;
  lp.base = full.base;
  lp.kp_s = pick4(l, full.kp_s[0], full.kp_s[1], full.kp_s[2], full.kp_s[3]);
  lp.kd_s = pick4(l, full.kd_s[0], full.kd_s[1], full.kd_s[2], full.kd_s[3]);
  lp.fric = pick4(l, full.fric[0], full.fric[1], full.fric[2], full.fric[3]);
#pragma unroll
  for (int q = 0; q < 3; ++q) lp.motor[q] = full.motor[q];
  const LegK K = leg_constants(l);
  float kp[3], kd[3], q0[3], eff[3];
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    kp[j] = pick4(l, a.kp[j], a.kp[3 + j], a.kp[6 + j], a.kp[9 + j]);
    kd[j] = pick4(l, a.kd[j], a.kd[3 + j], a.kd[6 + j], a.kd[9 + j]);
    q0[j] = pick4(l, a.q0[j], a.q0[3 + j], a.q0[6 + j], a.q0[9 + j]);
    eff[j] = pick4(l, a.effort[j], a.effort[3 + j], a.effort[6 + j], a.effort[9 + j]);
  }
  LegState st;
  leg_state_from_root(st, rec0, l);
  float sw[6];
  const float ws = seg_w[s];
#pragma unroll
  for (int q = 0; q < 6; ++q) sw[q] = sqrtf(fmaxf(cost.w[q] * ws, 0.0f));
  float* Jg = Jr + static_cast<size_t>(g) * H * SPI_STATE_DIM * L + k;
  bool ok = true;
  for (int t = 0; t < H; ++t) {
    const float* rec = rec0 + SPI_STATE_DIM + t * (12 + SPI_STATE_DIM);
    float* row = Jg + static_cast<size_t>(t) * SPI_STATE_DIM * L;
    if (ok) {
      DualT act[3];
#pragma unroll
      for (int j = 0; j < 3; ++j) act[j] = DualT(rec[3 * l + j]);
      for (int d = 0; d < a.decimation; ++d) {
        DualT tau[3];
        control_torques_leg(a, lp, kp, kd, q0, eff, act, st.q, st.qd, tau);
        physics_step_leg(a, lp, K, st, tau);
      }
      const float n = state_norm_leg(st).v;   // the same in the four lanes
      ok = finite_bits(n) && n < a.state_limit;
    }
    const float* tg = rec + 12;
    if (ok) {
      if (l == 0) {
        DualT r[13];
        leg_root(st, r);
#pragma unroll
        for (int i = 0; i < 13; ++i) {
          const int grp = i < 3 ? 0 : (i < 7 ? 1 : (i < 10 ? 2 : 3));
          row[i * L] = sw[grp] * (k < P ? r[i].d[0] : r[i].v - tg[i]);
        }
      }
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        const int iq = 13 + 3 * l + j, iv = 25 + 3 * l + j;
        row[iq * L] = sw[4] * (k < P ? st.q[j].d[0] : st.q[j].v - tg[iq]);
        row[iv * L] = sw[5] * (k < P ? st.qd[j].d[0] : st.qd[j].v - tg[iv]);
      }
    } else {
      if (l == 0) {
#pragma unroll
        for (int i = 0; i < 13; ++i) row[i * L] = 0.0f;
      }
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        row[(13 + 3 * l + j) * L] = 0.0f;
        row[(25 + 3 * l + j) * L] = 0.0f;
      }
    }
  }
  if (k == 0 && l == 0) status[g] = ok ? 0 : 1;
}

}  // namespace

extern "C" {

int spi_traj_sensitivity_legs(const SpiSimArgs* sim, const SpiTangentSpec* tan, const float* params_dev, int ld,
                              int n_samples, const float* init_dev, const float* actions_dev, int n_exc, int horizon,
                              float* J_dev, float* y_dev, int* status_dev, hipStream_t stream) {
  if (!sim_ok(sim)) return sfail(SPI_ERR_INVALID_ARG, "spi_traj_sensitivity_legs: invalid SpiSimArgs");
  if (!tan_ok(tan)) return sfail(SPI_ERR_INVALID_ARG, "spi_traj_sensitivity_legs: bad tangent spec");
  if (n_samples < 0 || ld < n_samples || n_exc < 0 || horizon < 1 || !params_dev || !init_dev || !actions_dev ||
      !J_dev || !status_dev)
    return sfail(SPI_ERR_INVALID_ARG, "spi_traj_sensitivity_legs: bad sizes or null pointer");
  if (n_samples == 0 || n_exc == 0) return SPI_OK;
  const size_t lds = (37 + static_cast<size_t>(horizon) * 12) * sizeof(float);
  if (lds > 64 * 1024) return sfail(SPI_ERR_TOO_LARGE, "spi_traj_sensitivity_legs: horizon too long for LDS staging");
  const long long units = static_cast<long long>(n_samples) * tan->n_tan;
  const long long bpe = (units + kSensThreads / 4 - 1) / (kSensThreads / 4);
  const long long blocks = bpe * n_exc;
  if (blocks > 0x7fffffffLL) return sfail(SPI_ERR_TOO_LARGE, "spi_traj_sensitivity_legs: grid too large");
  traj_sens_leg_kernel<<<static_cast<unsigned>(blocks), kSensThreads, lds, stream>>>(
      *sim, *tan, params_dev, ld, n_samples, init_dev, actions_dev, horizon, static_cast<int>(bpe), J_dev, y_dev,
      status_dev);
  return slaunch("traj_sens_leg_kernel");
}

int spi_residual_jacobian_legs(const SpiSimArgs* sim, const SpiCostArgs* cost, const SpiTangentSpec* tan,
                               const float* params_dev, int ld, int n_cand, const float* segments_dev,
                               const float* seg_weight_dev, int n_seg, int horizon, float* Jr_dev, int* status_dev,
                               hipStream_t stream) {
  if (!sim_ok(sim)) return sfail(SPI_ERR_INVALID_ARG, "spi_residual_jacobian_legs: invalid SpiSimArgs");
  if (!tan_ok(tan) || tan->n_tan + 1 > SPI_MAX_TANGENTS)
    return sfail(SPI_ERR_INVALID_ARG, "spi_residual_jacobian_legs: bad tangent spec (n_tan + 1 <= 32)");
  if (!cost || cost->mode != SPI_COST_TRACKING)
    return sfail(SPI_ERR_INVALID_ARG, "spi_residual_jacobian_legs: needs the tracking (sum of squares) cost");
  if (n_cand < 0 || ld < n_cand || n_seg < 0 || horizon < 1 || !params_dev || !segments_dev || !seg_weight_dev ||
      !Jr_dev || !status_dev)
    return sfail(SPI_ERR_INVALID_ARG, "spi_residual_jacobian_legs: bad sizes or null pointer");
  const long long G = static_cast<long long>(n_cand) * n_seg;
  if (G == 0) return SPI_OK;
  const int stride = SPI_STATE_DIM + horizon * (12 + SPI_STATE_DIM);
  const long long units = G * (tan->n_tan + 1);
  const long long blocks = (units + kSensThreads / 4 - 1) / (kSensThreads / 4);
  if (blocks > 0x7fffffffLL) return sfail(SPI_ERR_TOO_LARGE, "spi_residual_jacobian_legs: grid too large");
  residual_sens_leg_kernel<<<static_cast<unsigned>(blocks), kSensThreads, 0, stream>>>(
      *sim, *cost, *tan, params_dev, ld, n_cand, segments_dev, seg_weight_dev, n_seg, horizon, stride, Jr_dev,
      status_dev);
  return slaunch("residual_sens_leg_kernel");
}

}  // extern "C"
