// Differentiable rollouts with one lane per tangent (gfx950): the dynamics of spi_dyn.h
// instantiated with forward-mode arithmetic (spi_dual.h); one lane carries one tangent
// direction (one identified parameter) of one (parameter sample, excitation) group, so a group
// occupies n_tan neighbouring lanes and recomputes the same primal values.  Trajectory
// sensitivities, cost gradients (the same tangent lanes accumulate the tracking cost's
// derivative) and residual Jacobians.  Built with -ffinite-math-only and without
// reassociation: these kernels already spill, and reassociation keeps more values live.
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

// One block = gpb groups of n_tan lanes, all groups of one excitation e (its actions staged
// in LDS), consecutive parameter samples.
__global__ __launch_bounds__(kSensThreads) void traj_sens_kernel(SpiSimArgs a, SpiTangentSpec tan,
                                                                 const float* __restrict__ par, int ld, int n_samples,
                                                                 const float* __restrict__ init,
                                                                 const float* __restrict__ actions, int H, int gpb,
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
  const int gl = threadIdx.x / P, k = threadIdx.x % P;
  const int c = sb * gpb + gl;
  if (gl >= gpb || c >= n_samples) return;
  const size_t g = static_cast<size_t>(e) * n_samples + c;
  const LaneParams lp = load_params_tan(par, ld, c, tan.param[k]);
  State st;
  init_state(st, lds);
  float* Jg = J + g * static_cast<size_t>(H) * kFimDim * P;
  bool ok = true;
  for (int t = 0; t < H; ++t) {
    float* row = Jg + static_cast<size_t>(t) * kFimDim * P + k;
    if (ok) {
      DualT act[12];
#pragma unroll
      for (int j = 0; j < 12; ++j) act[j] = DualT(lds[37 + t * 12 + j]);
      for (int d = 0; d < a.decimation; ++d) {
        DualT tau[12];
        control_torques(a, lp, act, st, tau);
        physics_step(a, lp, st, tau, nullptr);
      }
      ok = state_ok(st, a.state_limit);
    }
    if (ok) {
      DualT r[13], q[12], qd[12];
      state_to_root(st, r, q, qd, 1);
#pragma unroll
      for (int i = 0; i < 13; ++i) row[i * P] = r[i].d[0];
#pragma unroll
      for (int j = 0; j < 12; ++j) row[(13 + j) * P] = q[j].d[0];
      if (Y && k == 0) {
        float* y = Y + (g * H + t) * kFimDim;
#pragma unroll
        for (int i = 0; i < 13; ++i) y[i] = r[i].v;
#pragma unroll
        for (int j = 0; j < 12; ++j) y[13 + j] = q[j].v;
      }
    } else {
#pragma unroll
      for (int i = 0; i < kFimDim; ++i) row[i * P] = 0.0f;
      if (Y && k == 0) {
        float* y = Y + (g * H + t) * kFimDim;
#pragma unroll
        for (int i = 0; i < kFimDim; ++i) y[i] = 0.0f;
      }
    }
  }
  if (k == 0) status[g] = ok ? 0 : 1;
}

// Cost + gradient: one block = gpb candidate groups of one segment (record staged in LDS).
__global__ __launch_bounds__(kSensThreads) void cost_grad_kernel(SpiSimArgs a, SpiCostArgs cost, SpiTangentSpec tan,
                                                                 const float* __restrict__ par, int ld, int n_cand,
                                                                 const float* __restrict__ seg, int H, int stride,
                                                                 int gpb, int blocks_per_seg, float* __restrict__ out,
                                                                 float* __restrict__ grad) {
  extern __shared__ float lds[];
  const int logical = xcd_remap(blockIdx.x, gridDim.x);
  const int s = logical / blocks_per_seg, cb = logical % blocks_per_seg;
  const float* src = seg + static_cast<size_t>(s) * stride;
  for (int i = threadIdx.x; i < stride; i += kSensThreads) lds[i] = src[i];
  __syncthreads();
  const int P = tan.n_tan;
  const int gl = threadIdx.x / P, k = threadIdx.x % P;
  const int c = cb * gpb + gl;
  if (gl >= gpb || c >= n_cand) return;
  apply_segment_gains(a, cost, s);
  const LaneParams lp = load_params_tan(par, ld, c, tan.param[k]);
  State st;
  init_state(st, lds);
  DualT acc(0.0f);
  bool ok = true;
  for (int t = 0; t < H; ++t) {
    const float* rec = lds + SPI_STATE_DIM + t * (12 + SPI_STATE_DIM);
    DualT act[12];
#pragma unroll
    for (int j = 0; j < 12; ++j) act[j] = DualT(rec[j]);
    for (int d = 0; d < a.decimation; ++d) {
      DualT tau[12];
      control_torques(a, lp, act, st, tau);
      physics_step(a, lp, st, tau, nullptr);
    }
    if (cost.mode == SPI_COST_TRACKING || t == H - 1) acc += step_cost(cost, st, rec + 12);
    if (!state_ok(st, a.state_limit)) { ok = false; break; }
  }
  ok = ok && finite_bits(acc.v) && finite_bits(acc.d[0]);
  const size_t o = static_cast<size_t>(s) * n_cand + c;
  grad[o * P + k] = ok ? acc.d[0] : 0.0f;
  if (k == 0) reinterpret_cast<unsigned int*>(out)[o] = ok ? __float_as_uint(acc.v) : 0x7f800000u;
}

// Weighted tracking residuals and their Jacobian for Gauss-Newton refinement.  Group
// g = c * n_seg + s (candidate c, segment s) has n_tan + 1 lanes: lane k < n_tan carries the
// tangent of parameter row tan.param[k], lane n_tan carries no tangent and writes the residual
// itself, so row t * 37 + i of the group's matrix is [d r / d theta | r] with
// r_i = sqrt(w_group(i) * seg_w) * (x_i - target_i) over the 37-float state after step t
// (x = [pos, quat xyzw, linvel, angvel (world), q, qd]; the tracking cost is sum r^2).
__global__ __launch_bounds__(kSensThreads) void residual_sens_kernel(SpiSimArgs a, SpiCostArgs cost,
                                                                     SpiTangentSpec tan, const float* __restrict__ par,
                                                                     int ld, int n_cand,
                                                                     const float* __restrict__ seg,
                                                                     const float* __restrict__ seg_w, int n_seg, int H,
                                                                     int stride, int gpb, float* __restrict__ Jr,
                                                                     int* __restrict__ status) {
  const int P = tan.n_tan, L = P + 1;
  const int gl = threadIdx.x / L, k = threadIdx.x % L;
  const long long g = static_cast<long long>(blockIdx.x) * gpb + gl;
  if (gl >= gpb || g >= static_cast<long long>(n_cand) * n_seg) return;
  const int c = static_cast<int>(g / n_seg), s = static_cast<int>(g % n_seg);
  apply_segment_gains(a, cost, s);
  const float* rec0 = seg + static_cast<size_t>(s) * stride;
  const LaneParams lp = load_params_tan(par, ld, c, k < P ? tan.param[k] : -1);
  State st;
  init_state(st, rec0);
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
      DualT act[12];
#pragma unroll
      for (int j = 0; j < 12; ++j) act[j] = DualT(rec[j]);
      for (int d = 0; d < a.decimation; ++d) {
        DualT tau[12];
        control_torques(a, lp, act, st, tau);
        physics_step(a, lp, st, tau, nullptr);
      }
      ok = state_ok(st, a.state_limit);
    }
    if (ok) {
      DualT r[13], q[12], qd[12];
      state_to_root(st, r, q, qd, 1);
      const float* tg = rec + 12;
#pragma unroll
      for (int i = 0; i < SPI_STATE_DIM; ++i) {
        const DualT& x = i < 13 ? r[i] : (i < 25 ? q[i - 13] : qd[i - 25]);
        const int grp = i < 3 ? 0 : (i < 7 ? 1 : (i < 10 ? 2 : (i < 13 ? 3 : (i < 25 ? 4 : 5))));
        row[i * L] = sw[grp] * (k < P ? x.d[0] : x.v - tg[i]);
      }
    } else {
#pragma unroll
      for (int i = 0; i < SPI_STATE_DIM; ++i) row[i * L] = 0.0f;
    }
  }
  if (k == 0) status[g] = ok ? 0 : 1;
}

}  // namespace

extern "C" {

int spi_traj_sensitivity(const SpiSimArgs* sim, const SpiTangentSpec* tan, const float* params_dev, int ld,
                         int n_samples, const float* init_dev, const float* actions_dev, int n_exc, int horizon,
                         float* J_dev, float* y_dev, int* status_dev, hipStream_t stream) {
  if (!sim_ok(sim)) return sfail(SPI_ERR_INVALID_ARG, "spi_traj_sensitivity: invalid SpiSimArgs");
  if (!tan_ok(tan)) return sfail(SPI_ERR_INVALID_ARG, "spi_traj_sensitivity: bad tangent spec");
  if (n_samples < 0 || ld < n_samples || n_exc < 0 || horizon < 1 || !params_dev || !init_dev || !actions_dev ||
      !J_dev || !status_dev)
    return sfail(SPI_ERR_INVALID_ARG, "spi_traj_sensitivity: bad sizes or null pointer");
  if (n_samples == 0 || n_exc == 0) return SPI_OK;
  const size_t lds = (37 + static_cast<size_t>(horizon) * 12) * sizeof(float);
  if (lds > 64 * 1024) return sfail(SPI_ERR_TOO_LARGE, "spi_traj_sensitivity: horizon too long for LDS staging");
  const int gpb = kSensThreads / tan->n_tan;
  const int bpe = (n_samples + gpb - 1) / gpb;
  const long long blocks = static_cast<long long>(bpe) * n_exc;
  if (blocks > 0x7fffffffLL) return sfail(SPI_ERR_TOO_LARGE, "spi_traj_sensitivity: grid too large");
  traj_sens_kernel<<<static_cast<unsigned>(blocks), kSensThreads, lds, stream>>>(
      *sim, *tan, params_dev, ld, n_samples, init_dev, actions_dev, horizon, gpb, bpe, J_dev, y_dev, status_dev);
  return slaunch("traj_sens_kernel");
}

int spi_rollout_cost_grad(const SpiSimArgs* sim, const SpiCostArgs* cost, const SpiTangentSpec* tan,
                          const float* params_dev, int ld, int n_cand, const float* segments_dev, int n_seg,
                          int horizon, float* cost_dev, float* grad_dev, hipStream_t stream) {
  if (!sim_ok(sim)) return sfail(SPI_ERR_INVALID_ARG, "spi_rollout_cost_grad: invalid SpiSimArgs");
  if (!tan_ok(tan)) return sfail(SPI_ERR_INVALID_ARG, "spi_rollout_cost_grad: bad tangent spec");
  if (!cost || n_cand < 0 || ld < n_cand || n_seg < 0 || horizon < 1 || !params_dev || !segments_dev || !cost_dev ||
      !grad_dev)
    return sfail(SPI_ERR_INVALID_ARG, "spi_rollout_cost_grad: bad sizes or null pointer");
  if (n_cand == 0 || n_seg == 0) return SPI_OK;
  const int stride = SPI_STATE_DIM + horizon * (12 + SPI_STATE_DIM);
  const size_t lds = static_cast<size_t>(stride) * sizeof(float);
  if (lds > 64 * 1024) return sfail(SPI_ERR_TOO_LARGE, "spi_rollout_cost_grad: horizon too long for LDS staging");
  const int gpb = kSensThreads / tan->n_tan;
  const int bps = (n_cand + gpb - 1) / gpb;
  const long long blocks = static_cast<long long>(bps) * n_seg;
  if (blocks > 0x7fffffffLL) return sfail(SPI_ERR_TOO_LARGE, "spi_rollout_cost_grad: grid too large");
  cost_grad_kernel<<<static_cast<unsigned>(blocks), kSensThreads, lds, stream>>>(
      *sim, *cost, *tan, params_dev, ld, n_cand, segments_dev, horizon, stride, gpb, bps, cost_dev, grad_dev);
  return slaunch("cost_grad_kernel");
}

int spi_residual_jacobian(const SpiSimArgs* sim, const SpiCostArgs* cost, const SpiTangentSpec* tan,
                          const float* params_dev, int ld, int n_cand, const float* segments_dev,
                          const float* seg_weight_dev, int n_seg, int horizon, float* Jr_dev, int* status_dev,
                          hipStream_t stream) {
  if (!sim_ok(sim)) return sfail(SPI_ERR_INVALID_ARG, "spi_residual_jacobian: invalid SpiSimArgs");
  if (!tan_ok(tan) || tan->n_tan + 1 > SPI_MAX_TANGENTS)
    return sfail(SPI_ERR_INVALID_ARG, "spi_residual_jacobian: bad tangent spec (n_tan + 1 <= 32)");
  if (!cost || cost->mode != SPI_COST_TRACKING)
    return sfail(SPI_ERR_INVALID_ARG, "spi_residual_jacobian: needs the tracking (sum of squares) cost");
  if (n_cand < 0 || ld < n_cand || n_seg < 0 || horizon < 1 || !params_dev || !segments_dev || !seg_weight_dev ||
      !Jr_dev || !status_dev)
    return sfail(SPI_ERR_INVALID_ARG, "spi_residual_jacobian: bad sizes or null pointer");
  const long long G = static_cast<long long>(n_cand) * n_seg;
  if (G == 0) return SPI_OK;
  const int stride = SPI_STATE_DIM + horizon * (12 + SPI_STATE_DIM);
  const int gpb = kSensThreads / (tan->n_tan + 1);
  const long long blocks = (G + gpb - 1) / gpb;
  if (blocks > 0x7fffffffLL) return sfail(SPI_ERR_TOO_LARGE, "spi_residual_jacobian: grid too large");
  residual_sens_kernel<<<static_cast<unsigned>(blocks), kSensThreads, 0, stream>>>(
      *sim, *cost, *tan, params_dev, ld, n_cand, segments_dev, seg_weight_dev, n_seg, horizon, stride, gpb, Jr_dev,
      status_dev);
  return slaunch("residual_sens_kernel");
}

}  // extern "C"
