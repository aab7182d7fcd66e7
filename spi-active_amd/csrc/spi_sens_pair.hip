// Trajectory sensitivities with two derivative directions per rollout (gfx950): the
// leg-parallel tangent rollout of spi_sens.hip (four lanes per unit, lane % 4 = leg) with a
// value and two tangents per lane, so one value path serves two parameters.  Per tangent it
// issues about a fifth fewer instructions than one tangent per unit; it needs about 500
// registers (one wavefront per SIMD).  Unit u = (sample, pair j) carries the tangents of
// parameter slots j and j + ceil(P / 2).  Same layout and semantics as spi_traj_sensitivity;
// every derivative component follows the same arithmetic as with one tangent per unit.
#include <hip/hip_runtime.h>

#include <cstdio>

#include "spi_dual.h"
#include "spi_types.h"

using DualT2 = spi_ad::Dual<2>;
#define SPI_REAL_T DualT2
#define SPI_REAL_HAS_MATH
#define SPI_DYN_NS spi_tan2
SPI_DUAL_MATH(spi_tan2, 2)
#include "spi_dyn.h"
#include "spi_dyn_leg.h"
#include "spi_launch.h"

using namespace spi_tan2;
using spi_launch::finite_bits;
using spi_launch::xcd_remap;

namespace {

int pair_fail(int code, const char* msg) {
  spi_set_error(msg);
  return code;
}

constexpr int kPairThreads = 256;
constexpr int kYDim = 25;   // y = root (13) + q (12)

template <class T> __device__ __forceinline__ T pick4(int l, const T& x0, const T& x1, const T& x2, const T& x3) {
  return l == 0 ? x0 : (l == 1 ? x1 : (l == 2 ? x2 : x3));
}

// Parameters of sample c with unit tangents on SoA rows seed0 (direction 0) and seed1
// (direction 1; -1: none).
__device__ __forceinline__ LaneParams load_params_pair(const float* __restrict__ par, int ld, int c, int seed0,
                                                       int seed1) {
  auto f = [&](int k) {
    DualT2 x(par[static_cast<size_t>(k) * ld + c]);
    x.d[0] = (k == seed0) ? 1.0f : 0.0f;
    x.d[1] = (k == seed1) ? 1.0f : 0.0f;
    return x;
  };
  LaneParams lp;
  lp.base = make_base_inertia(f(0), f(1), f(2), f(3), f(4), f(5), f(6), f(7), f(8), f(9));
#pragma unroll
  for (int l = 0; l < 4; ++l) {
    lp.kp_s[l] = f(10 + l);
    lp.kd_s[l] = f(14 + l);
    lp.fric[l] = f(18 + l);
  }
#pragma unroll
  for (int g = 0; g < 3; ++g) lp.motor[g] = f(22 + g);
  return lp;
}

// One block = 64 units of one excitation e (its initial state and actions staged in LDS).
__global__ __launch_bounds__(kPairThreads) void traj_sens_pair_kernel(SpiSimArgs a, SpiTangentSpec tan,
                                                                      const float* __restrict__ par, int ld,
                                                                      int n_samples, const float* __restrict__ init,
                                                                      const float* __restrict__ actions, int H,
                                                                      int blocks_per_exc, float* __restrict__ J,
                                                                      float* __restrict__ Y, int* __restrict__ status) {
  extern __shared__ float lds[];   // [37 init][H * 12 actions]
  const int logical = xcd_remap(blockIdx.x, gridDim.x);
  const int e = logical / blocks_per_exc, sb = logical % blocks_per_exc;
  const float* src = actions + static_cast<size_t>(e) * H * 12;
  for (int i = threadIdx.x; i < 37; i += kPairThreads) lds[i] = init[i];
  for (int i = threadIdx.x; i < H * 12; i += kPairThreads) lds[37 + i] = src[i];
  __syncthreads();
  const int P = tan.n_tan, Q = (P + 1) / 2;
  const long long u = static_cast<long long>(sb) * (kPairThreads / 4) + (threadIdx.x >> 2);
  const int l = threadIdx.x & 3;
  if (u >= static_cast<long long>(n_samples) * Q) return;  // whole 4-lane units leave together
  const int c = static_cast<int>(u / Q), j = static_cast<int>(u % Q);
  const int k0 = j, k1 = j + Q;
  const bool two = k1 < P;
  const size_t g = static_cast<size_t>(e) * n_samples + c;
  const LaneParams full = load_params_pair(par, ld, c, tan.param[k0], two ? tan.param[k1] : -1);
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
  for (int i = 0; i < 3; ++i) {
    kp[i] = pick4(l, a.kp[i], a.kp[3 + i], a.kp[6 + i], a.kp[9 + i]);
    kd[i] = pick4(l, a.kd[i], a.kd[3 + i], a.kd[6 + i], a.kd[9 + i]);
    q0[i] = pick4(l, a.q0[i], a.q0[3 + i], a.q0[6 + i], a.q0[9 + i]);
    eff[i] = pick4(l, a.effort[i], a.effort[3 + i], a.effort[6 + i], a.effort[9 + i]);
  }
  LegState st;
  leg_state_from_root(st, lds, l);
  float* Jg = J + g * static_cast<size_t>(H) * kYDim * P;
  bool ok = true;
  for (int t = 0; t < H; ++t) {
    float* row = Jg + static_cast<size_t>(t) * kYDim * P;
    if (ok) {
      DualT2 act[3];
#pragma unroll
      for (int i = 0; i < 3; ++i) act[i] = DualT2(lds[37 + t * 12 + 3 * l + i]);
      for (int d = 0; d < a.decimation; ++d) {
        DualT2 tau[3];
        control_torques_leg(a, lp, kp, kd, q0, eff, act, st.q, st.qd, tau);
        physics_step_leg(a, lp, K, st, tau);
      }
      const float n = state_norm_leg(st).v;   // the same in the four lanes
      ok = finite_bits(n) && n < a.state_limit;
    }
    float* y = (Y && j == 0) ? Y + (g * H + t) * kYDim : nullptr;
    if (ok) {
      if (l == 0) {
        DualT2 r[13];
        leg_root(st, r);
#pragma unroll
        for (int i = 0; i < 13; ++i) {
          row[i * P + k0] = r[i].d[0];
          if (two) row[i * P + k1] = r[i].d[1];
          if (y) y[i] = r[i].v;
        }
      }
#pragma unroll
      for (int i = 0; i < 3; ++i) {
        const int iq = 13 + 3 * l + i;
        row[iq * P + k0] = st.q[i].d[0];
        if (two) row[iq * P + k1] = st.q[i].d[1];
        if (y) y[iq] = st.q[i].v;
      }
    } else {
      if (l == 0) {
#pragma unroll
        for (int i = 0; i < 13; ++i) {
          row[i * P + k0] = 0.0f;
          if (two) row[i * P + k1] = 0.0f;
          if (y) y[i] = 0.0f;
        }
      }
#pragma unroll
      for (int i = 0; i < 3; ++i) {
        const int iq = 13 + 3 * l + i;
        row[iq * P + k0] = 0.0f;
        if (two) row[iq * P + k1] = 0.0f;
        if (y) y[iq] = 0.0f;
      }
    }
  }
  if (j == 0 && l == 0) status[g] = ok ? 0 : 1;
}

}  // namespace

extern "C" {

int spi_traj_sensitivity_pairs(const SpiSimArgs* sim, const SpiTangentSpec* tan, const float* params_dev, int ld,
                               int n_samples, const float* init_dev, const float* actions_dev, int n_exc, int horizon,
                               float* J_dev, float* y_dev, int* status_dev, hipStream_t stream) {
  if (!sim || sim->decimation < 1 || sim->substeps < 1 || !(sim->dt > 0.0f) || !(sim->state_limit > 0.0f))
    return pair_fail(SPI_ERR_INVALID_ARG, "spi_traj_sensitivity_pairs: invalid SpiSimArgs");
  bool tan_ok = tan && tan->n_tan >= 1 && tan->n_tan <= SPI_MAX_TANGENTS && tan->n_tan <= SPI_P_COUNT;
  for (int k = 0; tan_ok && k < tan->n_tan; ++k) tan_ok = tan->param[k] >= 0 && tan->param[k] < SPI_P_COUNT;
  if (!tan_ok) return pair_fail(SPI_ERR_INVALID_ARG, "spi_traj_sensitivity_pairs: bad tangent spec");
  if (n_samples < 0 || ld < n_samples || n_exc < 0 || horizon < 1 || !params_dev || !init_dev || !actions_dev ||
      !J_dev || !status_dev)
    return pair_fail(SPI_ERR_INVALID_ARG, "spi_traj_sensitivity_pairs: bad sizes or null pointer");
  if (n_samples == 0 || n_exc == 0) return SPI_OK;
  const size_t lds = (37 + static_cast<size_t>(horizon) * 12) * sizeof(float);
  if (lds > 64 * 1024)
    return pair_fail(SPI_ERR_TOO_LARGE, "spi_traj_sensitivity_pairs: horizon too long for LDS staging");
  const long long units = static_cast<long long>(n_samples) * ((tan->n_tan + 1) / 2);
  const long long bpe = (units + kPairThreads / 4 - 1) / (kPairThreads / 4);
  const long long blocks = bpe * n_exc;
  if (blocks > 0x7fffffffLL) return pair_fail(SPI_ERR_TOO_LARGE, "spi_traj_sensitivity_pairs: grid too large");
  traj_sens_pair_kernel<<<static_cast<unsigned>(blocks), kPairThreads, lds, stream>>>(
      *sim, *tan, params_dev, ld, n_samples, init_dev, actions_dev, horizon, static_cast<int>(bpe), J_dev, y_dev,
      status_dev);
  const hipError_t err = hipGetLastError();
  if (err == hipSuccess) return SPI_OK;
  char buf[256];
  std::snprintf(buf, sizeof(buf), "traj_sens_pair_kernel: %s", hipGetErrorString(err));
  spi_set_error(buf);
  return SPI_ERR_HIP;
}

}  // extern "C"
