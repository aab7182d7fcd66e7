// Reverse-mode gradient of the fused rollout cost on gfx950, four lanes per (candidate,
// segment) rollout (spi_adj_leg.h; one lane per rollout: spi_adjoint_lane.hip): the segment
// record staged in LDS as in the primal kernel, and a caller-provided [slot][19][lane] work
// buffer holding the block checkpoints and the physics-step states of the current block, so
// that every put / get is a coalesced access.
#include <hip/hip_runtime.h>

#include <cstdio>

#include "spi_adj_launch.h"
#include "spi_adj_leg.h"
#include "spi_launch.h"

using namespace spi;
using namespace spi_adj_launch;
using spi_launch::finite_bits;
using spi_launch::lane_params_valid;
using spi_launch::apply_segment_gains;
using spi_launch::xcd_remap;

namespace {

// Checkpoint store of the leg-parallel kernel: this lane's LegState (base + three joints).
constexpr int kLegStateDim = 19;
struct DevLegStore {
  float* base;
  size_t lanes, lane;
  __device__ void put(int slot, const LegState& s) {
    float* d = base + static_cast<size_t>(slot) * kLegStateDim * lanes + lane;
    d[0] = s.p.x; d[lanes] = s.p.y; d[2 * lanes] = s.p.z;
    d[3 * lanes] = s.qx; d[4 * lanes] = s.qy; d[5 * lanes] = s.qz; d[6 * lanes] = s.qw;
    d[7 * lanes] = s.w.x; d[8 * lanes] = s.w.y; d[9 * lanes] = s.w.z;
    d[10 * lanes] = s.v.x; d[11 * lanes] = s.v.y; d[12 * lanes] = s.v.z;
#pragma unroll
    for (int k = 0; k < 3; ++k) { d[(13 + k) * lanes] = s.q[k]; d[(16 + k) * lanes] = s.qd[k]; }
  }
  __device__ void get(int slot, LegState& s) {
    const float* d = base + static_cast<size_t>(slot) * kLegStateDim * lanes + lane;
    s.p = V3{d[0], d[lanes], d[2 * lanes]};
    s.qx = d[3 * lanes]; s.qy = d[4 * lanes]; s.qz = d[5 * lanes]; s.qw = d[6 * lanes];
    s.w = V3{d[7 * lanes], d[8 * lanes], d[9 * lanes]};
    s.v = V3{d[10 * lanes], d[11 * lanes], d[12 * lanes]};
#pragma unroll
    for (int k = 0; k < 3; ++k) { s.q[k] = d[(13 + k) * lanes]; s.qd[k] = d[(16 + k) * lanes]; }
  }
};

template <class T> __device__ __forceinline__ T pick4(int l, const T& x0, const T& x1, const T& x2, const T& x3) {
  return l == 0 ? x0 : (l == 1 ? x1 : (l == 2 ? x2 : x3));
}

// cost_adjoint_kernel with four lanes per rollout (lane % 4 = leg): one workgroup = one
// segment x 64 candidates.  Lane 0 of a rollout writes the cost, the base rows and the
// motor rows (summed over the four lanes); lane l writes its leg's Kp-scale, Kd-scale and
// friction rows.  A rollout's result is dropped (cost +inf, gradient 0) when any of its
// lanes holds a non-finite value.
__global__ __launch_bounds__(kAdjThreads) void cost_adjoint_leg_kernel(
    SpiSimArgs a, SpiCostArgs cost, const float* __restrict__ par, int ld, int n_cand, const float* __restrict__ seg,
    int H, int stride, int n_cblk, int every, float* __restrict__ work, float* __restrict__ out,
    float* __restrict__ grad) {
  extern __shared__ float lds[];
  const int logical = xcd_remap(blockIdx.x, gridDim.x);
  const int s = logical / n_cblk, cb = logical % n_cblk;
  const float* src = seg + static_cast<size_t>(s) * stride;
  for (int i = threadIdx.x; i < stride; i += kAdjThreads) lds[i] = src[i];
  __syncthreads();
  const int c = cb * (kAdjThreads / 4) + (threadIdx.x >> 2), l = threadIdx.x & 3;
  if (c >= n_cand) return;  // the four lanes of a rollout leave together
  apply_segment_gains(a, cost, s);
  const int n_seg = gridDim.x / n_cblk;
  const size_t rollout = static_cast<size_t>(s) * n_cand + c;
  LegParams lp;
  lp.base = make_base_inertia(par[0 * ld + c], par[1 * ld + c], par[2 * ld + c], par[3 * ld + c], par[4 * ld + c],
                              par[5 * ld + c], par[6 * ld + c], par[7 * ld + c], par[8 * ld + c], par[9 * ld + c]);
  lp.kp_s = par[(10 + l) * ld + c];
  lp.kd_s = par[(14 + l) * ld + c];
  lp.fric = par[(18 + l) * ld + c];
#pragma unroll
  for (int g = 0; g < 3; ++g) lp.motor[g] = par[(22 + g) * ld + c];
  const float raw[4] = {par[0 * ld + c], par[1 * ld + c], par[2 * ld + c], par[3 * ld + c]};
  const LegK K = leg_constants(l);
  LegGains G;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    G.kp[k] = pick4(l, a.kp[k], a.kp[3 + k], a.kp[6 + k], a.kp[9 + k]);
    G.kd[k] = pick4(l, a.kd[k], a.kd[3 + k], a.kd[6 + k], a.kd[9 + k]);
    G.q0[k] = pick4(l, a.q0[k], a.q0[3 + k], a.q0[6 + k], a.q0[9 + k]);
    G.effort[k] = pick4(l, a.effort[k], a.effort[3 + k], a.effort[6 + k], a.effort[9 + k]);
  }
  DevLegStore store{work, static_cast<size_t>(n_seg) * n_cand * 4, rollout * 4 + l};
  LegState ss[kMaxSubsteps];
  SubAcc sa[kMaxSubsteps];
  LegParBar pb;
  zero_par(pb);
  float cst = 0.0f;
  // the same verdict in the four lanes: an invalid candidate's quad skips the rollout (and its sums) together
  const bool valid = lane_params_valid(par, ld, c, a.motor_model);
  bool ok = valid && rollout_grad_leg(a, cost, lp, K, G, raw, lds, H, every, l, store, ss, sa, cst, pb);
  if (valid) {
#pragma unroll
    for (int g = 0; g < 3; ++g) pb.motor[g] = sum4(pb.motor[g]);
  }
  if (ok) {
    bool fin = finite_bits(cst) && finite_bits(pb.kp) && finite_bits(pb.kd) && finite_bits(pb.fric);
#pragma unroll
    for (int k = 0; k < 10; ++k) fin = fin && finite_bits(pb.base[k]);
#pragma unroll
    for (int g = 0; g < 3; ++g) fin = fin && finite_bits(pb.motor[g]);
    ok = sum4(fin ? 0.0f : 1.0f) == 0.0f;   // the same verdict in the four lanes (ok implies valid)
  }
  float* gr = grad + rollout * SPI_P_COUNT;
  gr[10 + l] = ok ? pb.kp : 0.0f;
  gr[14 + l] = ok ? pb.kd : 0.0f;
  gr[18 + l] = ok ? pb.fric : 0.0f;
  if (l == 0) {
#pragma unroll
    for (int k = 0; k < 10; ++k) gr[k] = ok ? pb.base[k] : 0.0f;
#pragma unroll
    for (int g = 0; g < 3; ++g) gr[22 + g] = ok ? pb.motor[g] : 0.0f;
    reinterpret_cast<unsigned int*>(out)[rollout] = ok ? __float_as_uint(cst) : 0x7f800000u;
  }
}

}  // namespace

extern "C" {

long long spi_adjoint_legs_work_floats(int n_cand, int n_seg, int horizon, int checkpoint_every, int decimation) {
  if (n_cand < 0 || n_seg < 0 || horizon < 1 || checkpoint_every < 1 || decimation < 1 || decimation > kMaxDecimation)
    return -1;
  const long long nb = (horizon + checkpoint_every - 1) / checkpoint_every;
  const long long slots = nb + static_cast<long long>(checkpoint_every) * decimation;
  return static_cast<long long>(n_cand) * n_seg * 4 * slots * kLegStateDim;
}

int spi_rollout_cost_adjoint_legs(const SpiSimArgs* sim, const SpiCostArgs* cost, const float* params_dev, int ld,
                                  int n_cand, const float* segments_dev, int n_seg, int horizon, int checkpoint_every,
                                  float* work_dev, long long work_floats, float* cost_dev, float* grad_dev,
                                  hipStream_t stream) {
  auto fail = [](int code, const char* msg) { spi_set_error(msg); return code; };
  const int rc = adjoint_args_check("spi_rollout_cost_adjoint_legs", sim, cost, params_dev, ld, n_cand, segments_dev,
                                    n_seg, horizon, checkpoint_every, work_dev, cost_dev, grad_dev);
  if (rc != SPI_OK) return rc;
  if (n_cand == 0 || n_seg == 0) return SPI_OK;
  if (work_floats < spi_adjoint_legs_work_floats(n_cand, n_seg, horizon, checkpoint_every, sim->decimation))
    return fail(SPI_ERR_INVALID_ARG,
                "spi_rollout_cost_adjoint_legs: work buffer smaller than spi_adjoint_legs_work_floats()");
  const int stride = SPI_STATE_DIM + horizon * (12 + SPI_STATE_DIM);
  const size_t lds = static_cast<size_t>(stride) * sizeof(float);
  if (lds > 64 * 1024)
    return fail(SPI_ERR_TOO_LARGE, "spi_rollout_cost_adjoint_legs: horizon too long for LDS staging");
  const int per_block = kAdjThreads / 4;
  const int n_cblk = (n_cand + per_block - 1) / per_block;
  const long long blocks = static_cast<long long>(n_cblk) * n_seg;
  if (blocks > 0x7fffffffLL || static_cast<long long>(n_seg) * n_cand * 4 > (1LL << 31))
    return fail(SPI_ERR_TOO_LARGE, "spi_rollout_cost_adjoint_legs: grid too large");
  cost_adjoint_leg_kernel<<<static_cast<unsigned>(blocks), kAdjThreads, lds, stream>>>(
      *sim, *cost, params_dev, ld, n_cand, segments_dev, horizon, stride, n_cblk, checkpoint_every, work_dev,
      cost_dev, grad_dev);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(SPI_ERR_HIP, hipGetErrorString(e));
  return SPI_OK;
}

}  // extern "C"
