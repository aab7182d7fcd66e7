// Reverse-mode gradient of the fused rollout cost (spi_adj.h) on gfx950, one lane per
// (candidate, segment) rollout: the segment record staged in LDS as in the primal kernel;
// checkpoints and the recomputed states of the current block live in a caller-provided work
// buffer laid out [slot][37][lane] so that every put / get is a coalesced access.  Apart from
// the four-lanes-per-rollout form (spi_adjoint.hip) because it is built without reassociation:
// it already spills, and reassociation keeps more values live (461 -> 540 ms per launch).
#include <hip/hip_runtime.h>

#include <cstdio>

#include "spi_adj.h"
#include "spi_adj_launch.h"
#include "spi_launch.h"

using namespace spi;
using namespace spi_adj_launch;
using spi_launch::finite_bits;
using spi_launch::lane_params_valid;
using spi_launch::apply_segment_gains;
using spi_launch::xcd_remap;

namespace {

struct DevStore {
  float* base;
  size_t lanes, lane;
  __device__ void put(int slot, const State& s) {
    float* d = base + static_cast<size_t>(slot) * SPI_STATE_DIM * lanes + lane;
    d[0] = s.p.x; d[lanes] = s.p.y; d[2 * lanes] = s.p.z;
    d[3 * lanes] = s.qx; d[4 * lanes] = s.qy; d[5 * lanes] = s.qz; d[6 * lanes] = s.qw;
    d[7 * lanes] = s.w.x; d[8 * lanes] = s.w.y; d[9 * lanes] = s.w.z;
    d[10 * lanes] = s.v.x; d[11 * lanes] = s.v.y; d[12 * lanes] = s.v.z;
#pragma unroll
    for (int j = 0; j < 12; ++j) { d[(13 + j) * lanes] = s.q[j]; d[(25 + j) * lanes] = s.qd[j]; }
  }
  __device__ void get(int slot, State& s) {
    const float* d = base + static_cast<size_t>(slot) * SPI_STATE_DIM * lanes + lane;
    s.p = V3{d[0], d[lanes], d[2 * lanes]};
    s.qx = d[3 * lanes]; s.qy = d[4 * lanes]; s.qz = d[5 * lanes]; s.qw = d[6 * lanes];
    s.w = V3{d[7 * lanes], d[8 * lanes], d[9 * lanes]};
    s.v = V3{d[10 * lanes], d[11 * lanes], d[12 * lanes]};
#pragma unroll
    for (int j = 0; j < 12; ++j) { s.q[j] = d[(13 + j) * lanes]; s.qd[j] = d[(25 + j) * lanes]; }
  }
};

__global__ __launch_bounds__(kAdjThreads) void cost_adjoint_kernel(
    SpiSimArgs a, SpiCostArgs cost, const float* __restrict__ par, int ld, int n_cand, const float* __restrict__ seg,
    int H, int stride, int n_cblk, int every, float* __restrict__ work, float* __restrict__ out,
    float* __restrict__ grad) {
  extern __shared__ float lds[];
  const int logical = xcd_remap(blockIdx.x, gridDim.x);
  const int s = logical / n_cblk, cb = logical % n_cblk;
  const float* src = seg + static_cast<size_t>(s) * stride;
  for (int i = threadIdx.x; i < stride; i += kAdjThreads) lds[i] = src[i];
  __syncthreads();
  const int c = cb * kAdjThreads + threadIdx.x;
  if (c >= n_cand) return;
  apply_segment_gains(a, cost, s);
  const int n_seg = gridDim.x / n_cblk;
  const size_t lane = static_cast<size_t>(s) * n_cand + c;
  LaneParams lp;
  lp.base = make_base_inertia(par[0 * ld + c], par[1 * ld + c], par[2 * ld + c], par[3 * ld + c], par[4 * ld + c],
                              par[5 * ld + c], par[6 * ld + c], par[7 * ld + c], par[8 * ld + c], par[9 * ld + c]);
#pragma unroll
  for (int l = 0; l < 4; ++l) {
    lp.kp_s[l] = par[(10 + l) * ld + c];
    lp.kd_s[l] = par[(14 + l) * ld + c];
    lp.fric[l] = par[(18 + l) * ld + c];
  }
#pragma unroll
  for (int g = 0; g < 3; ++g) lp.motor[g] = par[(22 + g) * ld + c];
  const float raw[4] = {par[0 * ld + c], par[1 * ld + c], par[2 * ld + c], par[3 * ld + c]};
  DevStore store{work, static_cast<size_t>(n_seg) * n_cand, lane};
  State ps[kMaxDecimation], ss[kMaxSubsteps];
  ParBar pb;
  float cst = 0.0f;
  bool ok = lane_params_valid(par, ld, c, a.motor_model) &&
            rollout_grad(a, cost, lp, raw, lds, H, every, store, ps, ss, cst, pb);
  if (ok) {
    ok = finite_bits(cst);
#pragma unroll
    for (int k = 0; k < SPI_P_COUNT; ++k) ok = ok && finite_bits(pb.g[k]);
  }
  float* g = grad + lane * SPI_P_COUNT;
#pragma unroll
  for (int k = 0; k < SPI_P_COUNT; ++k) g[k] = ok ? pb.g[k] : 0.0f;
  reinterpret_cast<unsigned int*>(out)[lane] = ok ? __float_as_uint(cst) : 0x7f800000u;
}

}  // namespace

extern "C" {

long long spi_adjoint_work_floats(int n_cand, int n_seg, int horizon, int checkpoint_every) {
  if (n_cand < 0 || n_seg < 0 || horizon < 1 || checkpoint_every < 1) return -1;
  const long long nb = (horizon + checkpoint_every - 1) / checkpoint_every;
  return static_cast<long long>(n_cand) * n_seg * (nb + checkpoint_every) * SPI_STATE_DIM;
}

int spi_rollout_cost_adjoint(const SpiSimArgs* sim, const SpiCostArgs* cost, const float* params_dev, int ld,
                             int n_cand, const float* segments_dev, int n_seg, int horizon, int checkpoint_every,
                             float* work_dev, long long work_floats, float* cost_dev, float* grad_dev,
                             hipStream_t stream) {
  auto fail = [](int code, const char* msg) { spi_set_error(msg); return code; };
  const int rc = adjoint_args_check("spi_rollout_cost_adjoint", sim, cost, params_dev, ld, n_cand, segments_dev, n_seg,
                                    horizon, checkpoint_every, work_dev, cost_dev, grad_dev);
  if (rc != SPI_OK) return rc;
  if (n_cand == 0 || n_seg == 0) return SPI_OK;
  if (work_floats < spi_adjoint_work_floats(n_cand, n_seg, horizon, checkpoint_every))
    return fail(SPI_ERR_INVALID_ARG, "spi_rollout_cost_adjoint: work buffer smaller than spi_adjoint_work_floats()");
  const int stride = SPI_STATE_DIM + horizon * (12 + SPI_STATE_DIM);
  const size_t lds = static_cast<size_t>(stride) * sizeof(float);
  if (lds > 64 * 1024) return fail(SPI_ERR_TOO_LARGE, "spi_rollout_cost_adjoint: horizon too long for LDS staging");
  const int n_cblk = (n_cand + kAdjThreads - 1) / kAdjThreads;
  const long long blocks = static_cast<long long>(n_cblk) * n_seg;
  if (blocks > 0x7fffffffLL) return fail(SPI_ERR_TOO_LARGE, "spi_rollout_cost_adjoint: grid too large");
  cost_adjoint_kernel<<<static_cast<unsigned>(blocks), kAdjThreads, lds, stream>>>(
      *sim, *cost, params_dev, ld, n_cand, segments_dev, horizon, stride, n_cblk, checkpoint_every, work_dev,
      cost_dev, grad_dev);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(SPI_ERR_HIP, hipGetErrorString(e));
  return SPI_OK;
}

}  // extern "C"
