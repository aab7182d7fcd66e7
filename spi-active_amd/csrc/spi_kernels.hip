// HIP kernels and the C ABI (include/spi_hip.h) of the gfx950 sysID engine.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>

#include "spi_dyn.h"
#include "spi_dyn_leg.h"
#include "spi_launch.h"

using namespace spi;

namespace {
thread_local char g_err[512] = "";
}  // namespace

void spi_set_error(const char* msg) { std::snprintf(g_err, sizeof(g_err), "%s", msg); }

namespace {

int fail(int code, const char* msg) {
  spi_set_error(msg);
  return code;
}
int hip_check(hipError_t e, const char* what) {
  if (e == hipSuccess) return SPI_OK;
  std::snprintf(g_err, sizeof(g_err), "%s: %s", what, hipGetErrorString(e));
  return SPI_ERR_HIP;
}
int launch_check(const char* what) { return hip_check(hipGetLastError(), what); }

using spi_launch::apply_segment_gains;
using spi_launch::xcd_remap;

__device__ __forceinline__ LaneParams load_params(const float* __restrict__ par, int ld, int i) {
  LaneParams lp;
  lp.base = make_base_inertia(par[0 * ld + i], par[1 * ld + i], par[2 * ld + i], par[3 * ld + i], par[4 * ld + i],
                              par[5 * ld + i], par[6 * ld + i], par[7 * ld + i], par[8 * ld + i], par[9 * ld + i]);
#pragma unroll
  for (int l = 0; l < 4; ++l) {
    lp.kp_s[l] = par[(10 + l) * ld + i];
    lp.kd_s[l] = par[(14 + l) * ld + i];
    lp.fric[l] = par[(18 + l) * ld + i];
  }
#pragma unroll
  for (int g = 0; g < 3; ++g) lp.motor[g] = par[(22 + g) * ld + i];
  return lp;
}

using spi_launch::finite_bits;
using spi_launch::lane_params_valid;

constexpr int kRolloutThreads = 256;

// One (candidate, segment) rollout of H control steps from the segment record `rec` (init
// state, then per step: action, target), returning the cost as float bits (+inf bits when
// the lane diverged).  `rec` is in LDS and read at uniform addresses.
template <int UNROLL>
__device__ __forceinline__ unsigned int rollout_lane(const SpiSimArgs& a, const SpiCostArgs& cost,
                                                     const LaneParams& lp, const float* rec0, int H) {
  State st;
  state_from_root(st, rec0, rec0 + 13, rec0 + 25, 1);
  float acc = 0.0f;
  bool ok = true;
  for (int t = 0; t < H; ++t) {
    const float* rec = rec0 + SPI_STATE_DIM + t * (12 + SPI_STATE_DIM);
    float act[12];
#pragma unroll
    for (int j = 0; j < 12; ++j) act[j] = rec[j];
#pragma unroll UNROLL
    for (int d = 0; d < a.decimation; ++d) {
      float tau[12];
      control_torques(a, lp, act, st, tau);
      physics_step(a, lp, st, tau, nullptr);
    }
    if (cost.mode == SPI_COST_TRACKING || t == H - 1) acc += step_cost(cost, st, rec + 12);
    const float nrm = state_norm(st);
    if (!finite_bits(nrm) || nrm >= a.state_limit) { ok = false; break; }
  }
  // diverged lanes get +inf, as a bit pattern (this file is built with finite-math)
  return (ok && finite_bits(acc)) ? __float_as_uint(acc) : 0x7f800000u;
}

// One workgroup = one segment x 256 candidates.  The segment record is staged once in LDS
// and read by every lane (uniform address: LDS broadcast); parameters are read coalesced
// from the SoA array.
__global__ __launch_bounds__(kRolloutThreads) void rollout_cost_kernel(
    SpiSimArgs a, SpiCostArgs cost, const float* __restrict__ par, int ld, int n_cand,
    const float* __restrict__ seg, int H, int stride, int n_cblk, float* __restrict__ out) {
  extern __shared__ float lds[];
  const int logical = xcd_remap(blockIdx.x, gridDim.x);
  const int s = logical / n_cblk, cb = logical % n_cblk;
  const float* src = seg + static_cast<size_t>(s) * stride;
  for (int i = threadIdx.x; i < stride; i += blockDim.x) lds[i] = src[i];
  __syncthreads();
  const int c = cb * kRolloutThreads + threadIdx.x;
  if (c >= n_cand) return;
  apply_segment_gains(a, cost, s);
  unsigned int* o = reinterpret_cast<unsigned int*>(out) + static_cast<size_t>(s) * n_cand + c;
  if (!lane_params_valid(par, ld, c, a.motor_model)) { *o = 0x7f800000u; return; }
  const LaneParams lp = load_params(par, ld, c);
  *o = rollout_lane<1>(a, cost, lp, lds, H);
}

// rollout_cost_kernel with four lanes per candidate, one per leg (spi_dyn_leg.h): one
// workgroup = one segment x 64 candidates.  For batches that leave most SIMDs idle (a mass
// landscape, a small job), the rollout's latency drops at the price of more instructions in
// total; same costs up to the order of the legs' sums.
template <class T> __device__ __forceinline__ T pick4(int l, const T& x0, const T& x1, const T& x2, const T& x3) {
  return l == 0 ? x0 : (l == 1 ? x1 : (l == 2 ? x2 : x3));
}
__global__ __launch_bounds__(kRolloutThreads) void rollout_cost_leg_kernel(
    SpiSimArgs a, SpiCostArgs cost, const float* __restrict__ par, int ld, int n_cand,
    const float* __restrict__ seg, int H, int stride, int n_cblk, float* __restrict__ out) {
  extern __shared__ float lds[];
  const int logical = xcd_remap(blockIdx.x, gridDim.x);
  const int s = logical / n_cblk, cb = logical % n_cblk;
  const float* src = seg + static_cast<size_t>(s) * stride;
  for (int i = threadIdx.x; i < stride; i += blockDim.x) lds[i] = src[i];
  __syncthreads();
  const int c = cb * (kRolloutThreads / 4) + (threadIdx.x >> 2), l = threadIdx.x & 3;
  if (c >= n_cand) return;  // the four lanes of a candidate leave together
  apply_segment_gains(a, cost, s);
  if (!lane_params_valid(par, ld, c, a.motor_model)) {   // the same verdict in the four lanes
    if (l == 0) reinterpret_cast<unsigned int*>(out)[static_cast<size_t>(s) * n_cand + c] = 0x7f800000u;
    return;
  }
  const LaneParams full = load_params(par, ld, c);
  LegParams lp;
  lp.base = full.base;
  lp.kp_s = pick4(l, full.kp_s[0], full.kp_s[1], full.kp_s[2], full.kp_s[3]);
  lp.kd_s = pick4(l, full.kd_s[0], full.kd_s[1], full.kd_s[2], full.kd_s[3]);
  lp.fric = pick4(l, full.fric[0], full.fric[1], full.fric[2], full.fric[3]);
#pragma unroll
  for (int g = 0; g < 3; ++g) lp.motor[g] = full.motor[g];
  const LegK K = leg_constants(l);
  float kp[3], kd[3], q0[3], eff[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    kp[k] = pick4(l, a.kp[k], a.kp[3 + k], a.kp[6 + k], a.kp[9 + k]);
    kd[k] = pick4(l, a.kd[k], a.kd[3 + k], a.kd[6 + k], a.kd[9 + k]);
    q0[k] = pick4(l, a.q0[k], a.q0[3 + k], a.q0[6 + k], a.q0[9 + k]);
    eff[k] = pick4(l, a.effort[k], a.effort[3 + k], a.effort[6 + k], a.effort[9 + k]);
  }
  LegState st;
  leg_state_from_root(st, lds, l);
  float acc = 0.0f;
  bool ok = true;
  for (int t = 0; t < H; ++t) {
    const float* rec = lds + SPI_STATE_DIM + t * (12 + SPI_STATE_DIM);
    float act[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) act[k] = rec[3 * l + k];
    for (int d = 0; d < a.decimation; ++d) {
      float tau[3];
      control_torques_leg(a, lp, kp, kd, q0, eff, act, st.q, st.qd, tau);
      physics_step_leg(a, lp, K, st, tau);
    }
    if (cost.mode == SPI_COST_TRACKING || t == H - 1) acc += step_cost_leg(cost, st, rec + 12, l);
    const float nrm = state_norm_leg(st);   // identical in the four lanes
    if (!finite_bits(nrm) || nrm >= a.state_limit) { ok = false; break; }
  }
  if (l == 0)
    reinterpret_cast<unsigned int*>(out)[static_cast<size_t>(s) * n_cand + c] =
        (ok && finite_bits(acc)) ? __float_as_uint(acc) : 0x7f800000u;
}

// Batched problems: a persistent grid pulls tasks (problem, segment, candidate block) from a
// device queue.  Tickets are handed out in table order (the host orders problems longest
// horizon first), every block leaves its loop once the tickets run out, and the last block
// to leave resets the queue so the same launch can be replayed (HIP graphs).
template <int BLOCK, int UNROLL>
__global__ __launch_bounds__(BLOCK) void rollout_cost_batch_kernel(SpiSimArgs a, const SpiBatchProblem* __restrict__ probs,
                                                                   const long long* __restrict__ prefix, int n_prob,
                                                                   long long n_tasks,
                                                                   unsigned long long* __restrict__ queue) {
  extern __shared__ float lds[];
  __shared__ long long ticket;
  for (;;) {
    if (threadIdx.x == 0) ticket = static_cast<long long>(atomicAdd(&queue[0], 1ULL));
    __syncthreads();
    const long long t = ticket;
    if (t >= n_tasks) break;
    int lo = 0, hi = n_prob - 1;             // last problem with prefix[p] <= t
    while (lo < hi) {
      int mid = (lo + hi + 1) >> 1;
      if (prefix[mid] <= t) lo = mid; else hi = mid - 1;
    }
    const SpiBatchProblem& P = probs[lo];
    const long long local = t - prefix[lo];
    const int n_cblk = (P.n_samples + BLOCK - 1) / BLOCK;
    const int s = static_cast<int>(local / n_cblk), cb = static_cast<int>(local % n_cblk);
    const int stride = SPI_STATE_DIM + P.horizon * (12 + SPI_STATE_DIM);
    const float* src = P.segments + static_cast<size_t>(s) * stride;
    for (int i = threadIdx.x; i < stride; i += BLOCK) lds[i] = src[i];
    __syncthreads();
    const int c = cb * BLOCK + threadIdx.x;
    if (c < P.n_samples) {
      SpiSimArgs at = a;
      apply_segment_gains(at, P.cost, s);
      unsigned int* o = reinterpret_cast<unsigned int*>(P.cost_seg) + static_cast<size_t>(s) * P.n_samples + c;
      if (lane_params_valid(P.theta, P.n_samples, c, at.motor_model)) {
        const LaneParams lp = load_params(P.theta, P.n_samples, c);
        *o = rollout_lane<UNROLL>(at, P.cost, lp, lds, P.horizon);
      } else {
        *o = 0x7f800000u;
      }
    }
    __syncthreads();                          // LDS and `ticket` are reused by the next task
  }
  if (threadIdx.x == 0) {
    __threadfence();                          // this block's last ticket precedes its exit count
    unsigned long long done = atomicAdd(&queue[1], 1ULL);
    if (done == gridDim.x - 1) {              // every other block has left its loop
      atomicExch(&queue[0], 0ULL);
      atomicExch(&queue[1], 0ULL);
    }
  }
}

__global__ __launch_bounds__(256) void physics_step_kernel(SpiSimArgs a, const float* __restrict__ par, int ld,
                                                           int n, float* __restrict__ root, float* __restrict__ dof,
                                                           const float* __restrict__ torques,
                                                           float* __restrict__ foot_force) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const LaneParams lp = load_params(par, ld, i);
  State st;
  float* r = root + static_cast<size_t>(i) * 13;
  float* d = dof + static_cast<size_t>(i) * 24;
  state_from_root(st, r, d, d + 1, 2);
  float tau[12];
#pragma unroll
  for (int j = 0; j < 12; ++j) tau[j] = torques[static_cast<size_t>(i) * 12 + j];
  float ff[12];
  physics_step(a, lp, st, tau, foot_force ? ff : nullptr);
  state_to_root(st, r, d, d + 1, 2);
  if (foot_force) {
#pragma unroll
    for (int k = 0; k < 12; ++k) foot_force[static_cast<size_t>(i) * 12 + k] = ff[k];
  }
}

__global__ __launch_bounds__(256) void forward_dynamics_kernel(SpiSimArgs a, const float* __restrict__ par, int ld,
                                                               int n, const float* __restrict__ root,
                                                               const float* __restrict__ dof,
                                                               const float* __restrict__ torques,
                                                               float* __restrict__ acc, float* __restrict__ foot_force) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const LaneParams lp = load_params(par, ld, i);
  State st;
  state_from_root(st, root + static_cast<size_t>(i) * 13, dof + static_cast<size_t>(i) * 24,
                  dof + static_cast<size_t>(i) * 24 + 1, 2);
  float tau[12], qdd[12], ff[12];
#pragma unroll
  for (int j = 0; j < 12; ++j) tau[j] = torques[static_cast<size_t>(i) * 12 + j];
  M6 a0;
  forward_dynamics(a, lp, st, tau, a0, qdd, foot_force ? ff : nullptr);
  float* o = acc + static_cast<size_t>(i) * 18;
  o[0] = a0.w.x; o[1] = a0.w.y; o[2] = a0.w.z; o[3] = a0.v.x; o[4] = a0.v.y; o[5] = a0.v.z;
#pragma unroll
  for (int j = 0; j < 12; ++j) o[6 + j] = qdd[j];
  if (foot_force) {
#pragma unroll
    for (int k = 0; k < 12; ++k) foot_force[static_cast<size_t>(i) * 12 + k] = ff[k];
  }
}

__global__ void compute_torques_kernel(SpiSimArgs a, const float* __restrict__ par, int ld, int n,
                                       const float* __restrict__ actions, const float* __restrict__ dof,
                                       float* __restrict__ torques) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const LaneParams lp = load_params(par, ld, i);
  State st;
#pragma unroll
  for (int j = 0; j < 12; ++j) {
    st.q[j] = dof[static_cast<size_t>(i) * 24 + 2 * j];
    st.qd[j] = dof[static_cast<size_t>(i) * 24 + 2 * j + 1];
  }
  float act[12], tau[12];
#pragma unroll
  for (int j = 0; j < 12; ++j) act[j] = actions[static_cast<size_t>(i) * 12 + j];
  control_torques(a, lp, act, st, tau);
#pragma unroll
  for (int j = 0; j < 12; ++j) torques[static_cast<size_t>(i) * 12 + j] = tau[j];
}

__global__ void set_rows_kernel(float* __restrict__ dst, const float* __restrict__ src, const int* __restrict__ idx,
                                int n_idx, int n_rows, int row_len) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n_idx * row_len) return;
  int k = t / row_len, e = t % row_len;
  int r = idx[k];
  if (r < 0 || r >= n_rows) return;
  dst[static_cast<size_t>(r) * row_len + e] = src[static_cast<size_t>(r) * row_len + e];
}

// Forward kinematics of the 19 IsaacGym bodies (rigid-body state tensor rows).
__global__ void rigid_body_kernel(const float* __restrict__ root, const float* __restrict__ dof, int n,
                                  float* __restrict__ out) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  State st;
  state_from_root(st, root + static_cast<size_t>(i) * 13, dof + static_cast<size_t>(i) * 24,
                  dof + static_cast<size_t>(i) * 24 + 1, 2);
  G3 R0 = quat_to_rot(st.qx, st.qy, st.qz, st.qw);
  // body poses (world) and spatial velocities (body coords) of the 13 folded bodies
  G3 Rb[13];
  V3 pb[13];
  M6 vb[13];
  Rb[0] = R0;
  pb[0] = st.p;
  vb[0] = M6{st.w, st.v};
#pragma unroll
  for (int b = 1; b < 13; ++b) {
    const int j = b - 1, parent = (j % 3 == 0) ? 0 : b - 1, K = SPI_JOINT_AXIS[j];
    CV3 r{float(SPI_JOINT_ORIGIN[3 * j]), float(SPI_JOINT_ORIGIN[3 * j + 1]), float(SPI_JOINT_ORIGIN[3 * j + 2])};
    float s, c;
    spi_sincos(st.q[j], &s, &c);
    M6 vp = vb[parent];
    M6 v;
    if (K == 0) v = xm<0>(vp, r, c, s);
    else if (K == 1) v = xm<1>(vp, r, c, s);
    else v = xm<2>(vp, r, c, s);
    if (K == 0) v.w.x += st.qd[j]; else if (K == 1) v.w.y += st.qd[j]; else v.w.z += st.qd[j];
    vb[b] = v;
    const G3& Rp = Rb[parent];
    pb[b] = pb[parent] + mul(Rp, cvec(r));
    // R_b = R_p * R_K(q): rotate the rows of R_p by R_K^T
    V3 x0, x1, x2;
    if (K == 0) { x0 = rotT<0>(V3{Rp.m00, Rp.m01, Rp.m02}, c, s); x1 = rotT<0>(V3{Rp.m10, Rp.m11, Rp.m12}, c, s); x2 = rotT<0>(V3{Rp.m20, Rp.m21, Rp.m22}, c, s); }
    else if (K == 1) { x0 = rotT<1>(V3{Rp.m00, Rp.m01, Rp.m02}, c, s); x1 = rotT<1>(V3{Rp.m10, Rp.m11, Rp.m12}, c, s); x2 = rotT<1>(V3{Rp.m20, Rp.m21, Rp.m22}, c, s); }
    else { x0 = rotT<2>(V3{Rp.m00, Rp.m01, Rp.m02}, c, s); x1 = rotT<2>(V3{Rp.m10, Rp.m11, Rp.m12}, c, s); x2 = rotT<2>(V3{Rp.m20, Rp.m21, Rp.m22}, c, s); }
    Rb[b] = G3{x0.x, x0.y, x0.z, x1.x, x1.y, x1.z, x2.x, x2.y, x2.z};
  }
  float* o = out + static_cast<size_t>(i) * SPI_NISAAC * 13;
#pragma unroll
  for (int k = 0; k < SPI_NISAAC; ++k) {
    const int b = SPI_ISAAC_BODY[k];
    V3 off{float(SPI_ISAAC_OFFSET[3 * k]), float(SPI_ISAAC_OFFSET[3 * k + 1]), float(SPI_ISAAC_OFFSET[3 * k + 2])};
    const G3& R = Rb[b];
    V3 p = pb[b] + mul(R, off);
    V3 vl = vb[b].v + cross(vb[b].w, off);
    V3 vw = mul(R, vl), ww = mul(R, vb[b].w);
    // rotation matrix -> quaternion (xyzw), Shepperd's method
    float tr = R.m00 + R.m11 + R.m22, qx, qy, qz, qw;
    if (tr > 0.f) {
      float S = sqrtf(tr + 1.f) * 2.f;
      qw = 0.25f * S; qx = (R.m21 - R.m12) / S; qy = (R.m02 - R.m20) / S; qz = (R.m10 - R.m01) / S;
    } else if (R.m00 > R.m11 && R.m00 > R.m22) {
      float S = sqrtf(1.f + R.m00 - R.m11 - R.m22) * 2.f;
      qw = (R.m21 - R.m12) / S; qx = 0.25f * S; qy = (R.m01 + R.m10) / S; qz = (R.m02 + R.m20) / S;
    } else if (R.m11 > R.m22) {
      float S = sqrtf(1.f + R.m11 - R.m00 - R.m22) * 2.f;
      qw = (R.m02 - R.m20) / S; qx = (R.m01 + R.m10) / S; qy = 0.25f * S; qz = (R.m12 + R.m21) / S;
    } else {
      float S = sqrtf(1.f + R.m22 - R.m00 - R.m11) * 2.f;
      qw = (R.m10 - R.m01) / S; qx = (R.m02 + R.m20) / S; qy = (R.m12 + R.m21) / S; qz = 0.25f * S;
    }
    if (qw < 0.f) { qx = -qx; qy = -qy; qz = -qz; qw = -qw; }
    float vals[13] = {p.x, p.y, p.z, qx, qy, qz, qw, vw.x, vw.y, vw.z, ww.x, ww.y, ww.z};
#pragma unroll
    for (int e = 0; e < 13; ++e) o[k * 13 + e] = vals[e];
  }
}

// Finite-difference FIM trace, active_sysid_openloop.py:402-426: per group of (1 main + P aux)
// envs, X = [(root_main - root_aux)/delta with origins removed from positions,
// (q_main - q_aux)/delta], value = trace(X^T X) = ||X||_F^2, repeated for the group.
__global__ void fd_fim_kernel(const float* __restrict__ root, const float* __restrict__ dof,
                              const float* __restrict__ origins, int n_main, int P, float delta,
                              float* __restrict__ out) {
  int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= n_main) return;
  const int g = P + 1, im = m * g;
  float acc = 0.0f;
  for (int k = 1; k <= P; ++k) {
    const int ia = im + k;
    for (int e = 0; e < 13; ++e) {
      float xm = root[static_cast<size_t>(im) * 13 + e], xa = root[static_cast<size_t>(ia) * 13 + e];
      if (e < 3) { xm -= origins[static_cast<size_t>(im) * 3 + e]; xa -= origins[static_cast<size_t>(ia) * 3 + e]; }
      float d = (xm - xa) / delta;
      acc += d * d;
    }
    for (int j = 0; j < 12; ++j) {
      float d = (dof[static_cast<size_t>(im) * 24 + 2 * j] - dof[static_cast<size_t>(ia) * 24 + 2 * j]) / delta;
      acc += d * d;
    }
  }
  for (int k = 0; k < g; ++k) out[im + k] = acc;
}

// Record mode: per-lane params, initial state and action sequence; state after every control step.
__global__ __launch_bounds__(256) void simulate_kernel(SpiSimArgs a, const float* __restrict__ par, int ld, int n,
                                                       const float* __restrict__ init,
                                                       const float* __restrict__ actions, int T,
                                                       float* __restrict__ states) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const LaneParams lp = load_params(par, ld, i);
  State st;
  const float* s0 = init + static_cast<size_t>(i) * SPI_STATE_DIM;
  state_from_root(st, s0, s0 + 13, s0 + 25, 1);
  for (int t = 0; t < T; ++t) {
    float act[12];
#pragma unroll
    for (int j = 0; j < 12; ++j) act[j] = actions[(static_cast<size_t>(i) * T + t) * 12 + j];
    for (int d = 0; d < a.decimation; ++d) {
      float tau[12];
      control_torques(a, lp, act, st, tau);
      physics_step(a, lp, st, tau, nullptr);
    }
    float* o = states + (static_cast<size_t>(i) * T + t) * SPI_STATE_DIM;
    state_to_root(st, o, o + 13, o + 25, 1);
  }
}

// Active-sysID FD-FIM rollout.  Group of G = n_param + 1 envs on G2 = pow2 >= G consecutive
// lanes of one wavefront (padding lanes shadow the main env and contribute nothing); states
// of the main env are read by the aux lanes with cross-lane shuffles.
__global__ __launch_bounds__(256) void fim_rollout_kernel(SpiSimArgs a, const float* __restrict__ par, int ld,
                                                          int n_group, int P, int G2, const float* __restrict__ init,
                                                          const float* __restrict__ actions, int T, float delta,
                                                          int ksync, int ksync_mode, float* __restrict__ fim) {
  const int gid = blockIdx.x * blockDim.x + threadIdx.x;
  const int group = gid / G2, member = gid % G2;
  const bool active = group < n_group;
  const int g = active ? group : n_group - 1;
  const int m = member <= P ? member : 0;              // padding lanes shadow the main env
  const int env = g * (P + 1) + m;
  const int lane = threadIdx.x & 63;
  const int main_lane = lane - member;                 // lane of this group's main env
  const LaneParams lp = load_params(par, ld, env);
  State st;
  const float* s0 = init + static_cast<size_t>(env) * SPI_STATE_DIM;
  state_from_root(st, s0, s0 + 13, s0 + 25, 1);
  const float inv_d = 1.0f / delta;
  float acc = 0.0f;
  bool use_main_dof = false;
  float mq[12], mqd[12];
  for (int t = 0; t < T; ++t) {
    float act[12];
#pragma unroll
    for (int j = 0; j < 12; ++j) act[j] = actions[(static_cast<size_t>(g) * T + t) * 12 + j];
    for (int d = 0; d < a.decimation; ++d) {
      float tau[12];
      if (use_main_dof && d == 0) {
        // reference ksync: the synced joint state only reaches the first torque computation
        State tmp = st;
#pragma unroll
        for (int j = 0; j < 12; ++j) { tmp.q[j] = mq[j]; tmp.qd[j] = mqd[j]; }
        control_torques(a, lp, act, tmp, tau);
      } else {
        control_torques(a, lp, act, st, tau);
      }
      physics_step(a, lp, st, tau, nullptr);
    }
    use_main_dof = false;
    // FD sensitivity against the main env: root in world frame (pos, quat, linvel, angvel) + q
    G3 R = quat_to_rot(st.qx, st.qy, st.qz, st.qw);
    V3 vw = mul(R, st.v), ww = mul(R, st.w);
    float x[25] = {st.p.x, st.p.y, st.p.z, st.qx, st.qy, st.qz, st.qw, vw.x, vw.y, vw.z, ww.x, ww.y, ww.z};
#pragma unroll
    for (int j = 0; j < 12; ++j) x[13 + j] = st.q[j];
    float e = 0.0f;
#pragma unroll
    for (int i = 0; i < 25; ++i) {
      float xm = __shfl(x[i], main_lane, 64);
      float dd = (xm - x[i]) * inv_d;
      e += dd * dd;
    }
    if (member == 0 || member > P) e = 0.0f;
    for (int off = 1; off < G2; off <<= 1) e += __shfl_xor(e, off, 64);
    acc += e;
    // the reference's step_idx of this command step: reset_all's zero-action step is step 1
    const int step_idx = t + 2;
    if (ksync > 0 && (ksync == 1 || step_idx % ksync == 1)) {
      if (ksync_mode == 1) {
        // full state copy from the main env
        st.p.x = __shfl(st.p.x, main_lane, 64); st.p.y = __shfl(st.p.y, main_lane, 64);
        st.p.z = __shfl(st.p.z, main_lane, 64);
        st.qx = __shfl(st.qx, main_lane, 64); st.qy = __shfl(st.qy, main_lane, 64);
        st.qz = __shfl(st.qz, main_lane, 64); st.qw = __shfl(st.qw, main_lane, 64);
        st.w.x = __shfl(st.w.x, main_lane, 64); st.w.y = __shfl(st.w.y, main_lane, 64);
        st.w.z = __shfl(st.w.z, main_lane, 64);
        st.v.x = __shfl(st.v.x, main_lane, 64); st.v.y = __shfl(st.v.y, main_lane, 64);
        st.v.z = __shfl(st.v.z, main_lane, 64);
#pragma unroll
        for (int j = 0; j < 12; ++j) {
          st.q[j] = __shfl(st.q[j], main_lane, 64);
          st.qd[j] = __shfl(st.qd[j], main_lane, 64);
        }
      } else {
#pragma unroll
        for (int j = 0; j < 12; ++j) {
          mq[j] = __shfl(st.q[j], main_lane, 64);
          mqd[j] = __shfl(st.qd[j], main_lane, 64);
        }
        use_main_dof = member != 0;
      }
    }
  }
  if (active && member == 0) fim[group] = acc;
}

bool sim_args_ok(const SpiSimArgs* a) {
  return a && a->dt > 0.f && a->decimation >= 1 && a->decimation <= 64 && a->substeps >= 1 && a->substeps <= 64 &&
         a->motor_model >= 0 && a->motor_model <= 3 && a->state_limit > 0.f;
}

}  // namespace

extern "C" {

const char* spi_last_error(void) { return g_err; }
int spi_abi_version(void) { return SPI_ABI_VERSION; }
int spi_segment_stride(int horizon) { return SPI_STATE_DIM + horizon * (12 + SPI_STATE_DIM); }

int spi_physics_step(const SpiSimArgs* sim, const float* params_dev, int ld, int n, float* root_dev,
                     float* dof_state_dev, const float* torques_dev, float* foot_force_dev, hipStream_t stream) {
  if (!sim_args_ok(sim)) return fail(SPI_ERR_INVALID_ARG, "spi_physics_step: invalid SpiSimArgs");
  if (n < 0 || ld < n || !params_dev || !root_dev || !dof_state_dev || !torques_dev)
    return fail(SPI_ERR_INVALID_ARG, "spi_physics_step: bad sizes or null pointer");
  if (n == 0) return SPI_OK;
  physics_step_kernel<<<(n + 255) / 256, 256, 0, stream>>>(*sim, params_dev, ld, n, root_dev, dof_state_dev,
                                                          torques_dev, foot_force_dev);
  return launch_check("physics_step_kernel");
}

int spi_forward_dynamics(const SpiSimArgs* sim, const float* params_dev, int ld, int n, const float* root_dev,
                         const float* dof_state_dev, const float* torques_dev, float* acc_dev, float* foot_force_dev,
                         hipStream_t stream) {
  if (!sim_args_ok(sim)) return fail(SPI_ERR_INVALID_ARG, "spi_forward_dynamics: invalid SpiSimArgs");
  if (n < 0 || ld < n || !params_dev || !root_dev || !dof_state_dev || !torques_dev || !acc_dev)
    return fail(SPI_ERR_INVALID_ARG, "spi_forward_dynamics: bad sizes or null pointer");
  if (n == 0) return SPI_OK;
  forward_dynamics_kernel<<<(n + 255) / 256, 256, 0, stream>>>(*sim, params_dev, ld, n, root_dev, dof_state_dev,
                                                              torques_dev, acc_dev, foot_force_dev);
  return launch_check("forward_dynamics_kernel");
}

int spi_set_rows_indexed(float* dst_dev, const float* src_dev, const int* idx_dev, int n_idx, int n_rows, int row_len,
                         hipStream_t stream) {
  if (n_idx < 0 || n_rows < 0 || row_len <= 0 || !dst_dev || !src_dev || (n_idx > 0 && !idx_dev))
    return fail(SPI_ERR_INVALID_ARG, "spi_set_rows_indexed: bad arguments");
  if (n_idx == 0 || dst_dev == src_dev) return SPI_OK;
  long long total = static_cast<long long>(n_idx) * row_len;
  if (total > 0x7fffffffLL) return fail(SPI_ERR_TOO_LARGE, "spi_set_rows_indexed: too many elements");
  set_rows_kernel<<<static_cast<int>((total + 255) / 256), 256, 0, stream>>>(dst_dev, src_dev, idx_dev, n_idx, n_rows,
                                                                              row_len);
  return launch_check("set_rows_kernel");
}

int spi_rigid_body_states(const float* root_dev, const float* dof_state_dev, int n, float* out_dev,
                          hipStream_t stream) {
  if (n < 0 || !root_dev || !dof_state_dev || !out_dev) return fail(SPI_ERR_INVALID_ARG, "spi_rigid_body_states");
  if (n == 0) return SPI_OK;
  rigid_body_kernel<<<(n + 127) / 128, 128, 0, stream>>>(root_dev, dof_state_dev, n, out_dev);
  return launch_check("rigid_body_kernel");
}

int spi_compute_torques(const SpiSimArgs* sim, const float* params_dev, int ld, int n, const float* actions_dev,
                        const float* dof_state_dev, float* torques_dev, hipStream_t stream) {
  if (!sim_args_ok(sim)) return fail(SPI_ERR_INVALID_ARG, "spi_compute_torques: invalid SpiSimArgs");
  if (n < 0 || ld < n || !params_dev || !actions_dev || !dof_state_dev || !torques_dev)
    return fail(SPI_ERR_INVALID_ARG, "spi_compute_torques: bad sizes or null pointer");
  if (n == 0) return SPI_OK;
  compute_torques_kernel<<<(n + 255) / 256, 256, 0, stream>>>(*sim, params_dev, ld, n, actions_dev, dof_state_dev,
                                                             torques_dev);
  return launch_check("compute_torques_kernel");
}

int spi_fd_fim_trace(const float* root_dev, const float* dof_state_dev, const float* env_origins_dev, int n_main,
                     int n_param, float delta, float* out_dev, hipStream_t stream) {
  if (n_main < 0 || n_param < 1 || !(delta != 0.0f) || !root_dev || !dof_state_dev || !env_origins_dev || !out_dev)
    return fail(SPI_ERR_INVALID_ARG, "spi_fd_fim_trace: bad arguments");
  if (n_main == 0) return SPI_OK;
  fd_fim_kernel<<<(n_main + 127) / 128, 128, 0, stream>>>(root_dev, dof_state_dev, env_origins_dev, n_main, n_param,
                                                          delta, out_dev);
  return launch_check("fd_fim_kernel");
}

int spi_rollout_cost(const SpiSimArgs* sim, const SpiCostArgs* cost, const float* params_dev, int ld, int n_cand,
                     const float* seg_dev, int n_seg, int horizon, float* cost_seg_dev, hipStream_t stream) {
  if (!sim_args_ok(sim) || !cost || (cost->mode != SPI_COST_TRACKING && cost->mode != SPI_COST_REFERENCE))
    return fail(SPI_ERR_INVALID_ARG, "spi_rollout_cost: invalid SpiSimArgs/SpiCostArgs");
  if (n_cand < 0 || n_seg < 0 || horizon < 1 || ld < n_cand || !params_dev || !seg_dev || !cost_seg_dev)
    return fail(SPI_ERR_INVALID_ARG, "spi_rollout_cost: bad sizes or null pointer");
  if (n_cand == 0 || n_seg == 0) return SPI_OK;
  const int stride = spi_segment_stride(horizon);
  const size_t lds_bytes = static_cast<size_t>(stride) * sizeof(float);
  if (lds_bytes > 64 * 1024) return fail(SPI_ERR_TOO_LARGE, "spi_rollout_cost: horizon too long for LDS staging");
  const int n_cblk = (n_cand + kRolloutThreads - 1) / kRolloutThreads;
  const long long nblk = static_cast<long long>(n_cblk) * n_seg;
  if (nblk > 0x7fffffffLL || static_cast<long long>(n_seg) * n_cand > (1LL << 31))
    return fail(SPI_ERR_TOO_LARGE, "spi_rollout_cost: grid too large");
  rollout_cost_kernel<<<static_cast<int>(nblk), kRolloutThreads, lds_bytes, stream>>>(
      *sim, *cost, params_dev, ld, n_cand, seg_dev, horizon, stride, n_cblk, cost_seg_dev);
  return launch_check("rollout_cost_kernel");
}

int spi_rollout_cost_legs(const SpiSimArgs* sim, const SpiCostArgs* cost, const float* params_dev, int ld,
                          int n_cand, const float* seg_dev, int n_seg, int horizon, float* cost_seg_dev,
                          hipStream_t stream) {
  if (!sim_args_ok(sim) || !cost || (cost->mode != SPI_COST_TRACKING && cost->mode != SPI_COST_REFERENCE))
    return fail(SPI_ERR_INVALID_ARG, "spi_rollout_cost_legs: invalid SpiSimArgs/SpiCostArgs");
  if (n_cand < 0 || n_seg < 0 || horizon < 1 || ld < n_cand || !params_dev || !seg_dev || !cost_seg_dev)
    return fail(SPI_ERR_INVALID_ARG, "spi_rollout_cost_legs: bad sizes or null pointer");
  if (n_cand == 0 || n_seg == 0) return SPI_OK;
  const int stride = spi_segment_stride(horizon);
  const size_t lds_bytes = static_cast<size_t>(stride) * sizeof(float);
  if (lds_bytes > 64 * 1024) return fail(SPI_ERR_TOO_LARGE, "spi_rollout_cost_legs: horizon too long for LDS staging");
  const int per_block = kRolloutThreads / 4;
  const int n_cblk = (n_cand + per_block - 1) / per_block;
  const long long nblk = static_cast<long long>(n_cblk) * n_seg;
  if (nblk > 0x7fffffffLL || static_cast<long long>(n_seg) * n_cand > (1LL << 31))
    return fail(SPI_ERR_TOO_LARGE, "spi_rollout_cost_legs: grid too large");
  rollout_cost_leg_kernel<<<static_cast<int>(nblk), kRolloutThreads, lds_bytes, stream>>>(
      *sim, *cost, params_dev, ld, n_cand, seg_dev, horizon, stride, n_cblk, cost_seg_dev);
  return launch_check("rollout_cost_leg_kernel");
}

int spi_fim_rollout(const SpiSimArgs* sim, const float* params_dev, int ld, int n_group, int n_param,
                    const float* init_dev, const float* actions_dev, int n_steps, float delta, int ksync,
                    int ksync_mode, float* fim_dev, hipStream_t stream) {
  if (!sim_args_ok(sim)) return fail(SPI_ERR_INVALID_ARG, "spi_fim_rollout: invalid SpiSimArgs");
  if (n_group < 0 || n_param < 1 || n_param + 1 > 64 || n_steps < 1 || !(delta != 0.0f) ||
      (ksync_mode != 0 && ksync_mode != 1) || !params_dev || !init_dev || !actions_dev || !fim_dev)
    return fail(SPI_ERR_INVALID_ARG, "spi_fim_rollout: bad arguments");
  const long long n_env = static_cast<long long>(n_group) * (n_param + 1);
  if (ld < n_env) return fail(SPI_ERR_INVALID_ARG, "spi_fim_rollout: ld < number of envs");
  if (n_group == 0) return SPI_OK;
  int G2 = 1;
  while (G2 < n_param + 1) G2 <<= 1;
  const long long lanes = static_cast<long long>(n_group) * G2;
  if (lanes > 0x7fffffffLL) return fail(SPI_ERR_TOO_LARGE, "spi_fim_rollout: too many groups");
  // blocks are whole wavefronts so groups never straddle a wavefront (G2 divides 64)
  const int blocks = static_cast<int>((lanes + 255) / 256);
  fim_rollout_kernel<<<blocks, 256, 0, stream>>>(*sim, params_dev, ld, n_group, n_param, G2, init_dev, actions_dev,
                                                 n_steps, delta, ksync, ksync_mode, fim_dev);
  return launch_check("fim_rollout_kernel");
}

int spi_batch_rollout_cost(const SpiSimArgs* sim, const SpiBatchProblem* probs_dev, const long long* task_prefix_dev,
                           int n_prob, long long n_tasks, int max_horizon, int block, int unroll, int n_blocks,
                           unsigned long long* queue_dev, hipStream_t stream) {
  if (!sim_args_ok(sim)) return fail(SPI_ERR_INVALID_ARG, "spi_batch_rollout_cost: invalid SpiSimArgs");
  if (n_prob < 1 || n_tasks < 0 || max_horizon < 1 || n_blocks < 1 || !probs_dev || !task_prefix_dev || !queue_dev)
    return fail(SPI_ERR_INVALID_ARG, "spi_batch_rollout_cost: bad arguments");
  if (n_tasks == 0) return SPI_OK;
  const size_t lds = static_cast<size_t>(spi_segment_stride(max_horizon)) * sizeof(float);
  if (lds > 64 * 1024) return fail(SPI_ERR_TOO_LARGE, "spi_batch_rollout_cost: horizon too long for LDS staging");
#define SPI_BATCH_LAUNCH(B, U)                                                                               \
  rollout_cost_batch_kernel<B, U><<<n_blocks, B, lds, stream>>>(*sim, probs_dev, task_prefix_dev, n_prob, n_tasks, \
                                                                queue_dev)
  if (block == 64 && unroll == 1) SPI_BATCH_LAUNCH(64, 1);
  else if (block == 128 && unroll == 1) SPI_BATCH_LAUNCH(128, 1);
  else if (block == 256 && unroll == 1) SPI_BATCH_LAUNCH(256, 1);
  else if (block == 64 && unroll == 2) SPI_BATCH_LAUNCH(64, 2);
  else if (block == 128 && unroll == 2) SPI_BATCH_LAUNCH(128, 2);
  else if (block == 256 && unroll == 2) SPI_BATCH_LAUNCH(256, 2);
  else return fail(SPI_ERR_INVALID_ARG, "spi_batch_rollout_cost: block must be 64/128/256, unroll 1/2");
#undef SPI_BATCH_LAUNCH
  return launch_check("rollout_cost_batch_kernel");
}

int spi_simulate(const SpiSimArgs* sim, const float* params_dev, int ld, int n, const float* init_dev,
                 const float* actions_dev, int n_steps, float* states_dev, hipStream_t stream) {
  if (!sim_args_ok(sim)) return fail(SPI_ERR_INVALID_ARG, "spi_simulate: invalid SpiSimArgs");
  if (n < 0 || ld < n || n_steps < 0 || !params_dev || !init_dev || !actions_dev || !states_dev)
    return fail(SPI_ERR_INVALID_ARG, "spi_simulate: bad sizes or null pointer");
  if (n == 0 || n_steps == 0) return SPI_OK;
  simulate_kernel<<<(n + 255) / 256, 256, 0, stream>>>(*sim, params_dev, ld, n, init_dev, actions_dev, n_steps,
                                                       states_dev);
  return launch_check("simulate_kernel");
}

}  // extern "C"
