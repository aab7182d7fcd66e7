// Policy-in-the-loop active-exploration rollout (gfx950): a locomotion policy drives groups of
// (1 main + P aux) envs through open-loop command sequences while the finite-difference Fisher
// information between each main env and its aux envs accumulates
// (spigym/agents/sysid/active_sysid.py:527-600 over spigym/envs/sysid/active_sysid_openloop.py).
//
// Per control step the host launches: policy_step_kernel (actions -> decimation physics steps,
// the joint-state sync quirk on the first torque), policy_reward_kernel (gravity termination
// OR-ed over each group, ||X||_F^2 or the termination reward), policy_obs_kernel (gait clock,
// observation terms, short history).  The policy network between steps lives in spi_mlp.hip.
// Built with -ffinite-math-only (dynamics): non-finite checks use exponent bits.
#include <hip/hip_runtime.h>

#include <cstdio>

#include "spi_dyn.h"
#include "spi_dyn_leg.h"
#include "spi_launch.h"

using namespace spi;
using spi_launch::finite_bits;

namespace {

constexpr int kPolThreads = 256;
constexpr int kTerms = 60;  // per-step observation terms, keys in sorted order
constexpr int kHist = 14;   // short-history length (newest first)
constexpr int kObs = kTerms * (1 + kHist);
constexpr int kKeys = 14;
constexpr float kTwoPi = 6.28318530717958647692f;

// Sorted observation keys of the go2_omni actor (config/obs/loco/go2_omni.yaml): offsets and
// widths inside the 60 current terms.
//  actions 12 | base_ang_vel 3 | clock_inputs 4 | command_ang_vel 1 | command_body_attitude 2 |
//  command_body_height 1 | command_footswing_height 1 | command_gait_freq 1 |
//  command_gait_phase 4 | command_lin_vel 2 | command_stance 2 | dof_pos 12 | dof_vel 12 |
//  projected_gravity 3
__constant__ int kOff[kKeys] = {0, 12, 15, 19, 20, 22, 23, 24, 25, 29, 31, 33, 45, 57};
__constant__ int kDim[kKeys] = {12, 3, 4, 1, 2, 1, 1, 1, 4, 2, 2, 12, 12, 3};
constexpr int kOffHost[kKeys] = {0, 12, 15, 19, 20, 22, 23, 24, 25, 29, 31, 33, 45, 57};
constexpr int kDimHost[kKeys] = {12, 3, 4, 1, 2, 1, 1, 1, 4, 2, 2, 12, 12, 3};

__device__ __forceinline__ LaneParams load_lane(const float* __restrict__ par, int ld, int i) {
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

// v expressed in the frame of the unit quaternion (x, y, z, w) (torch_utils.py quat_rotate_inverse).
__device__ __forceinline__ void quat_rotate_inverse(float qx, float qy, float qz, float qw, float vx, float vy,
                                                    float vz, float* o) {
  const float c2 = 2.0f * qw * qw - 1.0f;
  const float cx = qy * vz - qz * vy, cy = qz * vx - qx * vz, cz = qx * vy - qy * vx;
  const float d = 2.0f * (qx * vx + qy * vy + qz * vz);
  o[0] = vx * c2 - cx * qw * 2.0f + qx * d;
  o[1] = vy * c2 - cy * qw * 2.0f + qy * d;
  o[2] = vz * c2 - cz * qw * 2.0f + qz * d;
}

// One control step of every env.  Envs terminated at the previous step get zero actions
// (active_sysid.py:556-558); the applied (zeroed, clipped) actions are what the observation shows
// (legged_robot_base.py:185-198, 843-844).  sync_first: the previous step copied the main env's
// joint state into its aux envs' joint-state views without pushing it to the simulator
// (active_sysid_openloop.py:246-252, 316-330), so only the first torque of this step sees it.
// A lane whose state leaves the finite range or the divergence bound sets bad[i].
__global__ __launch_bounds__(kPolThreads) void policy_step_kernel(SpiSimArgs a, const float* __restrict__ par, int ld,
                                                                  int n, int group, const float* __restrict__ s_in,
                                                                  float* __restrict__ s_out,
                                                                  const float* __restrict__ actions,
                                                                  const int* __restrict__ done, int sync_first,
                                                                  float* __restrict__ applied, int* __restrict__ bad) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const LaneParams lp = load_lane(par, ld, i);
  const float* src = s_in + static_cast<size_t>(i) * SPI_STATE_DIM;
  State st;
  state_from_root(st, src, src + 13, src + 25, 1);
  float act[12];
  const bool stop = done[i] != 0;
#pragma unroll
  for (int j = 0; j < 12; ++j) {
    const float x = stop ? 0.0f : actions[static_cast<size_t>(i) * 12 + j];
    act[j] = fminf(fmaxf(x, -a.action_clip), a.action_clip);
    applied[static_cast<size_t>(i) * 12 + j] = act[j];
  }
  const int m = i - i % group;
  const bool view = sync_first && i != m;
  for (int d = 0; d < a.decimation; ++d) {
    float tau[12];
    if (d == 0 && view) {
      const float* ms = s_in + static_cast<size_t>(m) * SPI_STATE_DIM;
      State v = st;
#pragma unroll
      for (int j = 0; j < 12; ++j) { v.q[j] = ms[13 + j]; v.qd[j] = ms[25 + j]; }
      control_torques(a, lp, act, v, tau);
    } else {
      control_torques(a, lp, act, st, tau);
    }
    physics_step(a, lp, st, tau, nullptr);
  }
  const float nrm = state_norm(st);
  if (!finite_bits(nrm) || nrm >= a.state_limit) bad[i] = 1;
  float* dst = s_out + static_cast<size_t>(i) * SPI_STATE_DIM;
  state_to_root(st, dst, dst + 13, dst + 25, 1);
}

// policy_step_kernel with four lanes per env (lane % 4 = leg, spi_dyn_leg.h): the same step,
// for batches too small to fill the SIMDs with one env per lane.
__device__ __forceinline__ float pick4(int l, float x0, float x1, float x2, float x3) {
  return l == 0 ? x0 : (l == 1 ? x1 : (l == 2 ? x2 : x3));
}
__global__ __launch_bounds__(kPolThreads) void policy_step_leg_kernel(SpiSimArgs a, const float* __restrict__ par,
                                                                      int ld, int n, int group,
                                                                      const float* __restrict__ s_in,
                                                                      float* __restrict__ s_out,
                                                                      const float* __restrict__ actions,
                                                                      const int* __restrict__ done, int sync_first,
                                                                      float* __restrict__ applied,
                                                                      int* __restrict__ bad) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int i = t >> 2, l = t & 3;
  if (i >= n) return;  // the four lanes of an env leave together
  LegParams lp;
  lp.base = make_base_inertia(par[0 * ld + i], par[1 * ld + i], par[2 * ld + i], par[3 * ld + i], par[4 * ld + i],
                              par[5 * ld + i], par[6 * ld + i], par[7 * ld + i], par[8 * ld + i], par[9 * ld + i]);
  lp.kp_s = par[(10 + l) * ld + i];
  lp.kd_s = par[(14 + l) * ld + i];
  lp.fric = par[(18 + l) * ld + i];
#pragma unroll
  for (int g = 0; g < 3; ++g) lp.motor[g] = par[(22 + g) * ld + i];
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
  leg_state_from_root(st, s_in + static_cast<size_t>(i) * SPI_STATE_DIM, l);
  float act[3];
  const bool stop = done[i] != 0;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const float x = stop ? 0.0f : actions[static_cast<size_t>(i) * 12 + 3 * l + k];
    act[k] = fminf(fmaxf(x, -a.action_clip), a.action_clip);
    applied[static_cast<size_t>(i) * 12 + 3 * l + k] = act[k];
  }
  const int m = i - i % group;
  const bool view = sync_first && i != m;
  for (int d = 0; d < a.decimation; ++d) {
    float tau[3];
    if (d == 0 && view) {
      const float* ms = s_in + static_cast<size_t>(m) * SPI_STATE_DIM;
      float vq[3], vqd[3];
#pragma unroll
      for (int k = 0; k < 3; ++k) { vq[k] = ms[13 + 3 * l + k]; vqd[k] = ms[25 + 3 * l + k]; }
      control_torques_leg(a, lp, kp, kd, q0, eff, act, vq, vqd, tau);
    } else {
      control_torques_leg(a, lp, kp, kd, q0, eff, act, st.q, st.qd, tau);
    }
    physics_step_leg(a, lp, K, st, tau);
  }
  const float nrm = state_norm_leg(st);
  if (l == 0 && (!finite_bits(nrm) || nrm >= a.state_limit)) bad[i] = 1;
  leg_state_to_root(st, s_out + static_cast<size_t>(i) * SPI_STATE_DIM, l);
}

// Per group: gravity termination of every env (|g_b.x| or |g_b.y| above the limits,
// legged_robot_base.py:336-339) OR-ed over the group (active_sysid_openloop.py:259-272) and
// written to every env of the group; then the step's reward: the termination reward for a
// terminated group (active_sysid.py:566-568), else ||X||_F^2 with X = (x_main - x_aux) / delta
// over root state (13) and joint positions (12) of all aux envs (active_sysid_openloop.py:402-426).
// A non-finite reward or a diverged member marks the group bad.
__device__ __forceinline__ void group_reward(const float* __restrict__ s, int g, int group, float delta, float grav_x,
                                             float grav_y, float termination_rew, int* __restrict__ done,
                                             const int* __restrict__ bad, float* __restrict__ acc,
                                             int* __restrict__ group_bad) {
  const size_t m = static_cast<size_t>(g) * group;
  bool term = false, diverged = false;
  for (int k = 0; k < group; ++k) {
    const float* x = s + (m + k) * SPI_STATE_DIM;
    float pg[3];
    quat_rotate_inverse(x[3], x[4], x[5], x[6], 0.0f, 0.0f, -1.0f, pg);
    term = term || fabsf(pg[0]) > grav_x || fabsf(pg[1]) > grav_y;
    diverged = diverged || bad[m + k] != 0;
  }
  for (int k = 0; k < group; ++k) done[m + k] = term ? 1 : 0;
  const float inv = 1.0f / delta;
  const float* xm = s + m * SPI_STATE_DIM;
  float sum = 0.0f;
  for (int k = 1; k < group; ++k) {
    const float* xa = s + (m + k) * SPI_STATE_DIM;
#pragma unroll
    for (int e = 0; e < 25; ++e) {
      const float d = (xm[e] - xa[e]) * inv;
      sum += d * d;
    }
  }
  if (diverged || !finite_bits(sum)) group_bad[g] = 1;
  acc[g] += term ? termination_rew : sum;
}

__global__ void policy_reward_kernel(const float* __restrict__ s, int n_group, int group, float delta,
                                     float grav_x, float grav_y, float termination_rew, int* __restrict__ done,
                                     const int* __restrict__ bad, float* __restrict__ acc,
                                     int* __restrict__ group_bad) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= n_group) return;
  group_reward(s, g, group, delta, grav_x, grav_y, termination_rew, done, bad, acc, group_bad);
}

// Gait clock of go2_omni (go2_omni.py:348-379: gait index advanced by dt * frequency, foot
// indices from phase / offset / bound, warped by the stance duration) and the scaled
// observation terms in sorted key order (legged_robot_base.py:795-805, 843-844;
// go2_locomotion.py:273-277; go2_omni.py:627-668 with the evaluation-time body-height
// sinusoid).  cn: the commands the step ran with (gait clock); cx: the commands after the
// playback pointer advanced (observation, active_sysid_openloop.py:196-201).
__device__ __forceinline__ void obs_terms(const float* __restrict__ x, const float* __restrict__ cn,
                                          const float* __restrict__ cx, float dt_ctrl,
                                          const float* __restrict__ q_default, float* __restrict__ gait_i,
                                          const float* __restrict__ applied_i, int evaluating, float h_lo, float h_hi,
                                          float* cur) {
  float gi = *gait_i + dt_ctrl * cn[4];
  gi = gi - floorf(gi);
  *gait_i = gi;
  const float feet[4] = {gi + cn[5] + cn[6] + cn[7], gi + cn[6], gi + cn[7], gi + cn[5]};
  const float dur = cn[8];
  float clock[4];
#pragma unroll
  for (int f = 0; f < 4; ++f) {
    float idx = feet[f];
    const float r = idx - floorf(idx);
    if (r < dur) idx = r * (0.5f / dur);
    else if (r > dur) idx = 0.5f + (r - dur) * (0.5f / (1.0f - dur));
    clock[f] = sinf(kTwoPi * idx);
  }
  float ang[3], pg[3];
  quat_rotate_inverse(x[3], x[4], x[5], x[6], x[10], x[11], x[12], ang);
  quat_rotate_inverse(x[3], x[4], x[5], x[6], 0.0f, 0.0f, -1.0f, pg);
  float h = cx[3];
  if (evaluating) h = fminf(fmaxf(cx[3] + 0.20f * sinf(kTwoPi * gi), h_lo), h_hi);
#pragma unroll
  for (int j = 0; j < 12; ++j) {
    cur[j] = applied_i[j];                    // actions
    cur[33 + j] = x[13 + j] - q_default[j];   // dof_pos
    cur[45 + j] = 0.05f * x[25 + j];          // dof_vel
  }
  cur[12] = 0.25f * ang[0]; cur[13] = 0.25f * ang[1]; cur[14] = 0.25f * ang[2];  // base_ang_vel
#pragma unroll
  for (int f = 0; f < 4; ++f) cur[15 + f] = clock[f];                   // clock_inputs
  cur[19] = cx[2];                                                      // command_ang_vel
  cur[20] = 0.3f * cx[10]; cur[21] = 0.3f * cx[11];                     // command_body_attitude
  cur[22] = 2.0f * h;                                                   // command_body_height
  cur[23] = 0.15f * cx[9];                                              // command_footswing_height
  cur[24] = cx[4];                                                      // command_gait_freq
  cur[25] = cx[5]; cur[26] = cx[6]; cur[27] = cx[7]; cur[28] = cx[8];   // command_gait_phase
  cur[29] = cx[0]; cur[30] = cx[1];                                     // command_lin_vel
  cur[31] = cx[12]; cur[32] = cx[13];                                   // command_stance
  cur[57] = pg[0]; cur[58] = pg[1]; cur[59] = pg[2];                    // projected_gravity
}

// Dense observation: sorted terms + short history per key (legged_robot_base.py:503-527,
// 819-829), clipped (legged_base.yaml clip_observations), then the history shift
// (history_handler.py:36-40).  hist: [n][kHist][kTerms], newest first; obs [n][kObs].
__global__ __launch_bounds__(kPolThreads) void policy_obs_kernel(
    int n, int group, const float* __restrict__ s, const float* __restrict__ cmds, int T, int s_now, int s_next,
    float dt_ctrl, const float* __restrict__ q_default, float* __restrict__ gait, const float* __restrict__ applied,
    float* __restrict__ hist, float* __restrict__ obs, int evaluating, float h_lo, float h_hi, float obs_clip) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int g = i / group;
  float cur[kTerms];
  obs_terms(s + static_cast<size_t>(i) * SPI_STATE_DIM, cmds + (static_cast<size_t>(g) * T + s_now) * 14,
            cmds + (static_cast<size_t>(g) * T + s_next) * 14, dt_ctrl, q_default, gait + i,
            applied + static_cast<size_t>(i) * 12, evaluating, h_lo, h_hi, cur);
  float* o = obs + static_cast<size_t>(i) * kObs;
#pragma unroll
  for (int e = 0; e < kTerms; ++e) o[e] = fminf(fmaxf(cur[e], -obs_clip), obs_clip);
  float* hb = hist + static_cast<size_t>(i) * kHist * kTerms;
  int w = kTerms;
  for (int k = 0; k < kKeys; ++k) {
    const int off = kOff[k], dim = kDim[k];
    for (int t = 0; t < kHist; ++t)
      for (int e = 0; e < dim; ++e) o[w++] = fminf(fmaxf(hb[t * kTerms + off + e], -obs_clip), obs_clip);
  }
  for (int t = kHist - 1; t > 0; --t)
    for (int e = 0; e < kTerms; ++e) hb[t * kTerms + e] = hb[(t - 1) * kTerms + e];
#pragma unroll
  for (int e = 0; e < kTerms; ++e) hb[e] = cur[e];
}

// Ring observation: this step's terms go to slot `head` of ring[kHist + 1][kTerms][ld] (env
// index fastest, so the writes of a wavefront are coalesced).  The observation row is then
// [slot head | slot head - 1 - t of each key, t < kHist], which the actor's first layer
// gathers directly (spi_mlp_forward_ring with spi_policy_obs_kmap); nothing is shifted.
__global__ __launch_bounds__(kPolThreads) void policy_terms_kernel(
    int n, int group, const float* __restrict__ s, const float* __restrict__ cmds, int T, int s_now, int s_next,
    float dt_ctrl, const float* __restrict__ q_default, float* __restrict__ gait, const float* __restrict__ applied,
    float* __restrict__ ring, int ld, int head, int evaluating, float h_lo, float h_hi) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int g = i / group;
  float cur[kTerms];
  obs_terms(s + static_cast<size_t>(i) * SPI_STATE_DIM, cmds + (static_cast<size_t>(g) * T + s_now) * 14,
            cmds + (static_cast<size_t>(g) * T + s_next) * 14, dt_ctrl, q_default, gait + i,
            applied + static_cast<size_t>(i) * 12, evaluating, h_lo, h_hi, cur);
  float* r = ring + static_cast<size_t>(head) * kTerms * ld + i;
#pragma unroll
  for (int e = 0; e < kTerms; ++e) r[static_cast<size_t>(e) * ld] = cur[e];
}

// policy_reward_kernel and policy_terms_kernel in one launch (both read only the post-step
// state): one lane per env writes its terms; the lane of each group's main env also computes
// the group's termination and reward.
__global__ __launch_bounds__(kPolThreads) void policy_reward_terms_kernel(
    int n, int group, const float* __restrict__ s, const float* __restrict__ cmds, int T, int s_now, int s_next,
    float dt_ctrl, const float* __restrict__ q_default, float* __restrict__ gait, const float* __restrict__ applied,
    float* __restrict__ ring, int ld, int head, int evaluating, float h_lo, float h_hi, float delta, float grav_x,
    float grav_y, float termination_rew, int* __restrict__ done, const int* __restrict__ bad, float* __restrict__ acc,
    int* __restrict__ group_bad) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int g = i / group;
  if (i % group == 0) group_reward(s, g, group, delta, grav_x, grav_y, termination_rew, done, bad, acc, group_bad);
  float cur[kTerms];
  obs_terms(s + static_cast<size_t>(i) * SPI_STATE_DIM, cmds + (static_cast<size_t>(g) * T + s_now) * 14,
            cmds + (static_cast<size_t>(g) * T + s_next) * 14, dt_ctrl, q_default, gait + i,
            applied + static_cast<size_t>(i) * 12, evaluating, h_lo, h_hi, cur);
  float* r = ring + static_cast<size_t>(head) * kTerms * ld + i;
#pragma unroll
  for (int e = 0; e < kTerms; ++e) r[static_cast<size_t>(e) * ld] = cur[e];
}

// Full state copy main -> aux envs (the state-sync variant; the reference's sync only reaches the
// first torque, see policy_step_kernel).
__global__ void policy_sync_state_kernel(float* __restrict__ s, int n, int group) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || i % group == 0) return;
  const float* m = s + static_cast<size_t>(i - i % group) * SPI_STATE_DIM;
  float* d = s + static_cast<size_t>(i) * SPI_STATE_DIM;
  for (int e = 0; e < SPI_STATE_DIM; ++e) d[e] = m[e];
}

int pfail(const char* msg) {
  spi_set_error(msg);
  return SPI_ERR_INVALID_ARG;
}
int plaunch(const char* what) {
  const hipError_t e = hipGetLastError();
  if (e == hipSuccess) return SPI_OK;
  char buf[256];
  std::snprintf(buf, sizeof(buf), "%s: %s", what, hipGetErrorString(e));
  spi_set_error(buf);
  return SPI_ERR_HIP;
}
bool groups_ok(long long n, int group) { return n >= 0 && group >= 1 && n % group == 0 && n <= 0x7fffffffLL; }

}  // namespace

extern "C" {

int spi_policy_obs_dim(void) { return kObs; }

int spi_policy_step(const SpiSimArgs* sim, const float* params_dev, int ld, int n_env, int group,
                    const float* states_in_dev, float* states_out_dev, const float* actions_dev, const int* done_dev,
                    int sync_first, float* applied_dev, int* bad_dev, hipStream_t stream) {
  if (!sim || sim->decimation < 1 || sim->decimation > 64 || sim->substeps < 1 || sim->substeps > 64 ||
      !(sim->dt > 0.0f) || sim->motor_model < 0 || sim->motor_model > 3 || !(sim->state_limit > 0.0f))
    return pfail("spi_policy_step: invalid SpiSimArgs");
  if (!groups_ok(n_env, group) || ld < n_env || !params_dev || !states_in_dev || !states_out_dev || !actions_dev ||
      !done_dev || !applied_dev || !bad_dev || states_in_dev == states_out_dev)
    return pfail("spi_policy_step: bad sizes, null pointers or aliased state buffers");
  if (n_env == 0) return SPI_OK;
  policy_step_kernel<<<(n_env + kPolThreads - 1) / kPolThreads, kPolThreads, 0, stream>>>(
      *sim, params_dev, ld, n_env, group, states_in_dev, states_out_dev, actions_dev, done_dev, sync_first, applied_dev,
      bad_dev);
  return plaunch("policy_step_kernel");
}

int spi_policy_step_legs(const SpiSimArgs* sim, const float* params_dev, int ld, int n_env, int group,
                         const float* states_in_dev, float* states_out_dev, const float* actions_dev,
                         const int* done_dev, int sync_first, float* applied_dev, int* bad_dev, hipStream_t stream) {
  if (!sim || sim->decimation < 1 || sim->decimation > 64 || sim->substeps < 1 || sim->substeps > 64 ||
      !(sim->dt > 0.0f) || sim->motor_model < 0 || sim->motor_model > 3 || !(sim->state_limit > 0.0f))
    return pfail("spi_policy_step_legs: invalid SpiSimArgs");
  if (!groups_ok(n_env, group) || n_env > 0x1fffffff || ld < n_env || !params_dev || !states_in_dev ||
      !states_out_dev || !actions_dev || !done_dev || !applied_dev || !bad_dev || states_in_dev == states_out_dev)
    return pfail("spi_policy_step_legs: bad sizes, null pointers or aliased state buffers");
  if (n_env == 0) return SPI_OK;
  const long long lanes = 4LL * n_env;
  policy_step_leg_kernel<<<static_cast<int>((lanes + kPolThreads - 1) / kPolThreads), kPolThreads, 0, stream>>>(
      *sim, params_dev, ld, n_env, group, states_in_dev, states_out_dev, actions_dev, done_dev, sync_first,
      applied_dev, bad_dev);
  return plaunch("policy_step_leg_kernel");
}

int spi_policy_reward(const float* states_dev, int n_group, int group, float delta, float grav_x, float grav_y,
                      float termination_rew, int* done_dev, const int* bad_dev, float* acc_dev, int* group_bad_dev,
                      hipStream_t stream) {
  if (n_group < 0 || group < 2 || !(delta != 0.0f) || !states_dev || !done_dev || !bad_dev || !acc_dev ||
      !group_bad_dev || static_cast<long long>(n_group) * group > 0x7fffffffLL)
    return pfail("spi_policy_reward: bad arguments");
  if (n_group == 0) return SPI_OK;
  policy_reward_kernel<<<(n_group + 255) / 256, 256, 0, stream>>>(states_dev, n_group, group, delta, grav_x, grav_y,
                                                                  termination_rew, done_dev, bad_dev, acc_dev,
                                                                  group_bad_dev);
  return plaunch("policy_reward_kernel");
}

int spi_policy_obs(int n_env, int group, const float* states_dev, const float* commands_dev, int n_cmd_steps,
                   int step_now, int step_next, float dt_ctrl, const float* q_default_dev, float* gait_dev,
                   const float* applied_dev, float* history_dev, float* obs_dev, int evaluating, float height_lo,
                   float height_hi, float obs_clip, hipStream_t stream) {
  if (!groups_ok(n_env, group) || n_cmd_steps < 1 || step_now < 0 || step_next < 0 || step_now >= n_cmd_steps ||
      step_next >= n_cmd_steps || !(dt_ctrl > 0.0f) || !(obs_clip > 0.0f) || !states_dev || !commands_dev ||
      !q_default_dev || !gait_dev || !applied_dev || !history_dev || !obs_dev)
    return pfail("spi_policy_obs: bad arguments");
  if (n_env == 0) return SPI_OK;
  policy_obs_kernel<<<(n_env + kPolThreads - 1) / kPolThreads, kPolThreads, 0, stream>>>(
      n_env, group, states_dev, commands_dev, n_cmd_steps, step_now, step_next, dt_ctrl, q_default_dev, gait_dev,
      applied_dev, history_dev, obs_dev, evaluating, height_lo, height_hi, obs_clip);
  return plaunch("policy_obs_kernel");
}

int spi_policy_terms(int n_env, int group, const float* states_dev, const float* commands_dev, int n_cmd_steps,
                     int step_now, int step_next, float dt_ctrl, const float* q_default_dev, float* gait_dev,
                     const float* applied_dev, float* ring_dev, int ld, int head, int evaluating, float height_lo,
                     float height_hi, hipStream_t stream) {
  if (!groups_ok(n_env, group) || n_cmd_steps < 1 || step_now < 0 || step_next < 0 || step_now >= n_cmd_steps ||
      step_next >= n_cmd_steps || !(dt_ctrl > 0.0f) || ld < n_env || head < 0 || head > kHist || !states_dev ||
      !commands_dev || !q_default_dev || !gait_dev || !applied_dev || !ring_dev)
    return pfail("spi_policy_terms: bad arguments");
  if (n_env == 0) return SPI_OK;
  policy_terms_kernel<<<(n_env + kPolThreads - 1) / kPolThreads, kPolThreads, 0, stream>>>(
      n_env, group, states_dev, commands_dev, n_cmd_steps, step_now, step_next, dt_ctrl, q_default_dev, gait_dev,
      applied_dev, ring_dev, ld, head, evaluating, height_lo, height_hi);
  return plaunch("policy_terms_kernel");
}

int spi_policy_reward_terms(int n_env, int group, const float* states_dev, const float* commands_dev,
                            int n_cmd_steps, int step_now, int step_next, float dt_ctrl, const float* q_default_dev,
                            float* gait_dev, const float* applied_dev, float* ring_dev, int ld, int head,
                            int evaluating, float height_lo, float height_hi, float delta, float grav_x, float grav_y,
                            float termination_rew, int* done_dev, const int* bad_dev, float* acc_dev,
                            int* group_bad_dev, hipStream_t stream) {
  if (!groups_ok(n_env, group) || group < 2 || n_cmd_steps < 1 || step_now < 0 || step_next < 0 ||
      step_now >= n_cmd_steps || step_next >= n_cmd_steps || !(dt_ctrl > 0.0f) || ld < n_env || head < 0 ||
      head > kHist || !(delta != 0.0f) || !states_dev || !commands_dev || !q_default_dev || !gait_dev ||
      !applied_dev || !ring_dev || !done_dev || !bad_dev || !acc_dev || !group_bad_dev)
    return pfail("spi_policy_reward_terms: bad arguments");
  if (n_env == 0) return SPI_OK;
  policy_reward_terms_kernel<<<(n_env + kPolThreads - 1) / kPolThreads, kPolThreads, 0, stream>>>(
      n_env, group, states_dev, commands_dev, n_cmd_steps, step_now, step_next, dt_ctrl, q_default_dev, gait_dev,
      applied_dev, ring_dev, ld, head, evaluating, height_lo, height_hi, delta, grav_x, grav_y, termination_rew,
      done_dev, bad_dev, acc_dev, group_bad_dev);
  return plaunch("policy_reward_terms_kernel");
}

int spi_policy_ring_slots(void) { return kHist + 1; }

int spi_policy_obs_kmap(int* kmap_host) {
  if (!kmap_host) return pfail("spi_policy_obs_kmap: null output");
  int w = 0;
  for (int e = 0; e < kTerms; ++e, ++w) { kmap_host[2 * w] = 0; kmap_host[2 * w + 1] = e; }
  for (int k = 0; k < kKeys; ++k)
    for (int t = 0; t < kHist; ++t)
      for (int e = 0; e < kDimHost[k]; ++e, ++w) { kmap_host[2 * w] = 1 + t; kmap_host[2 * w + 1] = kOffHost[k] + e; }
  return w == kObs ? SPI_OK : pfail("spi_policy_obs_kmap: layout mismatch");
}

int spi_policy_sync_state(float* states_dev, int n_env, int group, hipStream_t stream) {
  if (!groups_ok(n_env, group) || !states_dev) return pfail("spi_policy_sync_state: bad arguments");
  if (n_env == 0) return SPI_OK;
  policy_sync_state_kernel<<<(n_env + 255) / 256, 256, 0, stream>>>(states_dev, n_env, group);
  return plaunch("policy_sync_state_kernel");
}

}  // extern "C"
