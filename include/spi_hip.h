/*
 * spi_hip.h — C ABI of the MI355X (gfx950) engine for SPI-Active's sysID hot path:
 * batched Go2 rollouts under sampled physical parameters, scored by a tracking cost and
 * driven by a CEM update.
 *
 * Plain pointers and sizes only.  Every pointer argument named *_dev is device memory
 * (e.g. a torch tensor's data_ptr()); every entry point takes the hipStream_t it is
 * enqueued on (pass torch.cuda.current_stream().cuda_stream) and never synchronises it
 * unless its name ends in _sync.  Return value: SPI_OK or an SpiStatus error code;
 * spi_last_error() describes the last failure of the calling thread.
 *
 * Parameter vectors are SoA: params_dev[p * ld + i] is parameter p of env/candidate i,
 * p in [0, SPI_P_COUNT).  Root states use IsaacGym's layout (isaacgym.py:567-572):
 * [pos(3), quat xyzw(4), linvel(3), angvel(3)] in world axes, linvel = velocity of the
 * base frame origin.  DOF state is [n][12][2] = (q, qd) as isaacgym.py:574-576.
 */
#ifndef SPI_HIP_H
#define SPI_HIP_H

#include <hip/hip_runtime_api.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SPI_ABI_VERSION 2
#define SPI_NUM_DOF 12
#define SPI_ROOT_DIM 13
#define SPI_STATE_DIM 37 /* root(13) + q(12) + qd(12) */
#define SPI_MAX_TANGENTS 32
#define SPI_NUM_FEET 4
#define SPI_NUM_ISAAC_BODIES 19
#define SPI_P_COUNT 25
#define SPI_MAX_CEM_SAMPLES 8192

typedef enum SpiStatus {
  SPI_OK = 0,
  SPI_ERR_INVALID_ARG = 1,
  SPI_ERR_HIP = 2,
  SPI_ERR_TOO_LARGE = 3
} SpiStatus;

typedef enum SpiMotorModel {
  SPI_MOTOR_NONE = 0,
  SPI_MOTOR_SCALAR = 1,    /* act2tau_scalar,    active_sysid_openloop.py:356-363 */
  SPI_MOTOR_VEC3 = 2,      /* act2tau_vec3,      active_sysid_openloop.py:365-379 */
  SPI_MOTOR_VEC3_TANH = 3  /* act2tau_vec3_tanh, active_sysid_openloop.py:381-400 */
} SpiMotorModel;

typedef enum SpiCostMode {
  SPI_COST_TRACKING = 0,   /* sum over steps of weighted squared errors of all 6 groups */
  SPI_COST_REFERENCE = 1   /* scripts/eval.py:287-292: horizon end, L2 norms of pos, quat, q */
} SpiCostMode;

/* Simulation settings shared by every lane of a launch. */
typedef struct SpiSimArgs {
  float dt;            /* physics step: 1 / sim.fps (isaacgym_active_sysid.yaml:15) */
  int decimation;      /* physics steps per control step (isaacgym_active_sysid.yaml:16) */
  int substeps;        /* integration sub-intervals per physics step, torque held */
  float gravity;       /* 9.81 along -z (isaacgym.py:71) */
  float kp[12];        /* PD stiffness per DOF (go2.yaml:78-90) */
  float kd[12];        /* PD damping per DOF (go2.yaml:92-104) */
  float q0[12];        /* default joint angles (go2.yaml:56-68) */
  float effort[12];    /* torque clip (go2.yaml:39) */
  float action_scale;  /* go2.yaml:107 */
  float action_clip;   /* go2.yaml:108 */
  float hip_scale;     /* 0.5 for go2_omni (go2_omni.py:437), 1.0 for legged_base */
  int motor_model;     /* SpiMotorModel */
  float k_contact;     /* penalty stiffness [N/m] */
  float c_contact;     /* normal damping [N s/m] */
  float mu;            /* Coulomb cone */
  float c_tangent;     /* tangential viscous slope below the cone [N s/m] */
  float fric_vel_inv;  /* 1 / velocity scale of the tanh joint friction [s/rad] */
  float state_limit;   /* a lane whose |v|,|w|,|qd| sum exceeds this is marked diverged */
} SpiSimArgs;

typedef struct SpiCostArgs {
  int mode;            /* SpiCostMode */
  float w[6];          /* weights: pos, quat, linvel, angvel, q, qd (reference mode uses 0, 1, 4) */
  /* Recorded PD gains, device [n_seg][2] = (kp, kd) per segment, or NULL.  When set, a rollout
   * of segment s runs its PD law with kp = seg_gains[2s] and kd = seg_gains[2s+1] on all 12
   * joints in place of SpiSimArgs kp/kd, as scripts/eval.py:192-201 (apply_pd_gains_batch)
   * sets env.p_gains/d_gains from the recorded pd_gain_kp/kd of each chunk (:245-250). */
  const float* seg_gains;
} SpiCostArgs;

/* One problem of a batched CEM launch (device-resident array of these).  Every pointer is
 * device memory; `iteration` is advanced on the device by spi_batch_cem_update, so a captured
 * graph of one iteration can be replayed.  theta/mean/std/lo/hi/min_std/best cover all
 * SPI_P_COUNT parameters (fixed ones have std 0 and min_std 0). */
typedef struct SpiBatchProblem {
  const float* segments;        /* [n_seg][spi_segment_stride(horizon)] */
  const float* seg_weight;      /* [n_seg]: cost[c] = sum_s seg_weight[s] * cost_seg[s][c] */
  int n_seg;
  int horizon;
  SpiCostArgs cost;
  int n_samples;                /* <= SPI_MAX_CEM_SAMPLES */
  int n_elite;
  float alpha;
  unsigned int iteration;
  unsigned long long seed;
  float* mean;
  float* std;
  const float* lo;
  const float* hi;
  const float* min_std;
  float* theta;                 /* [SPI_P_COUNT][n_samples] */
  float* cost_seg;              /* [n_seg][n_samples] */
  float* cost_out;              /* [n_samples] */
  float* best;                  /* [SPI_P_COUNT] */
  float* best_cost;             /* [1] */
} SpiBatchProblem;

/* Command sequences of the active-exploration optimizer (active_sysid.py:259-400). */
#define SPI_MAX_COMMAND_DIMS 16
typedef enum SpiCommandMode {
  SPI_COMMAND_CONSTANT = 0,     /* _sample_constant,   active_sysid.py:295-314 */
  SPI_COMMAND_POLYNOMIAL = 1,   /* _sample_polynomial, active_sysid.py:316-361 */
  SPI_COMMAND_BEZIER = 2        /* _sample_bezier,     active_sysid.py:363-400 */
} SpiCommandMode;
typedef struct SpiCommandSpec {
  int mode;                     /* SpiCommandMode */
  int n_dims;                   /* command width (14), <= SPI_MAX_COMMAND_DIMS */
  int n_sampled;                /* sampled dimensions */
  int sampled_idx[SPI_MAX_COMMAND_DIMS];
  float default_cmd[SPI_MAX_COMMAND_DIMS];
  float lo[SPI_MAX_COMMAND_DIMS], hi[SPI_MAX_COMMAND_DIMS];  /* command_ranges per dimension */
  int n_updates;                /* rollout_length // horizon_length */
  int steps_per_update;         /* int(horizon_length / dt) */
  int poly_degree;
  int n_bezier;
} SpiCommandSpec;

/* Tangent directions of a sensitivity rollout: slot k seeds parameter row param[k] of the
 * SoA parameter array (0 <= param[k] < SPI_P_COUNT) with a unit tangent. */
typedef struct SpiTangentSpec {
  int n_tan;                 /* 1 .. SPI_MAX_TANGENTS */
  int param[SPI_P_COUNT];
} SpiTangentSpec;

const char* spi_last_error(void);
int spi_abi_version(void);

/* Number of floats of one segment record: init state (37), then per control step the
 * action (12) and the target state after that step (37). */
int spi_segment_stride(int horizon);

/* ---- BaseSimulator surface (replaces the IsaacGym calls behind spigym/simulator) ---- */

/* One physics step for n envs in place: replaces gym.set_dof_actuation_force_tensor +
 * gym.simulate + refresh_dof_state_tensor (isaacgym.py:598-599, 622-626).
 * torques_dev [n][12]; foot_force_dev [n][4][3] world-frame contact forces or NULL. */
int spi_physics_step(const SpiSimArgs* sim, const float* params_dev, int ld, int n,
                     float* root_dev, float* dof_state_dev, const float* torques_dev,
                     float* foot_force_dev, hipStream_t stream);

/* Forward dynamics without integration: acc_dev [n][18] = base spatial acceleration in base
 * coordinates [angular(3), linear(3)] (gravity included) followed by qdd(12), for the motor
 * torques torques_dev [n][12] (joint friction added inside).  The quantity PhysX computes inside
 * gym.simulate (isaacgym.py:623); exposed for sensitivities and tests. */
int spi_forward_dynamics(const SpiSimArgs* sim, const float* params_dev, int ld, int n, const float* root_dev,
                         const float* dof_state_dev, const float* torques_dev, float* acc_dev, float* foot_force_dev,
                         hipStream_t stream);

/* Row scatter dst[idx[i]] = src[idx[i]] of row_len floats: replaces
 * gym.set_actor_root_state_tensor_indexed (isaacgym.py:601-607, row_len 13) and
 * gym.set_dof_state_tensor_indexed (isaacgym.py:615-620, row_len 24). */
int spi_set_rows_indexed(float* dst_dev, const float* src_dev, const int* idx_dev, int n_idx, int n_rows,
                         int row_len, hipStream_t stream);

/* Rigid-body states [n][19][13] (pos, quat xyzw, linvel, angvel; world) of the IsaacGym
 * body list (go2.yaml:44): replaces gym.refresh_rigid_body_state_tensor (isaacgym.py:582). */
int spi_rigid_body_states(const float* root_dev, const float* dof_state_dev, int n, float* out_dev,
                          hipStream_t stream);

/* PD + clip + motor model for n envs: the torque computation of
 * legged_robot_base.py:529-560 / go2_omni.py:423-465 + active_sysid_openloop.py:174-187.
 * actions_dev [n][12] (raw, clipped inside), torques_dev [n][12]. */
int spi_compute_torques(const SpiSimArgs* sim, const float* params_dev, int ld, int n, const float* actions_dev,
                        const float* dof_state_dev, float* torques_dev, hipStream_t stream);

/* Finite-difference Fisher-information trace of active_sysid_openloop.py:402-426.
 * Envs are grouped (1 main + n_param aux); out_dev[n] gets the group's ||X||_F^2 repeated
 * for every env of the group (repeat_interleave). */
int spi_fd_fim_trace(const float* root_dev, const float* dof_state_dev, const float* env_origins_dev,
                     int n_main, int n_param, float delta, float* out_dev, hipStream_t stream);

/* ---- fused hot path ---- */

/* Fused rollout + tracking cost: for every (segment s, candidate c) run `horizon` control
 * steps of PD + motor model + ABA + contact + semi-implicit Euler from the segment's initial
 * state and write cost_seg_dev[s * n_cand + c].  Replaces the eager loop of
 * scripts/eval.py:239-296 (and the per-candidate Python loop of mass_landscape.py:123-126).
 * seg_dev: [n_seg][spi_segment_stride(horizon)] floats.  Diverged lanes get +inf. */
int spi_rollout_cost(const SpiSimArgs* sim, const SpiCostArgs* cost, const float* params_dev, int ld, int n_cand,
                     const float* seg_dev, int n_seg, int horizon, float* cost_seg_dev, hipStream_t stream);

/* The same costs with four lanes per (candidate, segment), one per leg (base replicated, the
 * legs' articulated inertias summed across the four lanes): lower latency for batches that
 * leave most SIMDs idle, more instructions in total.  Same layout; equal to spi_rollout_cost up
 * to float rounding (the order of the legs' sums differs). */
int spi_rollout_cost_legs(const SpiSimArgs* sim, const SpiCostArgs* cost, const float* params_dev, int ld,
                          int n_cand, const float* seg_dev, int n_seg, int horizon, float* cost_seg_dev,
                          hipStream_t stream);

/* cost_dev[c] = sum_s seg_weight_dev[s] * cost_seg_dev[s * n_cand + c], fixed summation order. */
int spi_reduce_segments(const float* cost_seg_dev, const float* seg_weight_dev, int n_seg, int n_cand,
                        float* cost_dev, hipStream_t stream);

/* Record mode (synthetic ground truth / replay): n lanes each with their own params,
 * initial state init_dev [n][37] and actions actions_dev [n][T][12]; writes the state after
 * every control step to states_dev [n][T][37]. */
int spi_simulate(const SpiSimArgs* sim, const float* params_dev, int ld, int n, const float* init_dev,
                 const float* actions_dev, int n_steps, float* states_dev, hipStream_t stream);

/* Active-sysID objective over open-loop excitations (active_sysid_openloop.py:116-131, 196-252,
 * 316-330, 402-426; active_sysid.py:527-600): n_group groups of (1 main + n_param aux) envs,
 * env g*(n_param+1)+k uses params column g*(n_param+1)+k (aux k = main + delta on one
 * parameter), all envs of a group replay actions_dev [n_group][T][12] from their own initial
 * state init_dev [n_env][37].  After every control step the group's FD-FIM trace
 * ||X||_F^2 (X = (main - aux) / delta over root(13) + q(12)) is accumulated;
 * fim_dev [n_group] receives the sum over the T steps.  The T steps are the command steps
 * that follow reset_all's zero-action step (base_task.py:90-100), which the reference counts
 * as step_idx 1: command step t (0-based) has step_idx t + 2, and after it the aux envs are
 * synchronised to the main env when step_idx % ksync == 1, or after every step for
 * ksync == 1 (active_sysid_openloop.py:246-252); with ksync 5 that is after command steps
 * 5, 10, 15 (1-based).  init_dev holds the states after the reset step (and its sync).
 * ksync_mode 0 = as the reference does it (only the first physics step's PD torque sees the
 * main's joint state, active_sysid_openloop.py:283-290 writes views that are not pushed),
 * 1 = full state copy.  ksync <= 0 disables.  n_param + 1 <= 64. */
int spi_fim_rollout(const SpiSimArgs* sim, const float* params_dev, int ld, int n_group, int n_param,
                    const float* init_dev, const float* actions_dev, int n_steps, float delta, int ksync,
                    int ksync_mode, float* fim_dev, hipStream_t stream);

/* ---- env reset (active_sysid_openloop.py:116-162, 274-292; legged_robot_base.py:761-784) ---- */

/* What a reset draws around: q = q_default * U[q_lo, q_hi] per joint, qd = 0 (_reset_dofs,
 * active_sysid_openloop.py:274-292); base at root0 = (pos xyz, quat xyzw) of the init state
 * (go2.yaml init_state, plus the env origin), linear and angular velocity U[-vel_range, vel_range]
 * each (_reset_root_states, legged_robot_base.py:761-784). */
typedef struct SpiResetSpec {
  float q_default[12];
  float root0[7];
  float q_lo, q_hi;       /* 0.5, 1.5 in the reference */
  float vel_range;        /* 0.5 in the reference */
} SpiResetSpec;

/* The reset states of n_env envs drawn on the device, out_dev [n_env][37] = root(13) + q(12) +
 * qd(12).  Uniforms from Philox4x32-10 under key `seed`, counter (env, block, 'RSET'): env e's
 * draw depends only on (seed, e).  origins_dev [n_env][3] (env origins added to the base
 * position) or NULL.  Replaces the torch_rand_float draws of reset_all (base_task.py:90-100). */
int spi_reset_states(const SpiResetSpec* spec, int n_env, const float* origins_dev, unsigned long long seed,
                     float* out_dev, hipStream_t stream);

/* ---- CEM (replaces the Optuna sampler of algo/active_sysid.yaml:20 / mass_opt.py:216) ---- */

/* theta_dev[p * ld + k] = clamp(mean[p] + std[p] * z(seed, iter, p * n + k), lo[p], hi[p]),
 * z standard normal from Philox4x32-10 + Box-Muller.  mean/std/lo/hi: [n_param] device. */
int spi_cem_sample(const float* mean_dev, const float* std_dev, const float* lo_dev, const float* hi_dev, int n_param,
                   int n, unsigned long long seed, unsigned int iteration, float* theta_dev, int ld,
                   hipStream_t stream);

/* Elite selection + update: sorts (cost, index) of n <= SPI_MAX_CEM_SAMPLES samples in LDS
 * (bitonic), takes the n_elite lowest, writes new mean/std (std floored at min_std_dev[p],
 * blended with the old moments by alpha: new = alpha*old + (1-alpha)*elite), the best sample
 * and its cost (best_dev [n_param], best_cost_dev [1]: kept if not improved), and
 * elite_idx_dev [n_elite] (may be NULL). */
int spi_cem_update(const float* cost_dev, const float* theta_dev, int ld, int n_param, int n, int n_elite,
                   float alpha, const float* min_std_dev, float* mean_dev, float* std_dev, float* best_dev,
                   float* best_cost_dev, int* elite_idx_dev, hipStream_t stream);

/* ---- batched problems (many independent sysID problems per launch) ---- */

/* Sample every problem's candidates for its current iteration (as spi_cem_sample). */
int spi_batch_cem_sample(const SpiBatchProblem* probs_dev, int n_prob, int max_samples, hipStream_t stream);

/* Fused rollout + cost for every (problem, segment, candidate block) task with a persistent
 * grid of n_blocks workgroups of `block` lanes (64, 128 or 256; unroll 1 or 2 of the physics
 * step loop) pulling tasks from queue_dev (2 x u64, zero before first use; the launch leaves
 * it zero).  task_prefix_dev[p] = number of tasks of problems 0..p-1, a task of problem p
 * being (segment, block of `block` candidates); list long horizons first for balance.
 * Replaces the per-problem Python loops of scripts/mass_opt.py:203-219. */
int spi_batch_rollout_cost(const SpiSimArgs* sim, const SpiBatchProblem* probs_dev, const long long* task_prefix_dev,
                           int n_prob, long long n_tasks, int max_horizon, int block, int unroll, int n_blocks,
                           unsigned long long* queue_dev, hipStream_t stream);

/* Per problem (one workgroup each): cost_out = weighted sum over segments (fixed order), then
 * the CEM update of spi_cem_update, then iteration += 1. */
int spi_batch_cem_update(const SpiBatchProblem* probs_dev, int n_prob, hipStream_t stream);

/* ---- differentiable rollout and Fisher information (active exploration) ---- */

/* Trajectory sensitivities by forward (tangent) propagation through the exact discrete
 * rollout.  Group g = e * n_samples + c rolls out parameter sample c (params_dev[k * ld + c])
 * under excitation e (actions_dev [n_exc][horizon][12]) from init_dev [37]; after every
 * control step t it records y_t = [root (13, IsaacGym layout) | q (12)] -- the state the
 * reference's FIM reward differentiates (active_sysid_openloop.py:402-426) -- and
 * J_dev[g][t * 25 + i][k] = d y_t[i] / d theta[tan->param[k]].  y_dev [G][horizon][25] may be
 * NULL.  status_dev [G]: 0 ok, 1 the state left the finite / state_limit range (its
 * remaining rows of J are zero).  Replaces the (P + 1)-env finite-difference groups of
 * active_sysid_openloop.py:29-80 with exact derivatives. */
int spi_traj_sensitivity(const SpiSimArgs* sim, const SpiTangentSpec* tan, const float* params_dev, int ld,
                         int n_samples, const float* init_dev, const float* actions_dev, int n_exc, int horizon,
                         float* J_dev, float* y_dev, int* status_dev, hipStream_t stream);
/* The same sensitivities with four lanes per (sample, excitation, tangent), one per leg: the
 * tangent state of a lane is one leg and the base, which fits the registers (far fewer
 * spills) at the price of replicating the base work four times.  Same layout and results up to
 * float rounding (the order of the legs' sums into the base differs). */
int spi_traj_sensitivity_legs(const SpiSimArgs* sim, const SpiTangentSpec* tan, const float* params_dev, int ld,
                              int n_samples, const float* init_dev, const float* actions_dev, int n_exc, int horizon,
                              float* J_dev, float* y_dev, int* status_dev, hipStream_t stream);
/* The same sensitivities with four lanes per (sample, excitation, pair of tangents): each lane
 * carries the value and two tangents (parameter slots j and j + ceil(n_tan / 2)), so one value
 * path serves two parameters; about a fifth fewer instructions per tangent than
 * spi_traj_sensitivity_legs.  Same layout and results up to float rounding. */
int spi_traj_sensitivity_pairs(const SpiSimArgs* sim, const SpiTangentSpec* tan, const float* params_dev, int ld,
                               int n_samples, const float* init_dev, const float* actions_dev, int n_exc, int horizon,
                               float* J_dev, float* y_dev, int* status_dev, hipStream_t stream);

/* Fisher information of every group, F_dev[g] = J_g^T J_g ([n_tan][n_tan], n_tan <= 32),
 * with J_g = J_dev[g] of n_rows x n_tan, accumulated with FP32 MFMA (32x32x2). */
int spi_fim_assemble(const float* J_dev, long long n_groups, int n_rows, int n_tan, float* F_dev,
                     hipStream_t stream);

/* Per set s of set_size consecutive groups (one excitation): mean of the groups' F (FP64,
 * fixed order, groups with status != 0 skipped), scaled to D F D with D = diag(scale_dev)
 * (scale_dev [n_tan] may be NULL: D = I; with D = parameter magnitudes F is the information
 * about relative parameter changes), plus reg * trace / n_tan * I, FP64 Cholesky, logdet.
 * logdet_dev [n_sets] (-inf when not positive definite or no valid group), n_valid_dev
 * [n_sets], F_mean_dev [n_sets][n_tan][n_tan] (the scaled, regularised mean; may be NULL). */
int spi_fim_logdet(const float* F_dev, const int* status_dev, int n_sets, int set_size, int n_tan,
                   const double* scale_dev, double reg, double* F_mean_dev, double* logdet_dev, int* n_valid_dev,
                   hipStream_t stream);

/* Tracking cost and its exact gradient for every (segment, candidate) pair (layouts as
 * spi_rollout_cost): cost_dev [n_seg][n_cand], grad_dev [n_seg][n_cand][n_tan] =
 * d cost / d theta[tan->param[k]] (+inf cost and zero gradient for diverged lanes). */
int spi_rollout_cost_grad(const SpiSimArgs* sim, const SpiCostArgs* cost, const SpiTangentSpec* tan,
                          const float* params_dev, int ld, int n_cand, const float* segments_dev, int n_seg,
                          int horizon, float* cost_dev, float* grad_dev, hipStream_t stream);

/* Weighted tracking residuals and their Jacobian, for Gauss-Newton refinement of identified
 * parameters.  Group g = c * n_seg + s; Jr_dev [G][horizon * 37][n_tan + 1]: row t * 37 + i is
 * [d r_i / d theta[tan->param[k]] (k < n_tan) | r_i] with r_i = sqrt(w_group(i) * seg_w[s]) *
 * (x_i - target_i) over the 37-float state after control step t (tracking cost = sum r^2, the
 * objective of scripts/eval.py:287-292 in squared form).  status_dev [G] as above.
 * n_tan + 1 <= 32; tracking cost mode only. */
int spi_residual_jacobian(const SpiSimArgs* sim, const SpiCostArgs* cost, const SpiTangentSpec* tan,
                          const float* params_dev, int ld, int n_cand, const float* segments_dev,
                          const float* seg_weight_dev, int n_seg, int horizon, float* Jr_dev, int* status_dev,
                          hipStream_t stream);
/* spi_residual_jacobian with four lanes per (group, parameter) tangent, one per leg: the same
 * matrix up to the order of the legs' sums, for refinement steps whose few hundred groups would
 * leave most SIMDs idle with one lane per tangent. */
int spi_residual_jacobian_legs(const SpiSimArgs* sim, const SpiCostArgs* cost, const SpiTangentSpec* tan,
                               const float* params_dev, int ld, int n_cand, const float* segments_dev,
                               const float* seg_weight_dev, int n_seg, int horizon, float* Jr_dev, int* status_dev,
                               hipStream_t stream);

/* Levenberg-Marquardt step per candidate from the augmented Gram matrices gram_dev
 * [n_cand * n_seg][n_tan + 1][n_tan + 1] of spi_residual_jacobian's groups (spi_fim_assemble):
 * FP64 sums over segments, (A + lambda diag(A)) delta = -J^T r by FP64 Cholesky.  delta_dev
 * [n_cand][n_tan], cost_dev [n_cand] = r^T r (+inf if a group diverged), pred_dev [n_cand] = the
 * linear model's cost after the step.  Replaces the Optuna TPE refinement of mass_opt.py:203-219. */
int spi_gn_step(const float* gram_dev, const int* status_dev, int n_cand, int n_seg, int n_tan, const double* lambda_dev,
                double* delta_dev, double* cost_dev, double* pred_dev, hipStream_t stream);

/* Command sequences of n_trials trials from their parameters, params_dev [n_trials][n_params]
 * in the order the reference suggests them: constant [n_sampled][n_updates], polynomial
 * [n_updates][n_sampled][poly_degree + 1] (coefficients, tanh-squashed and mapped to the
 * range), bezier [n_updates][n_sampled][n_bezier] (control points, Bernstein basis, clipped).
 * out_dev [n_trials][n_updates * steps_per_update][n_dims]; unsampled dimensions hold
 * default_cmd.  Replaces the per-trial Python loops of active_sysid.py:295-400. */
int spi_command_sequences(const SpiCommandSpec* spec, const float* params_dev, int n_trials, int n_params,
                          float* out_dev, hipStream_t stream);

/* Sliding windows of one recorded trajectory packed as segment records (layout of
 * spi_rollout_cost): window w starts at t = first + w; init = states[t], then per step k
 * actions[t + k] and target states[t + k + 1].  states_dev [n_steps][37], actions_dev
 * [n_steps][12]; first + n_windows + horizon <= n_steps.  The device side of
 * scripts/eval.py:101-171 (one window per start index, stride 1). */
int spi_pack_windows(const float* states_dev, const float* actions_dev, int n_steps, int horizon, int first,
                     int n_windows, float* segments_dev, hipStream_t stream);
/* spi_pack_windows over many trajectories held back to back in one buffer, in one launch:
 * window w starts at row rows_dev[w] of states_dev [n_rows][37] / actions_dev [n_rows][12]
 * (the caller's table keeps each window inside its own trajectory).  A window whose rows
 * r .. r + horizon do not all lie below n_rows is written as NaN; nothing outside the buffers
 * is read.  The device side of the loop over recorded files in scripts/eval.py:101-171. */
int spi_pack_windows_rows(const float* states_dev, const float* actions_dev, int n_rows, const int* rows_dev,
                          int n_windows, int horizon, float* segments_dev, hipStream_t stream);

/* Tracking cost and its exact gradient with respect to all 25 parameter rows for every
 * (segment, candidate) pair, by reverse mode (adjoint) with checkpoints every
 * checkpoint_every control steps: cost_dev [n_seg][n_cand] (layout of spi_rollout_cost),
 * grad_dev [n_seg][n_cand][25] (+inf cost and zero gradient for diverged lanes).  work_dev:
 * at least spi_adjoint_work_floats() floats of scratch, owned by the caller.  Same derivative
 * as spi_rollout_cost_grad at about a tenth of its cost for 22 parameters; decimation and
 * substeps <= 8. */
long long spi_adjoint_work_floats(int n_cand, int n_seg, int horizon, int checkpoint_every);
int spi_rollout_cost_adjoint(const SpiSimArgs* sim, const SpiCostArgs* cost, const float* params_dev, int ld,
                             int n_cand, const float* segments_dev, int n_seg, int horizon, int checkpoint_every,
                             float* work_dev, long long work_floats, float* cost_dev, float* grad_dev,
                             hipStream_t stream);
/* spi_rollout_cost_adjoint with four lanes per rollout (one per leg; base quantities computed
 * in all four, the legs' contributions summed across them): the same outputs up to the order
 * of the legs' sums, a quarter of the registers per lane.  work_dev: at least
 * spi_adjoint_legs_work_floats() floats (block checkpoints and the physics-step states of one
 * block; decimation = sim->decimation). */
long long spi_adjoint_legs_work_floats(int n_cand, int n_seg, int horizon, int checkpoint_every, int decimation);
int spi_rollout_cost_adjoint_legs(const SpiSimArgs* sim, const SpiCostArgs* cost, const float* params_dev, int ld,
                                  int n_cand, const float* segments_dev, int n_seg, int horizon, int checkpoint_every,
                                  float* work_dev, long long work_floats, float* cost_dev, float* grad_dev,
                                  hipStream_t stream);

/* Excitations as Fourier series: actions_dev[e][t][j] = scale * sum_k coeff_dev[e][j][k] *
 * sin(2 pi (k + 1) t dt_ctrl / period), k < n_terms. */
int spi_fourier_excitations(const float* coeff_dev, int n_exc, int n_terms, int horizon, float dt_ctrl, float period,
                            float scale, float* actions_dev, hipStream_t stream);

/* ---- Policy-in-the-loop active exploration (active_sysid.py:527-600 evaluate_policy over
 * active_sysid_openloop.py).  Envs are numbered group-major: env g * group + k, k = 0 the main
 * env; states [n_env][37] in the root layout; actions [n_env][12]; int flags [n_env]. */

/* Observation width of the go2_omni actor: 60 current terms + 14 history steps of each. */
int spi_policy_obs_dim(void);

/* One control step (LeggedRobotBase.step, legged_robot_base.py:169-209, with the motor model of
 * active_sysid_openloop.py:174-187): envs with done_dev set get zero actions
 * (active_sysid.py:556-558); applied_dev receives the zeroed, clipped actions; with sync_first
 * the first torque of each aux env uses its main env's joint state (the reference's view-only
 * sync, active_sysid_openloop.py:316-330); bad_dev[i] is set when env i diverges.
 * states_in_dev and states_out_dev must not alias. */
int spi_policy_step(const SpiSimArgs* sim, const float* params_dev, int ld, int n_env, int group,
                    const float* states_in_dev, float* states_out_dev, const float* actions_dev, const int* done_dev,
                    int sync_first, float* applied_dev, int* bad_dev, hipStream_t stream);

/* The same step with four lanes per env, one per leg (the base replicated, the legs' articulated
 * inertias summed across the four lanes): about a third of the instructions per lane on the
 * critical path, for batches too small to fill the GPU with one env per lane.  Same arguments
 * and results up to float rounding (the order of the legs' sums differs). */
int spi_policy_step_legs(const SpiSimArgs* sim, const float* params_dev, int ld, int n_env, int group,
                         const float* states_in_dev, float* states_out_dev, const float* actions_dev,
                         const int* done_dev, int sync_first, float* applied_dev, int* bad_dev, hipStream_t stream);

/* Per group: gravity termination OR-ed over the group into done_dev of every member
 * (legged_robot_base.py:336-339, active_sysid_openloop.py:259-272); acc_dev[g] += termination_rew
 * for a terminated group, else ||X||_F^2 of the finite-difference states
 * (active_sysid_openloop.py:402-426); group_bad_dev[g] set on divergence. */
int spi_policy_reward(const float* states_dev, int n_group, int group, float delta, float grav_x, float grav_y,
                      float termination_rew, int* done_dev, const int* bad_dev, float* acc_dev, int* group_bad_dev,
                      hipStream_t stream);

/* Gait clock (go2_omni.py:348-379) with commands step_now, observation (legged_robot_base.py:
 * 503-527, 795-844; go2_omni.py:627-668) with commands step_next; commands_dev
 * [n_group][n_cmd_steps][14]; gait_dev [n_env]; history_dev [n_env][14][60] newest first;
 * obs_dev [n_env][spi_policy_obs_dim()]. */
int spi_policy_obs(int n_env, int group, const float* states_dev, const float* commands_dev, int n_cmd_steps,
                   int step_now, int step_next, float dt_ctrl, const float* q_default_dev, float* gait_dev,
                   const float* applied_dev, float* history_dev, float* obs_dev, int evaluating, float height_lo,
                   float height_hi, float obs_clip, hipStream_t stream);

/* Ring form of the same observation, without the history shift: this step's 60 terms go to
 * slot `head` of ring_dev [spi_policy_ring_slots()][60][ld] (env index fastest; ld >= n_env);
 * the actor's first layer gathers the 900-wide row from the ring (spi_mlp_forward_ring with the
 * table of spi_policy_obs_kmap).  Slot head = step % spi_policy_ring_slots(), ring zeroed at reset. */
int spi_policy_terms(int n_env, int group, const float* states_dev, const float* commands_dev, int n_cmd_steps,
                     int step_now, int step_next, float dt_ctrl, const float* q_default_dev, float* gait_dev,
                     const float* applied_dev, float* ring_dev, int ld, int head, int evaluating, float height_lo,
                     float height_hi, hipStream_t stream);
/* spi_policy_reward and spi_policy_terms in one launch for the ring layout (both read only the
 * post-step state; the arguments of the two): per control step, one launch fewer. */
int spi_policy_reward_terms(int n_env, int group, const float* states_dev, const float* commands_dev,
                            int n_cmd_steps, int step_now, int step_next, float dt_ctrl, const float* q_default_dev,
                            float* gait_dev, const float* applied_dev, float* ring_dev, int ld, int head,
                            int evaluating, float height_lo, float height_hi, float delta, float grav_x, float grav_y,
                            float termination_rew, int* done_dev, const int* bad_dev, float* acc_dev,
                            int* group_bad_dev, hipStream_t stream);
int spi_policy_ring_slots(void);
/* Host table kmap_host[900][2] = (slots back from head, term index) of every observation column. */
int spi_policy_obs_kmap(int* kmap_host);

/* Full-state copy of each main env into its aux envs (a corrected sync; the reference's reaches
 * only the next first torque, see spi_policy_step). */
int spi_policy_sync_state(float* states_dev, int n_env, int group, hipStream_t stream);

/* Actor MLP (modules.py:47-63, ppo_modules.py:73 act_inference): n_layers Linear layers of widths
 * dims[0] -> ... -> dims[n_layers], ELU between them.  params_dev packs, per layer, W [out][in]
 * (torch Linear layout) then b [out]; work_dev holds spi_mlp_work_floats() floats. */
long long spi_mlp_param_floats(int n_layers, const int* dims);
long long spi_mlp_work_floats(int M, int n_layers, const int* dims);
int spi_mlp_forward(const float* x_dev, int M, int n_layers, const int* dims, const float* params_dev, float* work_dev,
                    long long work_floats, float* y_dev, hipStream_t stream);
/* spi_mlp_forward without a caller workspace and without any allocation: no counterpart in
 * the reference; for callers that cannot allocate (code under HIP graph capture, serving loops,
 * harnesses that forbid hipMalloc).  The hidden activations live in a library-owned device
 * array of SPI_MLP_SCRATCH_FLOATS floats per GPU (held until the process ends), so
 * 2 * M * (widest hidden layer) must not exceed it (SPI_ERR_TOO_LARGE otherwise; 4096 rows of
 * the 512-wide actor).  Every call on one GPU shares that array: calls that overlap on one GPU
 * (other streams, other host threads) race.  Same results as spi_mlp_forward, bit for bit. */
#define SPI_MLP_SCRATCH_FLOATS (1 << 22)
int spi_mlp_forward_noalloc(const float* x_dev, int M, int n_layers, const int* dims, const float* params_dev,
                            float* y_dev, hipStream_t stream);
/* The same network with its input row m gathered from a ring (see spi_policy_terms):
 * x[m][k] = clip(ring_dev[((head - kmap[k][0]) mod n_slots) * n_terms + kmap[k][1]][m], +-clip),
 * kmap_dev [dims[0]][2] int (8-byte aligned); needs dims[0] % 4 == 0. */
int spi_mlp_forward_ring(const float* ring_dev, int ld, int n_slots, int n_terms, int head, const int* kmap_dev,
                         float clip, int M, int n_layers, const int* dims, const float* params_dev, float* work_dev,
                         long long work_floats, float* y_dev, hipStream_t stream);

#ifdef __cplusplus
}
#endif
#endif /* SPI_HIP_H */
