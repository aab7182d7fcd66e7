// Cross-entropy method on device: Gaussian sampling (Philox4x32-10 + Box-Muller), elite
// selection by an LDS bitonic sort, and the elite-moment update.  Replaces the Optuna
// samplers of the reference (algo/active_sysid.yaml:20 CmaEsSampler, mass_opt.py:216
// TPESampler) behind the same ask -> evaluate -> tell loop (active_sysid.py:196-217).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>

#include "spi_types.h"

namespace {

int cem_fail(int code, const char* msg) {
  spi_set_error(msg);
  return code;
}

// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11): ten rounds of two 32x32->64-bit
// multiplies, key bumped by the Weyl constants after every round.
__device__ __forceinline__ void philox4x32_10(uint32_t c[4], uint32_t k0, uint32_t k1) {
  const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u, W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    uint32_t hi0 = __umulhi(M0, c[0]), lo0 = M0 * c[0];
    uint32_t hi1 = __umulhi(M1, c[2]), lo1 = M1 * c[2];
    uint32_t n0 = hi1 ^ c[1] ^ k0, n2 = hi0 ^ c[3] ^ k1;
    c[0] = n0; c[1] = lo1; c[2] = n2; c[3] = lo0;
    k0 += W0; k1 += W1;
  }
}

__device__ __forceinline__ float u01_open(uint32_t x) {
  return (static_cast<float>(x) + 0.5f) * 2.3283064365386963e-10f;
}

// Block b of stream (seed, iteration) -> four standard normals (two Box-Muller pairs).
__device__ __forceinline__ void philox_normals(unsigned long long seed, uint32_t iteration, unsigned long long blk,
                                               float z[4]) {
  uint32_t c[4] = {static_cast<uint32_t>(blk), static_cast<uint32_t>(blk >> 32), iteration, 0x53504941u};
  philox4x32_10(c, static_cast<uint32_t>(seed), static_cast<uint32_t>(seed >> 32));
#pragma unroll
  for (int p = 0; p < 2; ++p) {
    float u1 = u01_open(c[2 * p]), u2 = u01_open(c[2 * p + 1]);
    float rad = sqrtf(-2.0f * logf(u1));
    float s, co;
    sincosf(6.283185307179586f * u2, &s, &co);
    z[2 * p] = rad * co;
    z[2 * p + 1] = rad * s;
  }
}

__global__ void cem_sample_kernel(const float* __restrict__ mean, const float* __restrict__ stdv,
                                  const float* __restrict__ lo, const float* __restrict__ hi, int n_param, int n,
                                  unsigned long long seed, uint32_t iteration, float* __restrict__ theta, int ld) {
  const unsigned long long blk = static_cast<unsigned long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  const unsigned long long total = static_cast<unsigned long long>(n_param) * n;
  const unsigned long long i0 = blk * 4;
  if (i0 >= total) return;
  float z[4];
  philox_normals(seed, iteration, blk, z);
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    unsigned long long idx = i0 + e;
    if (idx >= total) break;
    int p = static_cast<int>(idx / n), k = static_cast<int>(idx % n);
    float v = mean[p] + stdv[p] * z[e];
    v = fminf(fmaxf(v, lo[p]), hi[p]);
    theta[static_cast<size_t>(p) * ld + k] = v;
  }
}

// cost_dev[c] = sum_s w[s] * cost_seg[s][c]: one thread per candidate, segments in a fixed
// order, so the per-candidate cost (and the CEM ranking) is deterministic.  Built with strict
// IEEE semantics so +inf (diverged) costs propagate; a segment of weight 0 is skipped (its
// cost may be +inf).  The loads of eight segments are issued together (they are independent;
// one at a time, a 256-candidate landscape waited on 64 dependent load latencies), the sum
// keeps the segment order.
__global__ void reduce_segments_kernel(const float* __restrict__ cs, const float* __restrict__ w, int n_seg,
                                       int n_cand, float* __restrict__ out) {
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n_cand) return;
  float acc = 0.0f;
  int s = 0;
  for (; s + 8 <= n_seg; s += 8) {
    float ws[8], v[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      ws[k] = w[s + k];
      v[k] = cs[static_cast<size_t>(s + k) * n_cand + c];
    }
#pragma unroll
    for (int k = 0; k < 8; ++k)
      if (ws[k] != 0.0f) acc += ws[k] * v[k];
  }
  for (; s < n_seg; ++s) {
    float ws = w[s];
    if (ws != 0.0f) acc += ws * cs[static_cast<size_t>(s) * n_cand + c];
  }
  out[c] = acc;
}


constexpr int kCemThreads = 1024;

// (cost, index) pairs sorted ascending in LDS (bitonic network, ties broken by index, NaN
// treated as +inf), then per-parameter elite moments with wave reductions and the best-so-far.
// Called by one whole workgroup of kCemThreads threads; keys must already hold the costs.
__device__ void cem_sort_update(float* key, int* val, int n, int npow2, const float* __restrict__ theta, int ld,
                                int n_param, int n_elite, float alpha, const float* __restrict__ min_std,
                                float* __restrict__ mean, float* __restrict__ stdv, float* __restrict__ best,
                                float* __restrict__ best_cost, int* __restrict__ elite_idx) {
  const int tid = threadIdx.x;
  for (int i = tid; i < npow2; i += kCemThreads) {
    float c = i < n ? key[i] : INFINITY;
    key[i] = (c == c) ? c : INFINITY;
    val[i] = i;
  }
  __syncthreads();
  for (int k = 2; k <= npow2; k <<= 1) {
    for (int j = k >> 1; j > 0; j >>= 1) {
      for (int i = tid; i < npow2; i += kCemThreads) {
        int p = i ^ j;
        if (p > i) {
          bool up = (i & k) == 0;
          float ki = key[i], kp = key[p];
          int vi = val[i], vp = val[p];
          bool gt = (ki > kp) || (ki == kp && vi > vp);
          if (gt == up) { key[i] = kp; key[p] = ki; val[i] = vp; val[p] = vi; }
        }
      }
      __syncthreads();
    }
  }
  if (elite_idx)
    for (int e = tid; e < n_elite; e += kCemThreads) elite_idx[e] = val[e];
  // one wave per parameter: lanes stride over the elites, then a fixed butterfly
  const int wave = tid / 64, lane = tid % 64, n_waves = kCemThreads / 64;
  for (int p = wave; p < n_param; p += n_waves) {
    const float* th = theta + static_cast<size_t>(p) * ld;
    float s = 0.0f;
    for (int e = lane; e < n_elite; e += 64) s += th[val[e]];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    const float m = s / static_cast<float>(n_elite);
    float v = 0.0f;
    for (int e = lane; e < n_elite; e += 64) {
      float d = th[val[e]] - m;
      v += d * d;
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    if (lane == 0) {
      float sd = sqrtf(v / static_cast<float>(n_elite));
      float nm = alpha * mean[p] + (1.0f - alpha) * m;
      float ns = alpha * stdv[p] + (1.0f - alpha) * sd;
      mean[p] = nm;
      stdv[p] = fmaxf(ns, min_std[p]);
    }
  }
  // best sample so far: replaced only when this iteration's best strictly improves it
  __shared__ int improved;
  if (tid == 0) improved = key[0] < best_cost[0];
  __syncthreads();
  if (improved) {
    for (int p = tid; p < n_param; p += kCemThreads) best[p] = theta[static_cast<size_t>(p) * ld + val[0]];
    if (tid == 0) best_cost[0] = key[0];
  }
  __syncthreads();
}

__global__ __launch_bounds__(kCemThreads) void cem_update_kernel(
    const float* __restrict__ cost, const float* __restrict__ theta, int ld, int n_param, int n, int n_elite,
    float alpha, const float* __restrict__ min_std, float* __restrict__ mean, float* __restrict__ stdv,
    float* __restrict__ best, float* __restrict__ best_cost, int* __restrict__ elite_idx, int npow2) {
  __shared__ float key[SPI_MAX_CEM_SAMPLES];
  __shared__ int val[SPI_MAX_CEM_SAMPLES];
  for (int i = threadIdx.x; i < n; i += kCemThreads) key[i] = cost[i];
  __syncthreads();
  cem_sort_update(key, val, n, npow2, theta, ld, n_param, n_elite, alpha, min_std, mean, stdv, best, best_cost,
                  elite_idx);
}

// Batched: blockIdx.y = problem; four normals per thread as in cem_sample_kernel.
__global__ void cem_sample_batch_kernel(const SpiBatchProblem* __restrict__ probs) {
  const SpiBatchProblem& P = probs[blockIdx.y];
  const unsigned long long blk = static_cast<unsigned long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  const unsigned long long total = static_cast<unsigned long long>(SPI_P_COUNT) * P.n_samples;
  const unsigned long long i0 = blk * 4;
  if (i0 >= total) return;
  float z[4];
  philox_normals(P.seed, P.iteration, blk, z);
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    unsigned long long idx = i0 + e;
    if (idx >= total) break;
    int p = static_cast<int>(idx / P.n_samples), k = static_cast<int>(idx % P.n_samples);
    float v = P.mean[p] + P.std[p] * z[e];
    v = fminf(fmaxf(v, P.lo[p]), P.hi[p]);
    P.theta[static_cast<size_t>(p) * P.n_samples + k] = v;
  }
}

// Batched: one workgroup per problem; segment reduction (fixed order), sort, moments, best,
// iteration + 1.
__global__ __launch_bounds__(kCemThreads) void cem_update_batch_kernel(SpiBatchProblem* __restrict__ probs) {
  __shared__ float key[SPI_MAX_CEM_SAMPLES];
  __shared__ int val[SPI_MAX_CEM_SAMPLES];
  SpiBatchProblem& P = probs[blockIdx.x];
  const int n = P.n_samples;
  for (int c = threadIdx.x; c < n; c += kCemThreads) {
    float acc = 0.0f;
    for (int s = 0; s < P.n_seg; ++s) {
      float ws = P.seg_weight[s];
      if (ws != 0.0f) acc += ws * P.cost_seg[static_cast<size_t>(s) * n + c];
    }
    P.cost_out[c] = acc;
    key[c] = acc;
  }
  __syncthreads();
  int npow2 = 2;
  while (npow2 < n) npow2 <<= 1;
  cem_sort_update(key, val, n, npow2, P.theta, n, SPI_P_COUNT, P.n_elite, P.alpha, P.min_std, P.mean, P.std,
                  P.best, P.best_cost, nullptr);
  if (threadIdx.x == 0) P.iteration = P.iteration + 1u;
}

// Reset draw of one env: uniforms u_k, k = 0..17, word k % 4 of Philox block k / 4 of the
// counter (env, block, 'RSET') under key seed.  q_j = q_default_j * (q_lo + (q_hi - q_lo) u_j),
// qd = 0; base at root0's position (+ the env's origin) and orientation; linear then angular
// velocity -v + 2 v u_{12+i}.  An env's draw depends only on (seed, env).
__global__ void reset_states_kernel(SpiResetSpec spec, int n_env, const float* __restrict__ origins,
                                    unsigned long long seed, float* __restrict__ out) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n_env) return;
  float u[20];
#pragma unroll
  for (int b = 0; b < 5; ++b) {
    uint32_t c[4] = {static_cast<uint32_t>(e), 0u, static_cast<uint32_t>(b), 0x52534554u};
    philox4x32_10(c, static_cast<uint32_t>(seed), static_cast<uint32_t>(seed >> 32));
#pragma unroll
    for (int w = 0; w < 4; ++w) u[4 * b + w] = u01_open(c[w]);
  }
  float* o = out + static_cast<size_t>(e) * SPI_STATE_DIM;
#pragma unroll
  for (int i = 0; i < 7; ++i) o[i] = spec.root0[i];
  if (origins) {
#pragma unroll
    for (int i = 0; i < 3; ++i) o[i] += origins[3 * static_cast<size_t>(e) + i];
  }
  const float v = spec.vel_range;
#pragma unroll
  for (int i = 0; i < 6; ++i) o[7 + i] = -v + 2.0f * v * u[12 + i];
#pragma unroll
  for (int j = 0; j < 12; ++j) {
    o[13 + j] = spec.q_default[j] * (spec.q_lo + (spec.q_hi - spec.q_lo) * u[j]);
    o[25 + j] = 0.0f;
  }
}

}  // namespace

extern "C" {

int spi_reset_states(const SpiResetSpec* spec, int n_env, const float* origins_dev, unsigned long long seed,
                     float* out_dev, hipStream_t stream) {
  if (!spec || n_env < 0 || !out_dev || !(spec->q_lo <= spec->q_hi) || !(spec->vel_range >= 0.0f))
    return cem_fail(SPI_ERR_INVALID_ARG, "spi_reset_states: bad arguments");
  if (n_env == 0) return SPI_OK;
  reset_states_kernel<<<(n_env + 255) / 256, 256, 0, stream>>>(*spec, n_env, origins_dev, seed, out_dev);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return cem_fail(SPI_ERR_HIP, hipGetErrorString(e));
  return SPI_OK;
}

int spi_cem_sample(const float* mean_dev, const float* std_dev, const float* lo_dev, const float* hi_dev, int n_param,
                   int n, unsigned long long seed, unsigned int iteration, float* theta_dev, int ld,
                   hipStream_t stream) {
  if (n_param < 1 || n < 0 || ld < n || !mean_dev || !std_dev || !lo_dev || !hi_dev || !theta_dev)
    return cem_fail(SPI_ERR_INVALID_ARG, "spi_cem_sample: bad arguments");
  if (n == 0) return SPI_OK;
  unsigned long long total = static_cast<unsigned long long>(n_param) * n;
  unsigned long long blocks = (total + 3) / 4;
  unsigned long long grid = (blocks + 255) / 256;
  if (grid > 0x7fffffffULL) return cem_fail(SPI_ERR_TOO_LARGE, "spi_cem_sample: too many samples");
  cem_sample_kernel<<<static_cast<unsigned>(grid), 256, 0, stream>>>(mean_dev, std_dev, lo_dev, hi_dev, n_param, n,
                                                                     seed, iteration, theta_dev, ld);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return cem_fail(SPI_ERR_HIP, hipGetErrorString(e));
  return SPI_OK;
}

int spi_cem_update(const float* cost_dev, const float* theta_dev, int ld, int n_param, int n, int n_elite,
                   float alpha, const float* min_std_dev, float* mean_dev, float* std_dev, float* best_dev,
                   float* best_cost_dev, int* elite_idx_dev, hipStream_t stream) {
  if (n < 1 || n > SPI_MAX_CEM_SAMPLES) return cem_fail(SPI_ERR_TOO_LARGE, "spi_cem_update: n out of [1, 8192]");
  if (n_param < 1 || n_elite < 1 || n_elite > n || ld < n || !(alpha >= 0.0f && alpha < 1.0f) || !cost_dev ||
      !theta_dev || !min_std_dev || !mean_dev || !std_dev || !best_dev || !best_cost_dev)
    return cem_fail(SPI_ERR_INVALID_ARG, "spi_cem_update: bad arguments");
  int npow2 = 1;
  while (npow2 < n) npow2 <<= 1;
  if (npow2 < 2) npow2 = 2;
  cem_update_kernel<<<1, kCemThreads, 0, stream>>>(cost_dev, theta_dev, ld, n_param, n, n_elite, alpha, min_std_dev,
                                                   mean_dev, std_dev, best_dev, best_cost_dev, elite_idx_dev, npow2);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return cem_fail(SPI_ERR_HIP, hipGetErrorString(e));
  return SPI_OK;
}

int spi_reduce_segments(const float* cost_seg_dev, const float* seg_weight_dev, int n_seg, int n_cand,
                        float* cost_dev, hipStream_t stream) {
  if (n_seg < 0 || n_cand < 0 || !cost_seg_dev || !seg_weight_dev || !cost_dev)
    return cem_fail(SPI_ERR_INVALID_ARG, "spi_reduce_segments: bad arguments");
  if (n_cand == 0) return SPI_OK;
  reduce_segments_kernel<<<(n_cand + 255) / 256, 256, 0, stream>>>(cost_seg_dev, seg_weight_dev, n_seg, n_cand,
                                                                   cost_dev);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return cem_fail(SPI_ERR_HIP, hipGetErrorString(e));
  return SPI_OK;
}

int spi_batch_cem_sample(const SpiBatchProblem* probs_dev, int n_prob, int max_samples, hipStream_t stream) {
  if (n_prob < 1 || max_samples < 1 || max_samples > SPI_MAX_CEM_SAMPLES || !probs_dev)
    return cem_fail(SPI_ERR_INVALID_ARG, "spi_batch_cem_sample: bad arguments");
  if (n_prob > 65535) return cem_fail(SPI_ERR_TOO_LARGE, "spi_batch_cem_sample: too many problems");
  unsigned long long blocks = (static_cast<unsigned long long>(SPI_P_COUNT) * max_samples + 3) / 4;
  dim3 grid(static_cast<unsigned>((blocks + 255) / 256), static_cast<unsigned>(n_prob));
  cem_sample_batch_kernel<<<grid, 256, 0, stream>>>(probs_dev);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return cem_fail(SPI_ERR_HIP, hipGetErrorString(e));
  return SPI_OK;
}

int spi_batch_cem_update(const SpiBatchProblem* probs_dev, int n_prob, hipStream_t stream) {
  if (n_prob < 1 || !probs_dev) return cem_fail(SPI_ERR_INVALID_ARG, "spi_batch_cem_update: bad arguments");
  cem_update_batch_kernel<<<n_prob, kCemThreads, 0, stream>>>(const_cast<SpiBatchProblem*>(probs_dev));
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return cem_fail(SPI_ERR_HIP, hipGetErrorString(e));
  return SPI_OK;
}

}  // extern "C"
