// Fisher information of trajectory sensitivities and its log-determinant (gfx950).
//
// * Assembly: F = J^T J per (parameter sample, excitation) group with the 32x32x2 FP32 matrix
//   instruction, one wavefront per group -- the only matrix-core work of the engine.
// * logdet: FP64 mean over the parameter samples of one excitation, FP64 Cholesky.
// * Fourier excitations: the candidate command sequences scored by the objective.
// Strict IEEE semantics (no finite-math): -inf marks a matrix that is not positive definite.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstdio>

#include "spi_types.h"

namespace {

int ffail(int code, const char* msg) {
  spi_set_error(msg);
  return code;
}
int flaunch(const char* what) {
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) return SPI_OK;
  char buf[256];
  std::snprintf(buf, sizeof(buf), "%s: %s", what, hipGetErrorString(e));
  spi_set_error(buf);
  return SPI_ERR_HIP;
}

// F_g = J_g^T J_g: one wavefront per group.  Lane l supplies A[l % 32][l / 32] = B[l / 32][l % 32]
// = J[k0 + l / 32][l % 32] (zero beyond n_tan), so each instruction adds the outer products
// of two rows of J.
typedef float floatx16 __attribute__((ext_vector_type(16)));
__global__ __launch_bounds__(256) void fim_assemble_kernel(const float* __restrict__ J, long long G, int K, int P,
                                                           float* __restrict__ F) {
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const long long g = static_cast<long long>(blockIdx.x) * 4 + wave;
  if (g >= G) return;   // uniform per wavefront
  const float* Jg = J + static_cast<size_t>(g) * K * P;
  const int col = lane & 31, half = lane >> 5;
  const bool live = col < P;
  floatx16 acc;
#pragma unroll
  for (int r = 0; r < 16; ++r) acc[r] = 0.0f;
  int k0 = 0;
  for (; k0 + 8 <= K; k0 += 8) {
    float v[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) v[u] = live ? Jg[static_cast<size_t>(k0 + 2 * u + half) * P + col] : 0.0f;
#pragma unroll
    for (int u = 0; u < 4; ++u) acc = __builtin_amdgcn_mfma_f32_32x32x2f32(v[u], v[u], acc, 0, 0, 0);
  }
  for (; k0 < K; k0 += 2) {
    const int k = k0 + half;
    const float v = (live && k < K) ? Jg[static_cast<size_t>(k) * P + col] : 0.0f;
    acc = __builtin_amdgcn_mfma_f32_32x32x2f32(v, v, acc, 0, 0, 0);
  }
  // result layout: D[i][j], j = lane % 32, i = 8 (r / 4) + 4 (lane / 32) + r % 4
  float* Fg = F + static_cast<size_t>(g) * P * P;
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    const int i = 8 * (r >> 2) + 4 * half + (r & 3);
    if (i < P && col < P) Fg[i * P + col] = acc[r];
  }
}

// fim_assemble_kernel with the group's J streamed through LDS: a wavefront reads rows
// [64 c, 64 c + 64) of its group, a contiguous run of 64 P floats, with coalesced float4 loads
// (the next chunk in flight while the MFMAs consume this one), stores it to its LDS tile and
// feeds the MFMAs from there.  Same row pairs in the same order as fim_assemble_kernel: the same
// result.  Needs K * P % 4 == 0 and 16-byte aligned J; waves past the last group stream the last
// group and store nothing, so every wave reaches the barriers.
constexpr int kFimChunk = 64;
constexpr int kFimPer = kFimChunk * SPI_MAX_TANGENTS / 4 / 64;   // float4 per lane per chunk (P <= 32)
__global__ __launch_bounds__(256) void fim_assemble_vec_kernel(const float* __restrict__ J, long long G, int K, int P,
                                                               float* __restrict__ F) {
  __shared__ float4 tile[4][kFimChunk * SPI_MAX_TANGENTS / 4];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const long long g = static_cast<long long>(blockIdx.x) * 4 + wave;
  const long long gs = g < G ? g : G - 1;
  const float4* Jg = reinterpret_cast<const float4*>(J + static_cast<size_t>(gs) * K * P);
  const int n4 = K * P / 4;                 // float4 of the group
  const int c4 = kFimChunk * P / 4;         // float4 of a full chunk
  float4* T4 = tile[wave];
  const float* T = reinterpret_cast<const float*>(T4);
  const int col = lane & 31, half = lane >> 5;
  const bool live = col < P;
  floatx16 acc;
#pragma unroll
  for (int r = 0; r < 16; ++r) acc[r] = 0.0f;
  float4 reg[kFimPer];
  auto fetch = [&](int c) {
#pragma unroll
    for (int i = 0; i < kFimPer; ++i) {
      const int e = i * 64 + lane, idx = c * c4 + e;
      reg[i] = (e < c4 && idx < n4) ? Jg[idx] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    }
  };
  const int n_chunks = (K + kFimChunk - 1) / kFimChunk;
  fetch(0);
  for (int c = 0; c < n_chunks; ++c) {
#pragma unroll
    for (int i = 0; i < kFimPer; ++i)
      if (i * 64 + lane < c4) T4[i * 64 + lane] = reg[i];
    __syncthreads();
    if (c + 1 < n_chunks) fetch(c + 1);
    const int rows = min(kFimChunk, K - c * kFimChunk);
    for (int k = 0; k < rows; k += 2) {
      const int kk = k + half;
      const float v = (live && kk < rows) ? T[kk * P + col] : 0.0f;
      acc = __builtin_amdgcn_mfma_f32_32x32x2f32(v, v, acc, 0, 0, 0);
    }
    __syncthreads();
  }
  if (g >= G) return;
  float* Fg = F + static_cast<size_t>(g) * P * P;
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    const int i = 8 * (r >> 2) + 4 * half + (r & 3);
    if (i < P && col < P) Fg[i * P + col] = acc[r];
  }
}

// One workgroup per set: FP64 mean of the valid groups' F, + reg I, Cholesky, logdet.
__global__ __launch_bounds__(256) void fim_logdet_kernel(const float* __restrict__ F, const int* __restrict__ status,
                                                         int set_size, int P, const double* __restrict__ scale,
                                                         double reg, double* __restrict__ Fmean,
                                                         double* __restrict__ logdet, int* __restrict__ n_valid) {
  __shared__ double M[SPI_MAX_TANGENTS * SPI_MAX_TANGENTS];
  __shared__ int cnt;
  const int s = blockIdx.x;
  const size_t g0 = static_cast<size_t>(s) * set_size;
  if (threadIdx.x == 0) {
    int n = 0;
    for (int c = 0; c < set_size; ++c) n += status[g0 + c] == 0;
    cnt = n;
  }
  __syncthreads();
  const int n = cnt;
  for (int e = threadIdx.x; e < P * P; e += blockDim.x) {
    double acc = 0.0;
    for (int c = 0; c < set_size; ++c)
      if (status[g0 + c] == 0) acc += static_cast<double>(F[(g0 + c) * P * P + e]);
    const int i = e / P, j = e % P;
    double m = n > 0 ? acc / n : 0.0;
    if (scale) m *= scale[i] * scale[j];
    M[e] = m;
  }
  __syncthreads();
  if (threadIdx.x == 0) {            // relative ridge: reg * mean eigenvalue
    double tr = 0.0;
    for (int i = 0; i < P; ++i) tr += M[i * P + i];
    const double add = reg * (tr > 0.0 ? tr / P : 0.0);
    for (int i = 0; i < P; ++i) M[i * P + i] += add;
  }
  __syncthreads();
  if (Fmean)
    for (int e = threadIdx.x; e < P * P; e += blockDim.x) Fmean[static_cast<size_t>(s) * P * P + e] = M[e];
  __syncthreads();
  if (threadIdx.x == 0) {
    // symmetric part, then in-place Cholesky (lower) in FP64
    double ld = 0.0;
    bool pd = n > 0;
    for (int j = 0; j < P && pd; ++j) {
      double d = M[j * P + j];
      for (int k = 0; k < j; ++k) d -= M[j * P + k] * M[j * P + k];
      if (!(d > 0.0)) { pd = false; break; }
      const double ljj = sqrt(d);
      M[j * P + j] = ljj;
      ld += 2.0 * log(ljj);
      for (int i = j + 1; i < P; ++i) {
        double v = 0.5 * (M[i * P + j] + M[j * P + i]);
        for (int k = 0; k < j; ++k) v -= M[i * P + k] * M[j * P + k];
        M[i * P + j] = v / ljj;
      }
    }
    logdet[s] = pd ? ld : -INFINITY;
    n_valid[s] = n;
  }
}

__global__ void fourier_kernel(const float* __restrict__ coeff, int n_exc, int n_terms, int H, float dt, float period,
                               float scale, float* __restrict__ out) {
  const long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  const long long total = static_cast<long long>(n_exc) * H * 12;
  if (i >= total) return;
  const int j = static_cast<int>(i % 12), t = static_cast<int>((i / 12) % H), e = static_cast<int>(i / (12LL * H));
  const float* c = coeff + (static_cast<size_t>(e) * 12 + j) * n_terms;
  const double w = 2.0 * M_PI * t * static_cast<double>(dt) / period;
  double acc = 0.0;
  for (int k = 0; k < n_terms; ++k) acc += static_cast<double>(c[k]) * sin((k + 1) * w);
  out[i] = static_cast<float>(scale * acc);
}

// Gauss-Newton / Levenberg-Marquardt step per candidate from the augmented Gram matrices
// [J r]^T [J r] of its n_seg segment groups: FP64 sums A = J^T J, b = J^T r, rr = r^T r, then
// (A + lambda diag(A) + 1e-12 tr(A)/P I) delta = -b by FP64 Cholesky; pred = rr - model(delta).
// Every sum runs in a fixed order (segments in index order, k ascending), whichever thread
// computes it: the segment sums keep kGnBatch loads in flight per thread, and in Cholesky
// column j thread i > j computes L[i][j].
constexpr int kGnBatch = 16;
__global__ __launch_bounds__(256) void gn_step_kernel(const float* __restrict__ Gm, const int* __restrict__ status,
                                                      int n_seg, int P, const double* __restrict__ lambda,
                                                      double* __restrict__ delta, double* __restrict__ cost,
                                                      double* __restrict__ pred) {
  __shared__ double M[SPI_MAX_TANGENTS * SPI_MAX_TANGENTS];
  __shared__ double Ls[SPI_MAX_TANGENTS * SPI_MAX_TANGENTS];
  __shared__ double y[SPI_MAX_TANGENTS];
  const int c = blockIdx.x, L = P + 1, tid = threadIdx.x;
  const size_t g0 = static_cast<size_t>(c) * n_seg;
  int b = 0;
  for (int s = tid; s < n_seg; s += blockDim.x) b |= status[g0 + s];
  const bool bad = __syncthreads_or(b) != 0;
  const size_t LL = static_cast<size_t>(L) * L;
  for (int e = tid; e < L * L; e += blockDim.x) {
    const float* p = Gm + g0 * LL + e;
    double acc = 0.0;
    int s = 0;
    for (; s + kGnBatch <= n_seg; s += kGnBatch) {
      float v[kGnBatch];
#pragma unroll
      for (int u = 0; u < kGnBatch; ++u) v[u] = p[static_cast<size_t>(s + u) * LL];
#pragma unroll
      for (int u = 0; u < kGnBatch; ++u) acc += static_cast<double>(v[u]);
    }
    for (; s < n_seg; ++s) acc += static_cast<double>(p[static_cast<size_t>(s) * LL]);
    M[e] = acc;
  }
  __syncthreads();
  const double rr = M[P * L + P];
  double tr = 0.0;
  for (int i = 0; i < P; ++i) tr += M[i * L + i];
  const double lam = lambda[c], eps = 1e-12 * tr / P;
  bool pd = !bad;
  for (int j = 0; j < P && pd; ++j) {          // Cholesky of the damped normal matrix (lower)
    double d = M[j * L + j] * (1.0 + lam) + eps;   // every thread: the same bits, a uniform test
    for (int k = 0; k < j; ++k) d -= Ls[j * P + k] * Ls[j * P + k];
    if (!(d > 0.0)) { pd = false; break; }
    const double ljj = sqrt(d);
    if (tid == j) Ls[j * P + j] = ljj;
    if (tid > j && tid < P) {
      const int i = tid;
      double v = 0.5 * (M[i * L + j] + M[j * L + i]);
      for (int k = 0; k < j; ++k) v -= Ls[i * P + k] * Ls[j * P + k];
      Ls[i * P + j] = v / ljj;
    }
    __syncthreads();
  }
  if (tid != 0) return;
  cost[c] = bad ? INFINITY : rr;
  double* dc = delta + static_cast<size_t>(c) * P;
  if (!pd) {
    for (int i = 0; i < P; ++i) dc[i] = 0.0;
    pred[c] = bad ? INFINITY : rr;
    return;
  }
  for (int i = 0; i < P; ++i) {                // L y = -b
    double v = -0.5 * (M[i * L + P] + M[P * L + i]);
    for (int k = 0; k < i; ++k) v -= Ls[i * P + k] * y[k];
    y[i] = v / Ls[i * P + i];
  }
  for (int i = P - 1; i >= 0; --i) {           // L^T x = y
    double v = y[i];
    for (int k = i + 1; k < P; ++k) v -= Ls[k * P + i] * dc[k];
    dc[i] = v / Ls[i * P + i];
  }
  // predicted cost of the linear model: rr + 2 b.delta + delta^T A delta
  double q = rr;
  for (int i = 0; i < P; ++i) {
    double Ad = 0.0;
    for (int k = 0; k < P; ++k) Ad += M[i * L + k] * dc[k];
    q += 2.0 * M[i * L + P] * dc[i] + dc[i] * Ad;
  }
  pred[c] = q;
}

}  // namespace

extern "C" {

int spi_fim_assemble(const float* J_dev, long long n_groups, int n_rows, int n_tan, float* F_dev,
                     hipStream_t stream) {
  if (n_groups < 0 || n_rows < 1 || n_tan < 1 || n_tan > SPI_MAX_TANGENTS || !J_dev || !F_dev)
    return ffail(SPI_ERR_INVALID_ARG, "spi_fim_assemble: bad arguments");
  if (n_groups == 0) return SPI_OK;
  const long long blocks = (n_groups + 3) / 4;
  if (blocks > 0x7fffffffLL) return ffail(SPI_ERR_TOO_LARGE, "spi_fim_assemble: too many groups");
  if (static_cast<long long>(n_rows) * n_tan % 4 == 0 && reinterpret_cast<uintptr_t>(J_dev) % 16 == 0 &&
      static_cast<long long>(n_rows) * n_tan <= 0x7fffffffLL) {
    fim_assemble_vec_kernel<<<static_cast<unsigned>(blocks), 256, 0, stream>>>(J_dev, n_groups, n_rows, n_tan, F_dev);
    return flaunch("fim_assemble_vec_kernel");
  }
  fim_assemble_kernel<<<static_cast<unsigned>(blocks), 256, 0, stream>>>(J_dev, n_groups, n_rows, n_tan, F_dev);
  return flaunch("fim_assemble_kernel");
}

int spi_fim_logdet(const float* F_dev, const int* status_dev, int n_sets, int set_size, int n_tan,
                   const double* scale_dev, double reg, double* F_mean_dev, double* logdet_dev, int* n_valid_dev,
                   hipStream_t stream) {
  if (n_sets < 0 || set_size < 1 || n_tan < 1 || n_tan > SPI_MAX_TANGENTS || !(reg >= 0.0) || !F_dev ||
      !status_dev || !logdet_dev || !n_valid_dev)
    return ffail(SPI_ERR_INVALID_ARG, "spi_fim_logdet: bad arguments");
  if (n_sets == 0) return SPI_OK;
  fim_logdet_kernel<<<n_sets, 256, 0, stream>>>(F_dev, status_dev, set_size, n_tan, scale_dev, reg, F_mean_dev,
                                                logdet_dev, n_valid_dev);
  return flaunch("fim_logdet_kernel");
}

int spi_fourier_excitations(const float* coeff_dev, int n_exc, int n_terms, int horizon, float dt_ctrl, float period,
                            float scale, float* actions_dev, hipStream_t stream) {
  if (n_exc < 0 || n_terms < 0 || horizon < 0 || !(dt_ctrl > 0.0f) || !(period > 0.0f) || !coeff_dev || !actions_dev)
    return ffail(SPI_ERR_INVALID_ARG, "spi_fourier_excitations: bad arguments");
  const long long total = static_cast<long long>(n_exc) * horizon * 12;
  if (total == 0) return SPI_OK;
  const long long blocks = (total + 255) / 256;
  if (blocks > 0x7fffffffLL) return ffail(SPI_ERR_TOO_LARGE, "spi_fourier_excitations: too large");
  fourier_kernel<<<static_cast<unsigned>(blocks), 256, 0, stream>>>(coeff_dev, n_exc, n_terms, horizon, dt_ctrl,
                                                                    period, scale, actions_dev);
  return flaunch("fourier_kernel");
}

int spi_gn_step(const float* gram_dev, const int* status_dev, int n_cand, int n_seg, int n_tan, const double* lambda_dev,
                double* delta_dev, double* cost_dev, double* pred_dev, hipStream_t stream) {
  if (n_cand < 0 || n_seg < 1 || n_tan < 1 || n_tan + 1 > SPI_MAX_TANGENTS || !gram_dev || !status_dev ||
      !lambda_dev || !delta_dev || !cost_dev || !pred_dev)
    return ffail(SPI_ERR_INVALID_ARG, "spi_gn_step: bad arguments");
  if (n_cand == 0) return SPI_OK;
  gn_step_kernel<<<n_cand, 256, 0, stream>>>(gram_dev, status_dev, n_seg, n_tan, lambda_dev, delta_dev, cost_dev,
                                             pred_dev);
  return flaunch("gn_step_kernel");
}

}  // extern "C"
