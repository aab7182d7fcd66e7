// Command sequences of the active-exploration optimizer: one thread per (trial, step), all
// command dimensions of the step.  Strict IEEE float32 arithmetic in the order of the
// reference's numpy code, so the sequences match it to float rounding.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>

#include "spi_types.h"

namespace {

int cfail(int code, const char* msg) {
  spi_set_error(msg);
  return code;
}

int n_params_of(const SpiCommandSpec& c) {
  switch (c.mode) {
    case SPI_COMMAND_CONSTANT: return c.n_sampled * c.n_updates;
    case SPI_COMMAND_POLYNOMIAL: return c.n_updates * c.n_sampled * (c.poly_degree + 1);
    case SPI_COMMAND_BEZIER: return c.n_updates * c.n_sampled * c.n_bezier;
    default: return -1;
  }
}

__device__ float binom_f(int n, int k) {   // C(n, k) as the reference forms it: product, then int
  if (k > n - k) k = n - k;
  double r = 1.0;
  for (int i = 0; i < k; ++i) r *= static_cast<double>(n - i) / (i + 1);
  return static_cast<float>(static_cast<long long>(r));
}

constexpr int kMaxBasis = 8;    // polynomial degree + 1 or Bezier points held in registers

// The n_dims commands of pair idx = trial * T + step into o: the step's position t in its
// update window and its power or Bernstein basis are formed once (the same calls, in the same
// order, as per element).
template <typename Idx>
__device__ __forceinline__ void command_step(const SpiCommandSpec& c, const float* __restrict__ params,
                                             int n_params, Idx T, Idx idx, float* o) {
  const int b = static_cast<int>(idx / T);
  const int step = static_cast<int>(idx - static_cast<Idx>(b) * T);
  const float* p = params + static_cast<size_t>(b) * n_params;
  const int S = c.steps_per_update;
  const int u = step / S, i = step - u * S;
  // np.linspace(0, 1, S, dtype=float32): computed in float64, then rounded
  const float t = S > 1 ? static_cast<float>(static_cast<double>(i) / (S - 1)) : 0.0f;
  const int nb = c.mode == SPI_COMMAND_POLYNOMIAL ? c.poly_degree + 1 : c.mode == SPI_COMMAND_BEZIER ? c.n_bezier : 0;
  float basis[kMaxBasis];
  // unrolled over kMaxBasis with guards, so that the basis stays in registers
  if (c.mode == SPI_COMMAND_POLYNOMIAL) {
#pragma unroll
    for (int q = 0; q < kMaxBasis; ++q)
      if (q < nb) basis[q] = q == 0 ? 1.0f : powf(t, static_cast<float>(q));
  } else if (c.mode == SPI_COMMAND_BEZIER) {
    const int n = nb - 1;
    const float omt = 1.0f - t;
#pragma unroll
    for (int j = 0; j < kMaxBasis; ++j)
      if (j <= n)
        basis[j] = binom_f(n, j) * (j == 0 ? 1.0f : powf(t, static_cast<float>(j))) *
                   (n - j == 0 ? 1.0f : powf(omt, static_cast<float>(n - j)));
  }
  for (int k = 0; k < c.n_dims; ++k) {
    int d = -1;
    for (int m = 0; m < c.n_sampled; ++m)
      if (c.sampled_idx[m] == k) d = m;               // the last listed wins, as the reference's fill loop
    float v = c.default_cmd[k];
    if (d >= 0) {
      const float lo = c.lo[k], hi = c.hi[k];
      if (c.mode == SPI_COMMAND_CONSTANT) {
        v = p[d * c.n_updates + u];
      } else if (c.mode == SPI_COMMAND_POLYNOMIAL) {
        const float* cf = p + (static_cast<size_t>(u) * c.n_sampled + d) * nb;
        float poly = 0.0f;
#pragma unroll
        for (int q = 0; q < kMaxBasis; ++q)
          if (q < nb) poly = poly + cf[q] * basis[q];
        for (int q = kMaxBasis; q < nb; ++q) poly = poly + cf[q] * powf(t, static_cast<float>(q));
        poly = tanhf(poly);
        v = lo + (poly + 1.0f) * 0.5f * (hi - lo);
      } else {
        const int n = nb - 1;
        const float* cp = p + (static_cast<size_t>(u) * c.n_sampled + d) * nb;
        const float omt = 1.0f - t;
        float acc = 0.0f;
#pragma unroll
        for (int j = 0; j < kMaxBasis; ++j)
          if (j <= n) acc = acc + basis[j] * cp[j];
        for (int j = kMaxBasis; j <= n; ++j)
          acc = acc + binom_f(n, j) * powf(t, static_cast<float>(j)) *
                          (n - j == 0 ? 1.0f : powf(omt, static_cast<float>(n - j))) * cp[j];
        v = fminf(fmaxf(acc, lo), hi);
      }
    }
    o[k] = v;
  }
}

// One thread per (trial, step), its commands staged in LDS; the workgroup's 256 steps are one
// contiguous run of the output, written back with consecutive threads on consecutive words.
// Idx: int while the launch has fewer than 2^31 (trial, step) pairs, long long above.
template <typename Idx>
__global__ __launch_bounds__(256) void command_kernel(SpiCommandSpec c, const float* __restrict__ params,
                                                     int n_trials, int n_params, float* __restrict__ out) {
  __shared__ float stage[256 * SPI_MAX_COMMAND_DIMS];
  const Idx T = static_cast<Idx>(c.n_updates) * c.steps_per_update;
  const Idx pairs = static_cast<Idx>(n_trials) * T;
  const Idx base = static_cast<Idx>(blockIdx.x) * blockDim.x;
  const Idx idx = base + threadIdx.x;
  if (idx < pairs) command_step(c, params, n_params, T, idx, stage + threadIdx.x * c.n_dims);
  __syncthreads();
  const Idx left = pairs - base;
  const int n_out = static_cast<int>((left < static_cast<Idx>(blockDim.x) ? left : static_cast<Idx>(blockDim.x)) *
                                     c.n_dims);
  float* o = out + static_cast<size_t>(base) * c.n_dims;
  for (int e = threadIdx.x; e < n_out; e += blockDim.x) o[e] = stage[e];
}

}  // namespace

extern "C" {

int spi_command_sequences(const SpiCommandSpec* spec, const float* params_dev, int n_trials, int n_params,
                          float* out_dev, hipStream_t stream) {
  if (!spec || spec->n_dims < 1 || spec->n_dims > SPI_MAX_COMMAND_DIMS || spec->n_sampled < 0 ||
      spec->n_sampled > SPI_MAX_COMMAND_DIMS || spec->n_updates < 1 || spec->steps_per_update < 1 ||
      spec->poly_degree < 0 || spec->n_bezier < 1)
    return cfail(SPI_ERR_INVALID_ARG, "spi_command_sequences: bad spec");
  for (int m = 0; m < spec->n_sampled; ++m)
    if (spec->sampled_idx[m] < 0 || spec->sampled_idx[m] >= spec->n_dims)
      return cfail(SPI_ERR_INVALID_ARG, "spi_command_sequences: sampled index out of range");
  const int need = n_params_of(*spec);
  if (need < 0) return cfail(SPI_ERR_INVALID_ARG, "spi_command_sequences: unknown mode");
  if (n_trials < 0 || n_params != need || (n_trials > 0 && (!params_dev || !out_dev)))
    return cfail(SPI_ERR_INVALID_ARG, "spi_command_sequences: n_params does not match the spec");
  const long long pairs = static_cast<long long>(n_trials) * spec->n_updates * spec->steps_per_update;
  if (pairs == 0) return SPI_OK;
  const long long blocks = (pairs + 255) / 256;
  if (blocks > 0x7fffffffLL) return cfail(SPI_ERR_TOO_LARGE, "spi_command_sequences: too large");
  if (blocks * 256 < 0x7fffffffLL)
    command_kernel<int><<<static_cast<unsigned>(blocks), 256, 0, stream>>>(*spec, params_dev, n_trials, n_params,
                                                                           out_dev);
  else
    command_kernel<long long><<<static_cast<unsigned>(blocks), 256, 0, stream>>>(*spec, params_dev, n_trials,
                                                                                 n_params, out_dev);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return cfail(SPI_ERR_HIP, hipGetErrorString(e));
  return SPI_OK;
}

}  // extern "C"
