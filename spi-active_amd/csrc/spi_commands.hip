// Command sequences of the active-exploration optimizer: one thread per (trial, step,
// command dimension).  Strict IEEE float32 arithmetic in the order of the reference's numpy
// code, so the sequences match it to float rounding.
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

__global__ void command_kernel(SpiCommandSpec c, const float* __restrict__ params, int n_trials, int n_params,
                               float* __restrict__ out) {
  const long long T = static_cast<long long>(c.n_updates) * c.steps_per_update;
  const long long total = static_cast<long long>(n_trials) * T * c.n_dims;
  const long long idx = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (idx >= total) return;
  const int k = static_cast<int>(idx % c.n_dims);
  const long long step = (idx / c.n_dims) % T;
  const int b = static_cast<int>(idx / (static_cast<long long>(c.n_dims) * T));
  int d = -1;
  for (int m = 0; m < c.n_sampled; ++m)
    if (c.sampled_idx[m] == k) d = m;                 // the last listed wins, as the reference's fill loop
  if (d < 0) { out[idx] = c.default_cmd[k]; return; }
  const float* p = params + static_cast<size_t>(b) * n_params;
  const int S = c.steps_per_update;
  const int u = static_cast<int>(step / S), i = static_cast<int>(step % S);
  // np.linspace(0, 1, S, dtype=float32): computed in float64, then rounded
  const float t = S > 1 ? static_cast<float>(static_cast<double>(i) / (S - 1)) : 0.0f;
  const float lo = c.lo[k], hi = c.hi[k];
  float v;
  if (c.mode == SPI_COMMAND_CONSTANT) {
    v = p[d * c.n_updates + u];
  } else if (c.mode == SPI_COMMAND_POLYNOMIAL) {
    const int nc = c.poly_degree + 1;
    const float* cf = p + (static_cast<size_t>(u) * c.n_sampled + d) * nc;
    float poly = 0.0f;
    for (int q = 0; q < nc; ++q) poly = poly + cf[q] * (q == 0 ? 1.0f : powf(t, static_cast<float>(q)));
    poly = tanhf(poly);
    v = lo + (poly + 1.0f) * 0.5f * (hi - lo);
  } else {
    const int n = c.n_bezier - 1;
    const float* cp = p + (static_cast<size_t>(u) * c.n_sampled + d) * c.n_bezier;
    float acc = 0.0f;
    const float omt = 1.0f - t;
    for (int j = 0; j <= n; ++j) {
      const float bern = binom_f(n, j) * (j == 0 ? 1.0f : powf(t, static_cast<float>(j))) *
                         (n - j == 0 ? 1.0f : powf(omt, static_cast<float>(n - j)));
      acc = acc + bern * cp[j];
    }
    v = fminf(fmaxf(acc, lo), hi);
  }
  out[idx] = v;
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
  const long long total = static_cast<long long>(n_trials) * spec->n_updates * spec->steps_per_update * spec->n_dims;
  if (total == 0) return SPI_OK;
  const long long blocks = (total + 255) / 256;
  if (blocks > 0x7fffffffLL) return cfail(SPI_ERR_TOO_LARGE, "spi_command_sequences: too large");
  command_kernel<<<static_cast<unsigned>(blocks), 256, 0, stream>>>(*spec, params_dev, n_trials, n_params, out_dev);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return cfail(SPI_ERR_HIP, hipGetErrorString(e));
  return SPI_OK;
}

}  // extern "C"
