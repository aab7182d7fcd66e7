// Device-resident dataset windows: packs sliding windows of a recorded trajectory straight
// into the rollout kernels' segment records (init state, then per step actions + target).
#include <hip/hip_runtime.h>

#include <cstdio>

#include "spi_types.h"

namespace {

__global__ void pack_windows_kernel(const float* __restrict__ states, const float* __restrict__ actions, int H,
                                    int first, int n_windows, int stride, float* __restrict__ out) {
  const long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  const long long total = static_cast<long long>(n_windows) * stride;
  if (i >= total) return;
  const int w = static_cast<int>(i / stride), e = static_cast<int>(i % stride);
  const int t = first + w;
  float v;
  if (e < SPI_STATE_DIM) {
    v = states[static_cast<size_t>(t) * SPI_STATE_DIM + e];
  } else {
    const int r = e - SPI_STATE_DIM, k = r / (12 + SPI_STATE_DIM), c = r % (12 + SPI_STATE_DIM);
    v = c < 12 ? actions[static_cast<size_t>(t + k) * 12 + c]
               : states[static_cast<size_t>(t + k + 1) * SPI_STATE_DIM + (c - 12)];
  }
  out[i] = v;
}

}  // namespace

extern "C" {

int spi_pack_windows(const float* states_dev, const float* actions_dev, int n_steps, int horizon, int first,
                     int n_windows, float* segments_dev, hipStream_t stream) {
  if (n_steps < 0 || horizon < 1 || first < 0 || n_windows < 0 || first + n_windows + horizon > n_steps ||
      (n_windows > 0 && (!states_dev || !actions_dev || !segments_dev))) {
    spi_set_error("spi_pack_windows: windows must lie inside the trajectory (first + n + horizon <= n_steps)");
    return SPI_ERR_INVALID_ARG;
  }
  if (n_windows == 0) return SPI_OK;
  const int stride = SPI_STATE_DIM + horizon * (12 + SPI_STATE_DIM);
  const long long total = static_cast<long long>(n_windows) * stride;
  const long long blocks = (total + 255) / 256;
  if (blocks > 0x7fffffffLL) {
    spi_set_error("spi_pack_windows: too large");
    return SPI_ERR_TOO_LARGE;
  }
  pack_windows_kernel<<<static_cast<unsigned>(blocks), 256, 0, stream>>>(states_dev, actions_dev, horizon, first,
                                                                         n_windows, stride, segments_dev);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    spi_set_error(hipGetErrorString(e));
    return SPI_ERR_HIP;
  }
  return SPI_OK;
}

}  // extern "C"
