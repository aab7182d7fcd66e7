// Device-resident dataset windows: packs sliding windows of a recorded trajectory straight
// into the rollout kernels' segment records (init state, then per step actions + target).
#include <hip/hip_runtime.h>

#include <cstdio>

#include "spi_types.h"

namespace {

constexpr int kRow = SPI_STATE_DIM + 12;   // one staged row: state, then action
constexpr int kWinPerBlock = 64;            // consecutive windows per workgroup
constexpr int kMaxStagedH = 128;            // longer horizons read the rows directly

// Element e of the segment record of a window starting at row t, from rows addressed by
// st(row, i) (state component i) and ac(row, c) (action component c).
template <class St, class Ac>
__device__ __forceinline__ float record_elem(int e, int t, St st, Ac ac) {
  if (e < SPI_STATE_DIM) return st(t, e);
  const int r = e - SPI_STATE_DIM, k = r / kRow, c = r % kRow;
  return c < 12 ? ac(t + k, c) : st(t + k + 1, c - 12);
}

// Workgroup g packs windows g * 64 .. + 63; window w starts at row rows[w] (rows == nullptr:
// first + w).  When the group's windows start at consecutive rows r0, r0 + 1, ... and H <= 128,
// rows r0 .. r0 + 63 + H are staged once in LDS (consecutive windows share all but one row)
// and the group's records, one contiguous run of the output, are written with consecutive
// threads on consecutive words; otherwise each element is read from the buffers.  A window
// whose rows r .. r + H do not all lie below n_rows is written as NaN and nothing outside the
// buffers is read.
__global__ __launch_bounds__(256) void pack_windows_kernel(const float* __restrict__ states,
                                                           const float* __restrict__ actions,
                                                           const int* __restrict__ rows, int first, int n_rows,
                                                           int H, int n_windows, int stride,
                                                           float* __restrict__ out) {
  extern __shared__ float stage[];
  const int w0 = blockIdx.x * kWinPerBlock;
  const int nw = min(kWinPerBlock, n_windows - w0);
  const int r0 = rows ? rows[w0] : first + w0;
  const int j = threadIdx.x;
  // unsigned: "window j starts j rows after window 0" without signed overflow
  const unsigned rj = j < nw ? static_cast<unsigned>(rows ? rows[w0 + j] : first + w0 + j)
                             : static_cast<unsigned>(r0) + j;
  const bool staged = H <= kMaxStagedH && __syncthreads_and(rj == static_cast<unsigned>(r0) + j) && r0 >= 0 &&
                      static_cast<long long>(r0) + nw - 1 + H < n_rows;
  float* o = out + static_cast<size_t>(w0) * stride;
  if (staged) {
    const int n_stage = (nw + H) * kRow;
    for (int i = threadIdx.x; i < n_stage; i += blockDim.x) {
      const int row = i / kRow, c = i % kRow;
      stage[i] = c < SPI_STATE_DIM ? states[static_cast<size_t>(r0 + row) * SPI_STATE_DIM + c]
                                   : actions[static_cast<size_t>(r0 + row) * 12 + (c - SPI_STATE_DIM)];
    }
    __syncthreads();
    auto st = [&](int row, int i) { return stage[(row - r0) * kRow + i]; };
    auto ac = [&](int row, int c) { return stage[(row - r0) * kRow + SPI_STATE_DIM + c]; };
    for (int w = 0; w < nw; ++w)
      for (int e = threadIdx.x; e < stride; e += blockDim.x)
        o[static_cast<size_t>(w) * stride + e] = record_elem(e, r0 + w, st, ac);
    return;
  }
  auto st = [&](int row, int i) { return states[static_cast<size_t>(row) * SPI_STATE_DIM + i]; };
  auto ac = [&](int row, int c) { return actions[static_cast<size_t>(row) * 12 + c]; };
  for (int w = 0; w < nw; ++w) {
    const int t = rows ? rows[w0 + w] : first + w0 + w;
    const bool ok = t >= 0 && static_cast<long long>(t) + H < n_rows;
    for (int e = threadIdx.x; e < stride; e += blockDim.x)
      o[static_cast<size_t>(w) * stride + e] = ok ? record_elem(e, t, st, ac) : __int_as_float(0x7fc00000);
  }
}

int pack_launch(const float* states_dev, const float* actions_dev, const int* rows_dev, int first, int n_rows,
                int horizon, int n_windows, float* segments_dev, hipStream_t stream) {
  const int stride = SPI_STATE_DIM + horizon * kRow;
  const int blocks = (n_windows + kWinPerBlock - 1) / kWinPerBlock;
  const size_t lds = horizon <= kMaxStagedH ? static_cast<size_t>(kWinPerBlock + horizon) * kRow * sizeof(float) : 0;
  pack_windows_kernel<<<blocks, 256, lds, stream>>>(states_dev, actions_dev, rows_dev, first, n_rows, horizon,
                                                    n_windows, stride, segments_dev);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    spi_set_error(hipGetErrorString(e));
    return SPI_ERR_HIP;
  }
  return SPI_OK;
}

}  // namespace

extern "C" {

int spi_pack_windows(const float* states_dev, const float* actions_dev, int n_steps, int horizon, int first,
                     int n_windows, float* segments_dev, hipStream_t stream) {
  if (n_steps < 0 || horizon < 1 || first < 0 || n_windows < 0 ||
      static_cast<long long>(first) + n_windows + horizon > n_steps ||
      (n_windows > 0 && (!states_dev || !actions_dev || !segments_dev))) {
    spi_set_error("spi_pack_windows: windows must lie inside the trajectory (first + n + horizon <= n_steps)");
    return SPI_ERR_INVALID_ARG;
  }
  if (n_windows == 0) return SPI_OK;
  return pack_launch(states_dev, actions_dev, nullptr, first, n_steps, horizon, n_windows, segments_dev, stream);
}

int spi_pack_windows_rows(const float* states_dev, const float* actions_dev, int n_rows, const int* rows_dev,
                          int n_windows, int horizon, float* segments_dev, hipStream_t stream) {
  if (n_rows < 0 || horizon < 1 || n_windows < 0 ||
      (n_windows > 0 && (!states_dev || !actions_dev || !rows_dev || !segments_dev))) {
    spi_set_error("spi_pack_windows_rows: bad sizes or null pointer");
    return SPI_ERR_INVALID_ARG;
  }
  if (n_windows == 0) return SPI_OK;
  return pack_launch(states_dev, actions_dev, rows_dev, 0, n_rows, horizon, n_windows, segments_dev, stream);
}

}  // extern "C"
