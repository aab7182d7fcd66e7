// Policy network inference (gfx950): the actor MLP of the locomotion policy (Linear layers with
// ELU between them, spigym/agents/modules/modules.py:47-63; hidden sizes 512-256-128 from
// config/algo/ppo.yaml) as FP32 MFMA (v_mfma_f32_32x32x2f32) tiles.  One launch per layer:
// a workgroup of 4 wavefronts owns a 32-row x 128-column output tile; the row block of X and
// the 128 weight rows are staged through LDS in K chunks of 64, padded to 65 words per row so
// the column-wise operand reads hit distinct banks.  Strict IEEE float32 (no fast-math).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstdio>

#include "spi_types.h"

namespace {

typedef float floatx16 __attribute__((ext_vector_type(16)));

constexpr int kRows = 32;    // output rows per workgroup (MFMA M)
constexpr int kCols = 128;   // output columns per workgroup (4 wavefronts x MFMA N)
constexpr int kKc = 64;      // K chunk staged per pass
constexpr int kPad = kKc + 1;
constexpr int kMaxLayers = 8;

// Y[m][n] = act(sum_k X[m][k] W[n][k] + b[n]); W row-major [N][K] (torch Linear layout).
// MFMA operands: lane l supplies A[l % 32][k + l / 32] = X-tile element and
// B[k + l / 32][l % 32] = W[n0 + l % 32][k + l / 32]; accumulator r of lane l holds
// D[8 (r / 4) + 4 (l / 32) + r % 4][l % 32].  The next K chunk is fetched from global memory
// into registers while the MFMAs consume the current one from LDS.  kVec: rows of X and W are
// 16-byte aligned and K is a multiple of 4, so each thread moves float4s.
template <bool kElu, bool kVec>
__global__ __launch_bounds__(256) void linear_kernel(const float* __restrict__ X, int ldx, int M, int K,
                                                     const float* __restrict__ W, const float* __restrict__ b, int N,
                                                     float* __restrict__ Y, int ldy) {
  __shared__ float As[kRows * kPad];
  __shared__ float Ws[kCols * kPad];
  const int m0 = blockIdx.x * kRows, n0 = blockIdx.y * kCols;
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int col = lane & 31, half = lane >> 5;
  const bool live = n0 + wave * 32 < N;  // uniform per wavefront
  floatx16 acc;
#pragma unroll
  for (int r = 0; r < 16; ++r) acc[r] = 0.0f;
  constexpr int kV = kVec ? 4 : 1;                     // floats per load
  constexpr int kTPR = kKc / kV;                       // threads per tile row
  constexpr int kRowStep = 256 / kTPR;                 // rows covered by one pass of the block
  constexpr int kAPer = kRows / kRowStep, kWPer = kCols / kRowStep;
  const int r0 = tid / kTPR, c0 = (tid % kTPR) * kV;
  float ra[kAPer][kV], rw[kWPer][kV];
  auto fetch = [&](int k0) {
    const int gk = k0 + c0;
#pragma unroll
    for (int i = 0; i < kAPer; ++i) {
      const int gm = m0 + r0 + i * kRowStep;
      if (kVec) {
        float4 v = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        if (gm < M && gk < K) v = *reinterpret_cast<const float4*>(X + static_cast<size_t>(gm) * ldx + gk);
        ra[i][0] = v.x; ra[i][kV > 1 ? 1 : 0] = v.y; ra[i][kV > 2 ? 2 : 0] = v.z; ra[i][kV > 3 ? 3 : 0] = v.w;
      } else {
        ra[i][0] = (gm < M && gk < K) ? X[static_cast<size_t>(gm) * ldx + gk] : 0.0f;
      }
    }
#pragma unroll
    for (int i = 0; i < kWPer; ++i) {
      const int gn = n0 + r0 + i * kRowStep;
      if (kVec) {
        float4 v = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        if (gn < N && gk < K) v = *reinterpret_cast<const float4*>(W + static_cast<size_t>(gn) * K + gk);
        rw[i][0] = v.x; rw[i][kV > 1 ? 1 : 0] = v.y; rw[i][kV > 2 ? 2 : 0] = v.z; rw[i][kV > 3 ? 3 : 0] = v.w;
      } else {
        rw[i][0] = (gn < N && gk < K) ? W[static_cast<size_t>(gn) * K + gk] : 0.0f;
      }
    }
  };
  fetch(0);
  for (int k0 = 0; k0 < K; k0 += kKc) {
#pragma unroll
    for (int i = 0; i < kAPer; ++i)
#pragma unroll
      for (int v = 0; v < kV; ++v) As[(r0 + i * kRowStep) * kPad + c0 + v] = ra[i][v];
#pragma unroll
    for (int i = 0; i < kWPer; ++i)
#pragma unroll
      for (int v = 0; v < kV; ++v) Ws[(r0 + i * kRowStep) * kPad + c0 + v] = rw[i][v];
    __syncthreads();
    if (k0 + kKc < K) fetch(k0 + kKc);
    if (live) {
      const float* ap = As + col * kPad + half;
      const float* bp = Ws + (wave * 32 + col) * kPad + half;
#pragma unroll
      for (int kk = 0; kk < kKc; kk += 2) acc = __builtin_amdgcn_mfma_f32_32x32x2f32(ap[kk], bp[kk], acc, 0, 0, 0);
    }
    __syncthreads();
  }
  if (!live) return;
  const int gn = n0 + wave * 32 + col;
  if (gn >= N) return;
  const float bias = b[gn];
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    const int gm = m0 + 8 * (r >> 2) + 4 * half + (r & 3);
    if (gm < M) {
      float v = acc[r] + bias;
      if (kElu) v = v > 0.0f ? v : expm1f(v);
      Y[static_cast<size_t>(gm) * ldy + gn] = v;
    }
  }
}

int mfail(int code, const char* msg) {
  spi_set_error(msg);
  return code;
}

int launch_linear(const float* X, int ldx, int M, int K, const float* W, const float* b, int N, float* Y, int ldy,
                  bool elu, hipStream_t stream) {
  const dim3 grid((M + kRows - 1) / kRows, (N + kCols - 1) / kCols);
  const bool vec = K % 4 == 0 && ldx % 4 == 0 && reinterpret_cast<uintptr_t>(X) % 16 == 0 &&
                   reinterpret_cast<uintptr_t>(W) % 16 == 0;
  if (vec) {
    if (elu)
      linear_kernel<true, true><<<grid, 256, 0, stream>>>(X, ldx, M, K, W, b, N, Y, ldy);
    else
      linear_kernel<false, true><<<grid, 256, 0, stream>>>(X, ldx, M, K, W, b, N, Y, ldy);
  } else {
    if (elu)
      linear_kernel<true, false><<<grid, 256, 0, stream>>>(X, ldx, M, K, W, b, N, Y, ldy);
    else
      linear_kernel<false, false><<<grid, 256, 0, stream>>>(X, ldx, M, K, W, b, N, Y, ldy);
  }
  const hipError_t e = hipGetLastError();
  if (e == hipSuccess) return SPI_OK;
  char buf[256];
  std::snprintf(buf, sizeof(buf), "linear_kernel: %s", hipGetErrorString(e));
  spi_set_error(buf);
  return SPI_ERR_HIP;
}

}  // namespace

extern "C" {

long long spi_mlp_param_floats(int n_layers, const int* dims) {
  if (n_layers < 1 || n_layers > kMaxLayers || !dims) return -1;
  long long n = 0;
  for (int l = 0; l < n_layers; ++l) {
    if (dims[l] < 1 || dims[l + 1] < 1) return -1;
    n += static_cast<long long>(dims[l + 1]) * dims[l] + dims[l + 1];
  }
  return n;
}

long long spi_mlp_work_floats(int M, int n_layers, const int* dims) {
  if (M < 0 || spi_mlp_param_floats(n_layers, dims) < 0) return -1;
  int widest = 0;
  for (int l = 1; l < n_layers; ++l) widest = dims[l] > widest ? dims[l] : widest;
  return 2LL * M * widest;
}

int spi_mlp_forward(const float* x_dev, int M, int n_layers, const int* dims, const float* params_dev, float* work_dev,
                    long long work_floats, float* y_dev, hipStream_t stream) {
  const long long need = spi_mlp_work_floats(M, n_layers, dims);
  if (need < 0 || !x_dev || !params_dev || !y_dev || (n_layers > 1 && !work_dev))
    return mfail(SPI_ERR_INVALID_ARG, "spi_mlp_forward: bad arguments");
  if (work_floats < need) return mfail(SPI_ERR_INVALID_ARG, "spi_mlp_forward: work buffer too small");
  for (int l = 0; l <= n_layers; ++l)
    if (static_cast<long long>(M) * dims[l] > 0x7fffffffLL)
      return mfail(SPI_ERR_TOO_LARGE, "spi_mlp_forward: activations exceed 2^31 elements");
  if (M == 0) return SPI_OK;
  int widest = 0;
  for (int l = 1; l < n_layers; ++l) widest = dims[l] > widest ? dims[l] : widest;
  const float* in = x_dev;
  const float* p = params_dev;
  for (int l = 0; l < n_layers; ++l) {
    const int K = dims[l], N = dims[l + 1];
    const bool last = l == n_layers - 1;
    float* out = last ? y_dev : work_dev + static_cast<size_t>(l & 1) * M * widest;
    const int rc = launch_linear(in, K, M, K, p, p + static_cast<size_t>(N) * K, N, out, N, !last, stream);
    if (rc != SPI_OK) return rc;
    p += static_cast<size_t>(N) * K + N;
    in = out;
  }
  return SPI_OK;
}

}  // extern "C"
