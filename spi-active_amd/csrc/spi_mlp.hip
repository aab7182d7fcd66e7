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

// Layer input gathered from a ring of per-step observation terms, stored term-major
// ring[slot][term][m] (m = row, fastest): X[m][k] = clip(ring[slot_k][term_k][m]) with
// slot_k = (head - kmap[k].x) mod n_slots and term_k = kmap[k].y.
struct RingSrc {
  const float* ring;
  int ld;          // row stride of a (slot, term) line: >= M
  int n_slots, n_terms, head;
  const int2* kmap;
  float clip;
};

// Y[m][n] = act(sum_k X[m][k] W[n][k] + b[n]); W row-major [N][K] (torch Linear layout).
// MFMA operands: lane l supplies A[l % 32][k + l / 32] = X-tile element and
// B[k + l / 32][l % 32] = W[n0 + l % 32][k + l / 32]; accumulator r of lane l holds
// D[8 (r / 4) + 4 (l / 32) + r % 4][l % 32].  The next K chunk is fetched from global memory
// into registers while the MFMAs consume the current one from LDS.
// kSrc 0: X row-major, scalar loads; 1: X and W rows 16-byte aligned with K % 4 == 0, float4
// loads; 2: X gathered from a RingSrc (32 consecutive rows per load), W as in 1.
template <bool kElu, int kSrc>
__global__ __launch_bounds__(256) void linear_kernel(const float* __restrict__ X, int ldx, RingSrc rs, int M, int K,
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
  constexpr int kV = kSrc == 0 ? 1 : 4;                // floats per W (and row-major X) load
  constexpr int kTPR = kKc / kV;                       // threads per tile row
  constexpr int kRowStep = 256 / kTPR;                 // rows covered by one pass of the block
  constexpr int kAV = kSrc == 2 ? 1 : kV;
  constexpr int kAPer = kSrc == 2 ? kRows * kKc / 256 : kRows / kRowStep, kWPer = kCols / kRowStep;
  const int r0 = tid / kTPR, c0 = (tid % kTPR) * kV;
  const int gm_r = m0 + (tid & 31), kk_r = tid >> 5;  // ring gather: row, first K offset
  float ra[kAPer][kAV], rw[kWPer][kV];
  auto fetch = [&](int k0) {
    const int gk = k0 + c0;
#pragma unroll
    for (int i = 0; i < kAPer; ++i) {
      if (kSrc == 2) {
        const int k = k0 + kk_r + 8 * i;
        float v = 0.0f;
        if (gm_r < M && k < K) {
          const int2 km = rs.kmap[k];
          int slot = (rs.head - km.x) % rs.n_slots;           // any table entry stays inside the ring
          slot += slot < 0 ? rs.n_slots : 0;
          const int term = min(max(km.y, 0), rs.n_terms - 1);
          v = rs.ring[(static_cast<size_t>(slot) * rs.n_terms + term) * rs.ld + gm_r];
          v = fminf(fmaxf(v, -rs.clip), rs.clip);
        }
        ra[i][0] = v;
      } else {
        const int gm = m0 + r0 + i * kRowStep;
        if (kSrc == 1) {
          float4 v = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
          if (gm < M && gk < K) v = *reinterpret_cast<const float4*>(X + static_cast<size_t>(gm) * ldx + gk);
          ra[i][0] = v.x; ra[i][kAV > 1 ? 1 : 0] = v.y; ra[i][kAV > 2 ? 2 : 0] = v.z; ra[i][kAV > 3 ? 3 : 0] = v.w;
        } else {
          ra[i][0] = (gm < M && gk < K) ? X[static_cast<size_t>(gm) * ldx + gk] : 0.0f;
        }
      }
    }
#pragma unroll
    for (int i = 0; i < kWPer; ++i) {
      const int gn = n0 + r0 + i * kRowStep;
      if (kV == 4) {
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
    for (int i = 0; i < kAPer; ++i) {
      if (kSrc == 2) {
        As[(tid & 31) * kPad + kk_r + 8 * i] = ra[i][0];
      } else {
#pragma unroll
        for (int v = 0; v < kAV; ++v) As[(r0 + i * kRowStep) * kPad + c0 + v] = ra[i][v];
      }
    }
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

bool aligned16(const void* p) { return reinterpret_cast<uintptr_t>(p) % 16 == 0; }

// rs.ring != nullptr: X comes from the ring (K % 4 == 0 and W 16-byte aligned required).
int launch_linear(const float* X, int ldx, const RingSrc& rs, int M, int K, const float* W, const float* b, int N,
                  float* Y, int ldy, bool elu, hipStream_t stream) {
  const dim3 grid((M + kRows - 1) / kRows, (N + kCols - 1) / kCols);
  const bool wvec = K % 4 == 0 && aligned16(W);
#define SPI_LINEAR(E, S) linear_kernel<E, S><<<grid, 256, 0, stream>>>(X, ldx, rs, M, K, W, b, N, Y, ldy)
  if (rs.ring) {
    if (!wvec) return mfail(SPI_ERR_INVALID_ARG, "ring-gathered layer needs K % 4 == 0 and 16-byte aligned weights");
    if (elu) SPI_LINEAR(true, 2); else SPI_LINEAR(false, 2);
  } else if (wvec && ldx % 4 == 0 && aligned16(X)) {
    if (elu) SPI_LINEAR(true, 1); else SPI_LINEAR(false, 1);
  } else {
    if (elu) SPI_LINEAR(true, 0); else SPI_LINEAR(false, 0);
  }
#undef SPI_LINEAR
  const hipError_t e = hipGetLastError();
  if (e == hipSuccess) return SPI_OK;
  char buf[256];
  std::snprintf(buf, sizeof(buf), "linear_kernel: %s", hipGetErrorString(e));
  spi_set_error(buf);
  return SPI_ERR_HIP;
}

long long param_floats(int n_layers, const int* dims) {
  if (n_layers < 1 || n_layers > kMaxLayers || !dims) return -1;
  long long n = 0;
  for (int l = 0; l < n_layers; ++l) {
    if (dims[l] < 1 || dims[l + 1] < 1) return -1;
    n += static_cast<long long>(dims[l + 1]) * dims[l] + dims[l + 1];
  }
  return n;
}

int widest_hidden(int n_layers, const int* dims) {
  int w = 0;
  for (int l = 1; l < n_layers; ++l) w = dims[l] > w ? dims[l] : w;
  return w;
}

int mlp_layers(const float* x_dev, const RingSrc& rs, int M, int n_layers, const int* dims, const float* params_dev,
               float* work_dev, long long work_floats, float* y_dev, hipStream_t stream) {
  if (M < 0 || param_floats(n_layers, dims) < 0 || !params_dev || !y_dev || (n_layers > 1 && !work_dev))
    return mfail(SPI_ERR_INVALID_ARG, "spi_mlp_forward: bad arguments");
  const int widest = widest_hidden(n_layers, dims);
  if (work_floats < 2LL * M * widest) return mfail(SPI_ERR_INVALID_ARG, "spi_mlp_forward: work buffer too small");
  for (int l = 0; l <= n_layers; ++l)
    if (static_cast<long long>(M) * dims[l] > 0x7fffffffLL)
      return mfail(SPI_ERR_TOO_LARGE, "spi_mlp_forward: activations exceed 2^31 elements");
  if (M == 0) return SPI_OK;
  const float* in = x_dev;
  const float* p = params_dev;
  for (int l = 0; l < n_layers; ++l) {
    const int K = dims[l], N = dims[l + 1];
    const bool last = l == n_layers - 1;
    float* out = last ? y_dev : work_dev + static_cast<size_t>(l & 1) * M * widest;
    const RingSrc none{nullptr, 0, 0, 0, 0, nullptr, 0.0f};
    const int rc = launch_linear(in, K, l == 0 ? rs : none, M, K, p, p + static_cast<size_t>(N) * K, N, out, N,
                                 !last, stream);
    if (rc != SPI_OK) return rc;
    p += static_cast<size_t>(N) * K + N;
    in = out;
  }
  return SPI_OK;
}

}  // namespace

extern "C" {

long long spi_mlp_param_floats(int n_layers, const int* dims) { return param_floats(n_layers, dims); }

long long spi_mlp_work_floats(int M, int n_layers, const int* dims) {
  if (M < 0 || param_floats(n_layers, dims) < 0) return -1;
  return 2LL * M * widest_hidden(n_layers, dims);
}

int spi_mlp_forward(const float* x_dev, int M, int n_layers, const int* dims, const float* params_dev, float* work_dev,
                    long long work_floats, float* y_dev, hipStream_t stream) {
  if (!x_dev) return mfail(SPI_ERR_INVALID_ARG, "spi_mlp_forward: null input");
  const RingSrc none{nullptr, 0, 0, 0, 0, nullptr, 0.0f};
  return mlp_layers(x_dev, none, M, n_layers, dims, params_dev, work_dev, work_floats, y_dev, stream);
}

int spi_mlp_forward_ring(const float* ring_dev, int ld, int n_slots, int n_terms, int head, const int* kmap_dev,
                         float clip, int M, int n_layers, const int* dims, const float* params_dev, float* work_dev,
                         long long work_floats, float* y_dev, hipStream_t stream) {
  if (!ring_dev || !kmap_dev || ld < M || n_slots < 1 || n_terms < 1 || head < 0 || head >= n_slots ||
      !(clip > 0.0f) || reinterpret_cast<uintptr_t>(kmap_dev) % 8 != 0)
    return mfail(SPI_ERR_INVALID_ARG, "spi_mlp_forward_ring: bad ring arguments");
  const RingSrc rs{ring_dev, ld, n_slots, n_terms, head, reinterpret_cast<const int2*>(kmap_dev), clip};
  return mlp_layers(nullptr, rs, M, n_layers, dims, params_dev, work_dev, work_floats, y_dev, stream);
}

}  // extern "C"
