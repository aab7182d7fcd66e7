// Policy network inference (gfx950): the actor MLP of the locomotion policy (Linear layers with
// ELU between them, spigym/agents/modules/modules.py:47-63; hidden sizes 512-256-128 from
// config/algo/ppo.yaml) as FP32 MFMA (v_mfma_f32_32x32x2f32) tiles, one launch per layer with
// bias and ELU in the epilogue.  linear_kernel: a workgroup of 4 wavefronts owns a 32-row x
// 128-column output tile (many workgroups for small batches); rows of X and the weight rows are
// staged through LDS in K chunks of 64 (even columns, then odd ones, 65 words per row, so the
// operand reads of a wavefront hit 64 distinct banks).  linear_db_kernel: 128 x 128 (a 64 x 64
// quadrant per wavefront, K chunks of 32 double-buffered in LDS) for batches of 8192 rows and
// more.  Strict IEEE float32 (no fast-math).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <mutex>

#include "spi_types.h"

// Workspace of spi_mlp_forward_noalloc: the ping-pong hidden activations live in this
// file-scope device array (one copy per GPU, held until the process ends) instead of a buffer
// the caller allocates.  In host code the name stands for the symbol, not for a device
// address: the address of the current device's copy comes from hipGetSymbolAddress.
static __device__ __align__(256) float scratch[SPI_MLP_SCRATCH_FLOATS];

namespace {

typedef float floatx16 __attribute__((ext_vector_type(16)));

constexpr int kCols = 128;   // output columns per workgroup (4 wavefronts x MFMA N)
constexpr int kKc = 64;      // K chunk staged per pass
constexpr int kPad = kKc + 1;
constexpr int kMaxLayers = 8;
constexpr int kMaxRingK = 1024;  // widest ring-gathered input
constexpr int kBigM = 8192;      // batches from here on use linear_db_kernel (N >= 256)

// LDS slot of tile element (row, k): the even columns of a row first, then the odd ones, so
// the MFMA operand reads of a wavefront (lane -> row l % 32, column kk + l / 32) cover 64
// distinct banks: row * 65 + 32 (l / 32) + kk / 2.
__device__ __forceinline__ int lds_at(int row, int k) { return row * kPad + ((k & 1) << 5) + (k >> 1); }

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

// LDS slot of element (row, k) of a tile staged in K chunks of kC: the even columns of a row
// first, then the odd ones, rows padded to kC + 1 words (lds_at is the kC = 64 case).
template <int kC> __device__ __forceinline__ int lds_at_c(int row, int k) {
  return row * (kC + 1) + (k & 1) * (kC / 2) + (k >> 1);
}

// Y[m][n] = act(sum_k X[m][k] W[n][k] + b[n]); W row-major [N][K] (torch Linear layout).
// MFMA operands: lane l supplies A[l % 32][k + l / 32] = X-tile element and
// B[k + l / 32][l % 32] = W[n0 + l % 32][k + l / 32]; accumulator r of lane l holds
// D[8 (r / 4) + 4 (l / 32) + r % 4][l % 32].  The next K chunk is fetched from global memory
// into registers while the MFMAs consume the current one from LDS.
// kSrc 0: X row-major, scalar loads; 1: X and W rows 16-byte aligned with K % 4 == 0, float4
// loads; 2: X gathered from a RingSrc (32 consecutive rows of one column per load), W as in 1;
// 3: as 2 with float4 loads along the rows.
// Split K (kS groups of 4 wavefronts, 256 kS threads): group g multiplies chunks g, g + kS,
// g + 2 kS, ... of K (chunks of 64, of 32 for kS = 4) into its own accumulators from its own
// LDS tiles; at the end groups 1 .. kS-1 hand theirs to group 0 through LDS, which adds them in
// group order (a fixed order: the result does not depend on timing) and runs the epilogue.
// For small batches, whose 32 x 128 tiles leave SIMDs idle or alone: kS wavefronts per SIMD
// hide each other's load latency and each K loop is kS times shorter.  kS = 1 is the plain
// tile loop.
// kFuse: the layer's outputs (N <= 128: one column tile) never leave the workgroup; after the
// epilogue (bias, ELU) they go to LDS and the workgroup computes the next, last layer from them
// (W2 [N2][N], b2, N2 <= 32, no activation: for each row and output, multiply-add chains over
// the two halves of the N inputs in index order, then the two halves and the bias in that
// order), writing Y2 [M][N2].
template <bool kElu, int kSrc, int kS, bool kFuse = false>
__global__ __launch_bounds__(256 * kS) void linear_kernel(const float* __restrict__ X, int ldx, RingSrc rs, int M, int K,
                                                          const float* __restrict__ W, const float* __restrict__ b, int N,
                                                          float* __restrict__ Y, int ldy,
                                                          const float* __restrict__ W2 = nullptr,
                                                          const float* __restrict__ b2 = nullptr, int N2 = 0,
                                                          float* __restrict__ Y2 = nullptr) {
  constexpr int kRows = 32;
  constexpr int kKs = kS == 4 ? 32 : kKc;              // K chunk of one group
  constexpr int kPs = kKs + 1;
  constexpr int kGroupFloats = (kRows + kCols) * kPs;
  static_assert(kS == 1 || (kS - 1) * 16 * 256 <= kS * kGroupFloats, "reduction buffer fits the tiles");
  static_assert(kRows * (kCols + 1) + 2 * kRows * 32 <= kS * kGroupFloats, "a fused tail's tile fits the tiles");
  __shared__ float smem[kS * kGroupFloats];
  __shared__ int Kr[kSrc >= 2 ? kMaxRingK : 1];       // ring line of each input column
  const int m0 = blockIdx.x * kRows, n0 = blockIdx.y * kCols;
  const int grp = threadIdx.x >> 8, tid = threadIdx.x & 255, wave = tid >> 6, lane = tid & 63;
  float* As = smem + grp * kGroupFloats;
  float* Ws = As + kRows * kPs;
  const int col = lane & 31, half = lane >> 5;
  const bool live = n0 + wave * 32 < N;  // uniform per wavefront
  if (kSrc >= 2) {
    // any table entry is folded into the ring: slots back in [0, n_slots), term in [0, n_terms)
    for (int k = threadIdx.x; k < K; k += 256 * kS) {
      const int2 km = rs.kmap[k];
      const int back = min(max(km.x, 0), rs.n_slots - 1);
      const int slot = rs.head >= back ? rs.head - back : rs.head - back + rs.n_slots;
      Kr[k] = slot * rs.n_terms + min(max(km.y, 0), rs.n_terms - 1);
    }
    __syncthreads();
  }
  floatx16 acc;
#pragma unroll
  for (int r = 0; r < 16; ++r) acc[r] = 0.0f;
  constexpr int kV = kSrc == 0 ? 1 : 4;                // floats per W (and row-major X) load
  constexpr int kTPR = kKs / kV;                       // threads per tile row
  constexpr int kRowStep = 256 / kTPR;                 // rows covered by one pass of the group
  constexpr int kAV = kSrc == 2 ? 1 : kV;
  // ring gather: thread -> row (tid % 32), columns kk_r + i * 8; float4 ring gather (kSrc 3):
  // thread -> rows 4 (tid % 8) .. + 3, columns kk4 + i * 32
  constexpr int kGStep = 256 / kRows;
  constexpr int kG4 = kRows / 4, kG4Step = 256 / kG4;
  constexpr int kAPer = kSrc == 2 ? kRows * kKs / 256 : (kSrc == 3 ? kRows * kKs / 1024 : kRows / kRowStep);
  constexpr int kWPer = kCols / kRowStep;
  static_assert(kAPer >= 1 && kWPer >= 1, "tile staging covers the chunk");
  const int r0 = tid / kTPR, c0 = (tid % kTPR) * kV;
  const int gm_r = m0 + tid % kRows, kk_r = tid / kRows;
  const int m4 = (tid % kG4) * 4, kk4 = tid / kG4;
  float ra[kAPer][kAV], rw[kWPer][kV];
  auto fetch = [&](int k0) {
    const int gk = k0 + c0;
#pragma unroll
    for (int i = 0; i < kAPer; ++i) {
      if (kSrc == 2) {
        const int k = k0 + kk_r + kGStep * i;
        float v = 0.0f;
        if (gm_r < M && k < K) {
          v = rs.ring[static_cast<size_t>(Kr[k]) * rs.ld + gm_r];
          v = fminf(fmaxf(v, -rs.clip), rs.clip);
        }
        ra[i][0] = v;
      } else if (kSrc == 3) {
        const int k = k0 + kk4 + kG4Step * i;
        float4 v = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        if (m0 + m4 < M && k < K)
          v = *reinterpret_cast<const float4*>(rs.ring + static_cast<size_t>(Kr[k]) * rs.ld + m0 + m4);
        ra[i][0] = fminf(fmaxf(v.x, -rs.clip), rs.clip);
        ra[i][kAV > 1 ? 1 : 0] = fminf(fmaxf(v.y, -rs.clip), rs.clip);
        ra[i][kAV > 2 ? 2 : 0] = fminf(fmaxf(v.z, -rs.clip), rs.clip);
        ra[i][kAV > 3 ? 3 : 0] = fminf(fmaxf(v.w, -rs.clip), rs.clip);
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
  // every group runs the same number of passes (chunks past K load zeros), so the barriers match
  const int n_pass = (K + kKs * kS - 1) / (kKs * kS);
  fetch(grp * kKs);
  for (int it = 0; it < n_pass; ++it) {
#pragma unroll
    for (int i = 0; i < kAPer; ++i) {
      if (kSrc == 2) {
        As[lds_at_c<kKs>(tid % kRows, kk_r + kGStep * i)] = ra[i][0];
      } else if (kSrc == 3) {
#pragma unroll
        for (int v = 0; v < 4; ++v) As[lds_at_c<kKs>(m4 + v, kk4 + kG4Step * i)] = ra[i][v];
      } else {
#pragma unroll
        for (int v = 0; v < kAV; ++v) As[lds_at_c<kKs>(r0 + i * kRowStep, c0 + v)] = ra[i][v];
      }
    }
#pragma unroll
    for (int i = 0; i < kWPer; ++i)
#pragma unroll
      for (int v = 0; v < kV; ++v) Ws[lds_at_c<kKs>(r0 + i * kRowStep, c0 + v)] = rw[i][v];
    __syncthreads();
    if (it + 1 < n_pass) fetch(((it + 1) * kS + grp) * kKs);
    if (live) {
      const float* ap = As + col * kPs + (kKs / 2) * half;
      const float* bp = Ws + (wave * 32 + col) * kPs + (kKs / 2) * half;
#pragma unroll
      for (int kk = 0; kk < kKs; kk += 2)
        acc = __builtin_amdgcn_mfma_f32_32x32x2f32(ap[kk / 2], bp[kk / 2], acc, 0, 0, 0);
    }
    __syncthreads();
  }
  if (kS > 1) {
    // the loop ended on a barrier: the tiles are free; groups 1 .. kS-1 -> [kS-1][16][256]
    float* red = smem;
    if (grp > 0 && live) {
#pragma unroll
      for (int r = 0; r < 16; ++r) red[((grp - 1) * 16 + r) * 256 + tid] = acc[r];
    }
    __syncthreads();
    if (grp > 0 && !kFuse) return;
    if (grp == 0 && live) {
      for (int g = 1; g < kS; ++g)
#pragma unroll
        for (int r = 0; r < 16; ++r) acc[r] += red[((g - 1) * 16 + r) * 256 + tid];
    }
  }
  if (kFuse) {
    // every thread of every group reaches both barriers; group 0 holds the finished tile
    constexpr int kHs = kCols + 1;
    float* h = smem;                                   // [32][kHs] after the reduction reads
    __syncthreads();
    if (grp == 0) {
      const int gn = wave * 32 + col;
      const float bias = gn < N ? b[gn] : 0.0f;
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const int row = 8 * (r >> 2) + 4 * half + (r & 3);
        float v = 0.0f;
        if (live && gn < N) {
          v = acc[r] + bias;
          if (kElu) v = v > 0.0f ? v : expm1f(v);
        }
        h[row * kHs + gn] = v;
      }
    }
    __syncthreads();
    // every thread: partial sums over the two halves of the N inputs, (half, row, output) per
    // thread, then half 0 + half 1 + bias in that order
    float* part = h + kRows * kHs;                     // [2][32][32]
    const int nh = (N + 1) / 2;
    for (int u = threadIdx.x; u < 2 * kRows * N2; u += blockDim.x) {
      const int q = u / (kRows * N2), rem = u - q * (kRows * N2), row = rem / N2, j = rem - row * N2;
      const int k0 = q * nh, k1 = min(N, k0 + nh);
      const float* hr = h + row * kHs;
      const float* w2 = W2 + static_cast<size_t>(j) * N;
      float sum = 0.0f;
#pragma unroll 8
      for (int k = k0; k < k1; ++k) sum = fmaf(hr[k], w2[k], sum);
      part[(q * kRows + row) * 32 + j] = sum;
    }
    __syncthreads();
    for (int u = threadIdx.x; u < kRows * N2; u += blockDim.x) {
      const int row = u / N2, j = u - row * N2, gm = m0 + row;
      if (gm < M) Y2[static_cast<size_t>(gm) * N2 + j] = (part[row * 32 + j] + part[(kRows + row) * 32 + j]) + b2[j];
    }
    return;
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

// Groups of 4 wavefronts per 32 x 128 tile: split K while the tiles alone would leave SIMDs
// idle or with one wavefront (kS needs K >= 128: a chunk per group).
int split_for(int M, int K, int N) {
  const long long tiles = static_cast<long long>((M + 31) / 32) * ((N + kCols - 1) / kCols);
  if (K < 128 || tiles >= 512) return 1;
  return tiles <= 128 ? 4 : 2;
}

// Large batches: a workgroup of 4 wavefronts owns a 128 x 128 output tile, each wavefront a
// 64 x 64 quadrant (2 x 2 MFMA blocks, 4 accumulators), so every LDS operand read feeds two
// MFMAs.  K is staged in chunks of 32 through two LDS buffers (rows of 32 floats + 2 of padding,
// 68 KB in all, so two workgroups share a CU): while the wavefronts multiply chunk c out of one
// buffer, chunk c + 1 goes from registers into the other and chunk c + 2 is fetched into
// registers, with one barrier per chunk.  In MFMA step s of a chunk the lanes of half h supply
// column 16 h + s of both operands, read as float2 pairs (s, s + 1); in the last, ragged chunk
// only the steps that reach a column below K run.  kSrc 0: X row-major (rows 16-byte aligned,
// K % 4 == 0), float4 loads along the rows; 1: X gathered from a RingSrc whose lines are
// 16-byte aligned with ld % 4 == 0, float4 loads of 4 rows of one column (lane l of a wavefront
// covers rows 4 (l % 8) .. + 3 and column l / 8, so that the 64 lanes' LDS stores of one row
// offset fall in 64 distinct banks); 2: any RingSrc, one row's four consecutive columns per
// thread in scalar loads.  The accumulation order over K depends only on the tile layout, not
// on the source, so the dense and the ring input give the same bits.
constexpr int kDbRows = 128, kDbCols = 128, kDbK = 32, kDbStride = kDbK + 2, kDbTile = 128 * kDbStride;
template <bool kElu, int kSrc>
__global__ __launch_bounds__(256) void linear_db_kernel(const float* __restrict__ X, int ldx, RingSrc rs, int M,
                                                        int K, const float* __restrict__ W,
                                                        const float* __restrict__ b, int N, float* __restrict__ Y,
                                                        int ldy) {
  __shared__ __attribute__((aligned(16))) float lds[2 * 2 * kDbTile];   // [buffer][A | W][row][34]
  constexpr bool kRing = kSrc != 0;
  __shared__ int Kr[kRing ? kMaxRingK : 1];
  const int m0 = blockIdx.x * kDbRows, n0 = blockIdx.y * kDbCols;
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int col = lane & 31, half = lane >> 5;
  const int wr = (wave >> 1) * 64, wc = (wave & 1) * 64;   // this wavefront's quadrant
  const bool live = n0 + wc < N;                            // uniform per wavefront
  if (kRing) {
    for (int k = tid; k < K; k += 256) {
      const int2 km = rs.kmap[k];
      const int back = min(max(km.x, 0), rs.n_slots - 1);
      const int slot = rs.head >= back ? rs.head - back : rs.head - back + rs.n_slots;
      Kr[k] = slot * rs.n_terms + min(max(km.y, 0), rs.n_terms - 1);
    }
    __syncthreads();
  }
  floatx16 acc[2][2];
#pragma unroll
  for (int p = 0; p < 2; ++p)
#pragma unroll
    for (int q = 0; q < 2; ++q)
#pragma unroll
      for (int r = 0; r < 16; ++r) acc[p][q][r] = 0.0f;
  // W (and dense X) staging: thread -> rows (tid >> 3) + 32 i, columns 4 (tid & 7) .. + 3
  const int sr = tid >> 3, sc = (tid & 7) * 4;
  // float4 ring staging: thread -> rows 32 wave + 4 (lane % 8) .. + 3, columns lane / 8 + 8 i
  const int fr = 32 * wave + 4 * (lane & 7), fk = lane >> 3;
  // scalar ring staging: thread -> row tid & 127, column quads (tid >> 7) + 2 i
  const int gr = tid & 127, gq = tid >> 7;
  float4 ra[4], rw[4];
  auto fetch = [&](int k0) {
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      if (kSrc == 1) {
        // rows up to M rounded to 4 exist in every line (ld >= that); rows >= M read as 0
        const int gm = m0 + fr, k = k0 + fk + 8 * i;
        float4 v = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        if (gm < M && k < K) v = *reinterpret_cast<const float4*>(rs.ring + static_cast<size_t>(Kr[k]) * rs.ld + gm);
        ra[i].x = fminf(fmaxf(v.x, -rs.clip), rs.clip);
        ra[i].y = gm + 1 < M ? fminf(fmaxf(v.y, -rs.clip), rs.clip) : 0.0f;
        ra[i].z = gm + 2 < M ? fminf(fmaxf(v.z, -rs.clip), rs.clip) : 0.0f;
        ra[i].w = gm + 3 < M ? fminf(fmaxf(v.w, -rs.clip), rs.clip) : 0.0f;
      } else if (kSrc == 2) {
        const int gm = m0 + gr, kq = k0 + 4 * (gq + 2 * i);
        float v[4];
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const int k = kq + e;
          v[e] = 0.0f;
          if (gm < M && k < K)
            v[e] = fminf(fmaxf(rs.ring[static_cast<size_t>(Kr[k]) * rs.ld + gm], -rs.clip), rs.clip);
        }
        ra[i] = make_float4(v[0], v[1], v[2], v[3]);
      } else {
        const int gm = m0 + sr + 32 * i, gk = k0 + sc;
        ra[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        if (gm < M && gk < K) ra[i] = *reinterpret_cast<const float4*>(X + static_cast<size_t>(gm) * ldx + gk);
      }
      const int gn = n0 + sr + 32 * i, gk = k0 + sc;
      rw[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
      if (gn < N && gk < K) rw[i] = *reinterpret_cast<const float4*>(W + static_cast<size_t>(gn) * K + gk);
    }
  };
  auto stage = [&](int buf) {
    float* A = lds + buf * 2 * kDbTile;
    float* B = A + kDbTile;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      if (kSrc == 1) {
        float* pa = A + fr * kDbStride + fk + 8 * i;
        pa[0] = ra[i].x;
        pa[kDbStride] = ra[i].y;
        pa[2 * kDbStride] = ra[i].z;
        pa[3 * kDbStride] = ra[i].w;
      } else {
        float2* pa = reinterpret_cast<float2*>(kSrc == 2 ? A + gr * kDbStride + 4 * (gq + 2 * i)
                                                         : A + (sr + 32 * i) * kDbStride + sc);
        pa[0] = make_float2(ra[i].x, ra[i].y);
        pa[1] = make_float2(ra[i].z, ra[i].w);
      }
      float2* pb = reinterpret_cast<float2*>(B + (sr + 32 * i) * kDbStride + sc);
      pb[0] = make_float2(rw[i].x, rw[i].y);
      pb[1] = make_float2(rw[i].z, rw[i].w);
    }
  };
  auto compute = [&](int buf, int npair) {
    const float* A = lds + buf * 2 * kDbTile;
    const float* ap = A + (wr + col) * kDbStride + 16 * half;
    const float* bp = A + kDbTile + (wc + col) * kDbStride + 16 * half;
#pragma unroll
    for (int jp = 0; jp < kDbK / 4; ++jp) {          // steps s = 2 jp, 2 jp + 1
      if (jp < npair) {
        const float2 a0 = *reinterpret_cast<const float2*>(ap + 2 * jp);
        const float2 a1 = *reinterpret_cast<const float2*>(ap + 32 * kDbStride + 2 * jp);
        const float2 b0 = *reinterpret_cast<const float2*>(bp + 2 * jp);
        const float2 b1 = *reinterpret_cast<const float2*>(bp + 32 * kDbStride + 2 * jp);
        acc[0][0] = __builtin_amdgcn_mfma_f32_32x32x2f32(a0.x, b0.x, acc[0][0], 0, 0, 0);
        acc[0][1] = __builtin_amdgcn_mfma_f32_32x32x2f32(a0.x, b1.x, acc[0][1], 0, 0, 0);
        acc[1][0] = __builtin_amdgcn_mfma_f32_32x32x2f32(a1.x, b0.x, acc[1][0], 0, 0, 0);
        acc[1][1] = __builtin_amdgcn_mfma_f32_32x32x2f32(a1.x, b1.x, acc[1][1], 0, 0, 0);
        acc[0][0] = __builtin_amdgcn_mfma_f32_32x32x2f32(a0.y, b0.y, acc[0][0], 0, 0, 0);
        acc[0][1] = __builtin_amdgcn_mfma_f32_32x32x2f32(a0.y, b1.y, acc[0][1], 0, 0, 0);
        acc[1][0] = __builtin_amdgcn_mfma_f32_32x32x2f32(a1.y, b0.y, acc[1][0], 0, 0, 0);
        acc[1][1] = __builtin_amdgcn_mfma_f32_32x32x2f32(a1.y, b1.y, acc[1][1], 0, 0, 0);
      }
    }
  };
  const int nch = (K + kDbK - 1) / kDbK;
  fetch(0);
  stage(0);
  if (nch > 1) fetch(kDbK);
  __syncthreads();
  for (int c = 0; c < nch; ++c) {
    // columns 16 + s of half 1 lie past K whenever column s does
    const int klen = min(kDbK, K - c * kDbK);
    const int npair = klen >= kDbK / 2 ? kDbK / 4 : (klen + 1) / 2;
    if (live) compute(c & 1, npair);
    if (c + 1 < nch) stage((c + 1) & 1);
    if (c + 2 < nch) fetch((c + 2) * kDbK);
    __syncthreads();
  }
  if (!live) return;
#pragma unroll
  for (int q = 0; q < 2; ++q) {
    const int gn = n0 + wc + 32 * q + col;
    if (gn >= N) continue;
    const float bias = b[gn];
#pragma unroll
    for (int p = 0; p < 2; ++p)
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const int gm = m0 + wr + 32 * p + 8 * (r >> 2) + 4 * half + (r & 3);
        if (gm < M) {
          float v = acc[p][q][r] + bias;
          if (kElu) v = v > 0.0f ? v : expm1f(v);
          Y[static_cast<size_t>(gm) * ldy + gn] = v;
        }
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
  const bool wvec = K % 4 == 0 && aligned16(W);
  const bool xvec = rs.ring ? true : (wvec && ldx % 4 == 0 && aligned16(X));
  if (rs.ring && !wvec)
    return mfail(SPI_ERR_INVALID_ARG, "ring-gathered layer needs K % 4 == 0 and 16-byte aligned weights");
  // rows of the ring read as float4: needs ld % 4 == 0 (then rows up to round4(M) <= ld exist)
  const bool ring4 = rs.ring && rs.ld % 4 == 0 && aligned16(rs.ring);
  if (M >= kBigM && N >= 256 && wvec && xvec) {
    const dim3 grid((M + kDbRows - 1) / kDbRows, (N + kDbCols - 1) / kDbCols);
#define SPI_DB(E, S) linear_db_kernel<E, S><<<grid, 256, 0, stream>>>(X, ldx, rs, M, K, W, b, N, Y, ldy)
    if (ring4) {
      if (elu) SPI_DB(true, 1); else SPI_DB(false, 1);
    } else if (rs.ring) {
      if (elu) SPI_DB(true, 2); else SPI_DB(false, 2);
    } else {
      if (elu) SPI_DB(true, 0); else SPI_DB(false, 0);
    }
#undef SPI_DB
  } else {
    const dim3 grid((M + 31) / 32, (N + kCols - 1) / kCols);
    const int split = split_for(M, K, N);
#define SPI_LINEAR_S(E, S, P) linear_kernel<E, S, P><<<grid, 256 * P, 0, stream>>>(X, ldx, rs, M, K, W, b, N, Y, ldy)
#define SPI_LINEAR(E, S)            \
  do {                              \
    if (split == 4)                 \
      SPI_LINEAR_S(E, S, 4);        \
    else if (split == 2)            \
      SPI_LINEAR_S(E, S, 2);        \
    else                            \
      SPI_LINEAR_S(E, S, 1);        \
  } while (0)
    if (ring4) {
      if (elu) SPI_LINEAR(true, 3); else SPI_LINEAR(false, 3);
    } else if (rs.ring) {
      if (elu) SPI_LINEAR(true, 2); else SPI_LINEAR(false, 2);
    } else if (xvec) {
      if (elu) SPI_LINEAR(true, 1); else SPI_LINEAR(false, 1);
    } else {
      if (elu) SPI_LINEAR(true, 0); else SPI_LINEAR(false, 0);
    }
#undef SPI_LINEAR
#undef SPI_LINEAR_S
  }
  const hipError_t e = hipGetLastError();
  if (e == hipSuccess) return SPI_OK;
  char buf[256];
  std::snprintf(buf, sizeof(buf), "linear_kernel: %s", hipGetErrorString(e));
  spi_set_error(buf);
  return SPI_ERR_HIP;
}

// The last two layers in one launch (linear_kernel with kFuse) when the second-to-last has at
// most 128 outputs (one column tile) and the last at most 32, and the input is a dense buffer
// with 16-byte rows; returns false (nothing launched) when that does not hold.
bool fused_tail(const float* X, int ldx, int M, int K, const float* W, const float* b, int N, const float* W2,
                const float* b2, int N2, float* Y2, hipStream_t stream, int* rc) {
  if (N > kCols || N2 > 32 || K % 4 != 0 || ldx % 4 != 0 || !aligned16(W) || !aligned16(X)) return false;
  const RingSrc none{nullptr, 0, 0, 0, 0, nullptr, 0.0f};
  const dim3 grid((M + 31) / 32, 1);
  const int split = split_for(M, K, N);
  if (split == 4)
    linear_kernel<true, 1, 4, true><<<grid, 1024, 0, stream>>>(X, ldx, none, M, K, W, b, N, nullptr, 0, W2, b2, N2, Y2);
  else if (split == 2)
    linear_kernel<true, 1, 2, true><<<grid, 512, 0, stream>>>(X, ldx, none, M, K, W, b, N, nullptr, 0, W2, b2, N2, Y2);
  else
    linear_kernel<true, 1, 1, true><<<grid, 256, 0, stream>>>(X, ldx, none, M, K, W, b, N, nullptr, 0, W2, b2, N2, Y2);
  const hipError_t e = hipGetLastError();
  *rc = e == hipSuccess ? SPI_OK : mfail(SPI_ERR_HIP, hipGetErrorString(e));
  return true;
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
    if (l > 0 && l == n_layers - 2) {                 // hidden layer + output layer in one launch
      const float* p2 = p + static_cast<size_t>(N) * K + N;
      const int N2 = dims[l + 2];
      int rc = SPI_OK;
      if (fused_tail(in, K, M, K, p, p + static_cast<size_t>(N) * K, N, p2, p2 + static_cast<size_t>(N2) * N, N2,
                     y_dev, stream, &rc))
        return rc;
    }
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

int spi_mlp_forward_noalloc(const float* x_dev, int M, int n_layers, const int* dims, const float* params_dev,
                            float* y_dev, hipStream_t stream) {
  // device address of this GPU's scratch array, looked up once per device with it current
  static std::mutex mu;
  static float* addr[64] = {};
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess || dev < 0 || dev >= 64)
    return mfail(SPI_ERR_HIP, "spi_mlp_forward_noalloc: no current device");
  float* work = nullptr;
  {
    std::lock_guard<std::mutex> lock(mu);
    if (!addr[dev]) {
      void* ptr = nullptr;
      e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
      if (e != hipSuccess || !ptr) {
        char buf[256];
        std::snprintf(buf, sizeof(buf), "spi_mlp_forward_noalloc: hipGetSymbolAddress: %s", hipGetErrorString(e));
        return mfail(SPI_ERR_HIP, buf);
      }
      addr[dev] = static_cast<float*>(ptr);
    }
    work = addr[dev];
  }
  if (!x_dev) return mfail(SPI_ERR_INVALID_ARG, "spi_mlp_forward_noalloc: null input");
  if (M < 0 || param_floats(n_layers, dims) < 0)
    return mfail(SPI_ERR_INVALID_ARG, "spi_mlp_forward_noalloc: bad arguments");
  if (2LL * M * widest_hidden(n_layers, dims) > SPI_MLP_SCRATCH_FLOATS)
    return mfail(SPI_ERR_TOO_LARGE, "spi_mlp_forward_noalloc: 2 * M * (widest hidden layer) exceeds "
                                    "SPI_MLP_SCRATCH_FLOATS; use spi_mlp_forward with a caller workspace");
  const RingSrc none{nullptr, 0, 0, 0, 0, nullptr, 0.0f};
  return mlp_layers(x_dev, none, M, n_layers, dims, params_dev, work, SPI_MLP_SCRATCH_FLOATS, y_dev, stream);
}

int spi_mlp_forward_ring(const float* ring_dev, int ld, int n_slots, int n_terms, int head, const int* kmap_dev,
                         float clip, int M, int n_layers, const int* dims, const float* params_dev, float* work_dev,
                         long long work_floats, float* y_dev, hipStream_t stream) {
  if (!ring_dev || !kmap_dev || ld < M || n_slots < 1 || n_terms < 1 || head < 0 || head >= n_slots ||
      !(clip > 0.0f) || reinterpret_cast<uintptr_t>(kmap_dev) % 8 != 0 || !dims || dims[0] > kMaxRingK)
    return mfail(SPI_ERR_INVALID_ARG, "spi_mlp_forward_ring: bad ring arguments (input width <= 1024)");
  const RingSrc rs{ring_dev, ld, n_slots, n_terms, head, reinterpret_cast<const int2*>(kmap_dev), clip};
  return mlp_layers(nullptr, rs, M, n_layers, dims, params_dev, work_dev, work_floats, y_dev, stream);
}

}  // extern "C"
