// Host-CPU baseline with the GPU kernel's own algorithm: spi_dyn.h (PD law + motor model,
// articulated-body dynamics specialised to the Go2 tree, penalty contact, semi-implicit Euler
// with torque held over the sub-intervals) built for x86 and run over OpenMP threads, one
// (candidate, segment) rollout per loop iteration, scored like rollout_cost_kernel.  bench.py
// times it beside the GPU; nothing in the library runs it.
#include <cstring>

#include <omp.h>

#include "spi_dyn.h"

using namespace spi;

namespace {
LaneParams lane_params(const float* par, int ld, int i) {
  LaneParams lp;
  lp.base = make_base_inertia(par[0 * ld + i], par[1 * ld + i], par[2 * ld + i], par[3 * ld + i], par[4 * ld + i],
                              par[5 * ld + i], par[6 * ld + i], par[7 * ld + i], par[8 * ld + i], par[9 * ld + i]);
  for (int l = 0; l < 4; ++l) {
    lp.kp_s[l] = par[(10 + l) * ld + i];
    lp.kd_s[l] = par[(14 + l) * ld + i];
    lp.fric[l] = par[(18 + l) * ld + i];
  }
  for (int g = 0; g < 3; ++g) lp.motor[g] = par[(22 + g) * ld + i];
  return lp;
}

// exponent-bit test behind an empty asm: this file is built with -ffinite-math-only, under which
// the compiler may fold a plain test of a float to "finite"
bool finite_f(float x) {
  unsigned u;
  std::memcpy(&u, &x, 4);
  asm volatile("" : "+r"(u));
  return (u & 0x7f800000u) != 0x7f800000u;
}
// +inf as a bit pattern: this file is built with finite-math, where the INFINITY macro is undefined
float inf_bits() {
  const unsigned u = 0x7f800000u;
  float x;
  std::memcpy(&x, &u, 4);
  return x;
}
}  // namespace

// cost_out[s * n_cand + c]: the cost of candidate c (column of the SoA params [25][ld]) on
// segment s (records [n_seg][37 + 49 H]); +inf where the rollout diverged.  n_threads <= 0:
// OpenMP's default.  Returns the number of threads used.
extern "C" int spi_cpu_rollout_cost(const SpiSimArgs* a, const SpiCostArgs* cost, const float* par, int ld,
                                    int n_cand, const float* seg, int n_seg, int H, float* cost_out, int n_threads) {
  if (n_threads > 0) omp_set_num_threads(n_threads);
  const int stride = SPI_STATE_DIM + H * (12 + SPI_STATE_DIM);
  const long long total = static_cast<long long>(n_cand) * n_seg;
  int used = 1;
#pragma omp parallel
  {
#pragma omp single
    used = omp_get_num_threads();
#pragma omp for schedule(dynamic, 16)
    for (long long lane = 0; lane < total; ++lane) {
      const int s = static_cast<int>(lane / n_cand), c = static_cast<int>(lane % n_cand);
      const float* rec0 = seg + static_cast<size_t>(s) * stride;
      SpiSimArgs as = *a;
      if (cost->seg_gains) {
        for (int j = 0; j < 12; ++j) {
          as.kp[j] = cost->seg_gains[2 * s];
          as.kd[j] = cost->seg_gains[2 * s + 1];
        }
      }
      const LaneParams lp = lane_params(par, ld, c);
      State st;
      state_from_root(st, rec0, rec0 + 13, rec0 + 25, 1);
      float acc = 0.0f;
      bool ok = true;
      for (int t = 0; t < H; ++t) {
        const float* rec = rec0 + SPI_STATE_DIM + t * (12 + SPI_STATE_DIM);
        for (int d = 0; d < as.decimation; ++d) {
          float tau[12];
          control_torques(as, lp, rec, st, tau);
          physics_step(as, lp, st, tau, nullptr);
        }
        if (cost->mode == SPI_COST_TRACKING || t == H - 1) acc += step_cost(*cost, st, rec + 12);
        const float nrm = state_norm(st);
        if (!finite_f(nrm) || nrm >= as.state_limit) {
          ok = false;
          break;
        }
      }
      cost_out[lane] = (ok && finite_f(acc)) ? acc : inf_bits();
    }
  }
  return used;
}
