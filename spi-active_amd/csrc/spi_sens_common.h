// Launch helpers, parameter loading and argument checks shared by the tangent-rollout sources
// (spi_sens.hip: four lanes per tangent unit; spi_sens_lane.hip: one lane per tangent).
// Included after the forward-mode instantiation of spi_dyn.h (namespace spi_tan, DualT).
#pragma once

namespace {

int sfail(int code, const char* msg) {
  spi_set_error(msg);
  return code;
}
int slaunch(const char* what) {
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) return SPI_OK;
  char buf[256];
  std::snprintf(buf, sizeof(buf), "%s: %s", what, hipGetErrorString(e));
  spi_set_error(buf);
  return SPI_ERR_HIP;
}

constexpr int kSensThreads = 256;
constexpr int kFimDim = 25;   // y = root (13) + q (12)

// Parameters of candidate c with a unit tangent on SoA row `seed` (-1: none).
__device__ __forceinline__ LaneParams load_params_tan(const float* __restrict__ par, int ld, int c, int seed) {
  auto f = [&](int k) {
    DualT x(par[static_cast<size_t>(k) * ld + c]);
    x.d[0] = (k == seed) ? 1.0f : 0.0f;
    return x;
  };
  LaneParams lp;
  lp.base = make_base_inertia(f(0), f(1), f(2), f(3), f(4), f(5), f(6), f(7), f(8), f(9));
#pragma unroll
  for (int l = 0; l < 4; ++l) {
    lp.kp_s[l] = f(10 + l);
    lp.kd_s[l] = f(14 + l);
    lp.fric[l] = f(18 + l);
  }
#pragma unroll
  for (int g = 0; g < 3; ++g) lp.motor[g] = f(22 + g);
  return lp;
}

__device__ __forceinline__ void init_state(State& st, const float* s37) {
  DualT r[13], q[12], qd[12];
#pragma unroll
  for (int i = 0; i < 13; ++i) r[i] = DualT(s37[i]);
#pragma unroll
  for (int j = 0; j < 12; ++j) { q[j] = DualT(s37[13 + j]); qd[j] = DualT(s37[25 + j]); }
  state_from_root(st, r, q, qd, 1);
}

__device__ __forceinline__ bool state_ok(const State& st, float limit) {
  const float n = state_norm(st).v;
  return finite_bits(n) && n < limit;
}

bool tan_ok(const SpiTangentSpec* t) {
  if (!t || t->n_tan < 1 || t->n_tan > SPI_MAX_TANGENTS || t->n_tan > SPI_P_COUNT) return false;
  for (int k = 0; k < t->n_tan; ++k)
    if (t->param[k] < 0 || t->param[k] >= SPI_P_COUNT) return false;
  return true;
}
bool sim_ok(const SpiSimArgs* a) {
  return a && a->decimation >= 1 && a->substeps >= 1 && a->dt > 0.0f && a->state_limit > 0.0f;
}

}  // namespace
