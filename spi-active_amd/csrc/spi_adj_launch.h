// Launch constants and the argument check shared by the two reverse-mode sources:
// spi_adjoint.hip (four lanes per rollout) and spi_adjoint_lane.hip (one lane per rollout).
#pragma once
#include <cstdio>

#include "spi_types.h"

namespace spi_adj_launch {

constexpr int kAdjThreads = 256;
constexpr int kMaxDecimation = 8;
constexpr int kMaxSubsteps = 8;

inline int adjoint_args_check(const char* who, const SpiSimArgs* sim, const SpiCostArgs* cost, const float* params_dev,
                              int ld, int n_cand, const float* segments_dev, int n_seg, int horizon, int checkpoint_every,
                              const float* work_dev, const float* cost_dev, const float* grad_dev) {
  char msg[160];
  if (!sim || sim->decimation < 1 || sim->decimation > kMaxDecimation || sim->substeps < 1 ||
      sim->substeps > kMaxSubsteps || !(sim->dt > 0.0f) || !(sim->state_limit > 0.0f)) {
    std::snprintf(msg, sizeof(msg), "%s: invalid SpiSimArgs (decimation, substeps <= 8)", who);
    spi_set_error(msg);
    return SPI_ERR_INVALID_ARG;
  }
  if (!cost || n_cand < 0 || ld < n_cand || n_seg < 0 || horizon < 1 || checkpoint_every < 1 || !params_dev ||
      !segments_dev || !cost_dev || !grad_dev || !work_dev) {
    std::snprintf(msg, sizeof(msg), "%s: bad sizes or null pointer", who);
    spi_set_error(msg);
    return SPI_ERR_INVALID_ARG;
  }
  return SPI_OK;
}

}  // namespace spi_adj_launch
