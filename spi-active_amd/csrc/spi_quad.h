// Lane exchange inside a group of four consecutive lanes (a DPP quad permutation): the value
// of lane (this lane xor M), M = 1 or 2.  A VALU move with no LDS round trip, for the 4-lane
// sums of the leg-parallel kernels.  All four lanes of the group must be active.
#pragma once
#include <hip/hip_runtime.h>

template <int M> __device__ __forceinline__ float spi_quad_xor(float x) {
  static_assert(M == 1 || M == 2, "swaps inside a quad: xor 1 or xor 2");
  constexpr int kCtrl = M == 1 ? 0xB1 : 0x4E;   // quad_perm [1,0,3,2] / [2,3,0,1]
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), kCtrl, 0xF, 0xF, false));
}

// In a source built with -fassociative-math (which the build marks with SPI_REASSOCIATE=1) an
// empty asm the optimiser cannot look through: a value passed through it is a leaf for
// reassociation, so a quad sum adds the same two operands in each lane (in opposite order,
// which float addition does not notice) and leaves the same bits in all four.  Without
// reassociation the sums cannot be regrouped and the value passes unchanged (the asm would
// only constrain scheduling).
__device__ __forceinline__ float spi_quad_pin(float x) {
#if SPI_REASSOCIATE
  asm("" : "+v"(x));
#endif
  return x;
}
