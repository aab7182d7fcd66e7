// Launch-shape and bit-level helpers shared by the kernel translation units.
#pragma once
#include <hip/hip_runtime.h>

namespace spi_launch {

// Workgroups are dispatched round-robin over the 8 XCDs of an MI355X.  Remap the hardware
// block id so that logically consecutive blocks (same segment or excitation, neighbouring
// candidate blocks) run on the same XCD and share its L2.
__device__ __forceinline__ int xcd_remap(int b, int nb) {
  constexpr int NX = 8;
  int q = nb / NX, r = nb % NX;
  int x = b % NX, i = b / NX;
  int base = x < r ? x * (q + 1) : r * (q + 1) + (x - r) * q;
  return base + i;
}

// Finiteness from the exponent bits: stays correct when a file is compiled with
// -ffinite-math-only (which lets the compiler fold the model's many x*0 terms).
__device__ __forceinline__ bool finite_bits(float x) { return (__float_as_uint(x) & 0x7f800000u) != 0x7f800000u; }

}  // namespace spi_launch
