// ORACLE — test infrastructure only.
// Philox4x32-10 counter-based generator, restated from the published algorithm
// (J. K. Salmon, M. A. Moraes, R. O. Dror, D. E. Shaw, "Parallel Random Numbers:
// As Easy as 1, 2, 3", SC'11, section 3): ten rounds of two 32x32->64 multiplies
// with the round keys bumped by the Weyl constants; plus Box-Muller normals.
#pragma once
#include <cmath>
#include <cstdint>

namespace spio {

struct Philox4 { uint32_t v[4]; };

inline Philox4 philox4x32_10(Philox4 ctr, uint32_t k0, uint32_t k1) {
  const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
  const uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
  for (int r = 0; r < 10; ++r) {
    uint64_t p0 = uint64_t(M0) * ctr.v[0];
    uint64_t p1 = uint64_t(M1) * ctr.v[2];
    uint32_t hi0 = uint32_t(p0 >> 32), lo0 = uint32_t(p0);
    uint32_t hi1 = uint32_t(p1 >> 32), lo1 = uint32_t(p1);
    Philox4 n;
    n.v[0] = hi1 ^ ctr.v[1] ^ k0;
    n.v[1] = lo1;
    n.v[2] = hi0 ^ ctr.v[3] ^ k1;
    n.v[3] = lo0;
    ctr = n;
    k0 += W0;
    k1 += W1;
  }
  return ctr;
}

// uint32 -> uniform in the open interval (0, 1)
inline double u01_open(uint32_t x) { return (double(x) + 0.5) * (1.0 / 4294967296.0); }

// Standard normal draw number `idx` of stream (seed, iteration): one Philox block yields
// four uint32, i.e. two Box-Muller pairs -> four normals.
inline double philox_normal(uint64_t seed, uint32_t iteration, uint64_t idx) {
  Philox4 c;
  uint64_t blk = idx >> 2;
  c.v[0] = uint32_t(blk);
  c.v[1] = uint32_t(blk >> 32);
  c.v[2] = iteration;
  c.v[3] = 0x53504941u;  // stream tag
  Philox4 r = philox4x32_10(c, uint32_t(seed), uint32_t(seed >> 32));
  int lane = int(idx & 3);
  int pair = lane >> 1;
  double u1 = u01_open(r.v[2 * pair]), u2 = u01_open(r.v[2 * pair + 1]);
  double rad = std::sqrt(-2.0 * std::log(u1));
  double ang = 6.283185307179586 * u2;
  return (lane & 1) ? rad * std::sin(ang) : rad * std::cos(ang);
}

}  // namespace spio
