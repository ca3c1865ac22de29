// Counter-based random numbers: Philox4x32-10 (Salmon, Moraes, Dror, Shaw,
// "Parallel random numbers: as easy as 1, 2, 3", SC'11). Every draw is a pure
// function of (seed, stream, epoch, index, draw), so momenta, the die roll and
// Haar links do not depend on launch shape, batching or device.
//
// Stream layout (shared with the CPU oracle):
//   key = seed, counter = {index, draw, epoch_lo32, stream | epoch_hi32 << 8}
//   stream 1 momentum (index = link, epoch = trajectory)
//   stream 2 Haar links (index = link), 3 accept die roll (index = trajectory)
//   stream 4 demon proposals (index = link, epoch = sweep)
#pragma once
#include <stdint.h>

#include "su3.hpp"

namespace mag {

enum : uint32_t { RNG_MOMENTUM = 1, RNG_HAAR = 2, RNG_ACCEPT = 3, RNG_DEMON = 4, RNG_HEATBATH = 5, RNG_LOCAL = 6 };

struct u32x4 {
  uint32_t x, y, z, w;
};

MAG_HD u32x4 philox4x32_10(u32x4 c, uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    if (r) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
    const uint64_t p0 = (uint64_t)0xD2511F53u * c.x;
    const uint64_t p1 = (uint64_t)0xCD9E8D57u * c.z;
    const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c.y ^ k0;
    const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c.w ^ k1;
    c.x = n0;
    c.y = (uint32_t)p1;
    c.z = n2;
    c.w = (uint32_t)p0;
  }
  return c;
}

MAG_HD u32x4 rng_draw(uint64_t seed, uint32_t stream, uint64_t epoch, uint32_t index, uint32_t d) {
  u32x4 c = {index, d, (uint32_t)epoch, stream | ((uint32_t)(epoch >> 32) << 8)};
  return philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
}

// [0,1) from 53 random bits
MAG_HD double rng_u01(uint32_t hi, uint32_t lo) {
  const uint64_t x = (((uint64_t)hi << 32) | lo) >> 11;
  return (double)x * (1. / 9007199254740992.);
}
// (0,1]
MAG_HD double rng_u01_open0(uint32_t hi, uint32_t lo) {
  const uint64_t x = (((uint64_t)hi << 32) | lo) >> 11;
  return ((double)x + 1.) * (1. / 9007199254740992.);
}
// Box-Muller: one Philox block -> two N(0,1) doubles.
MAG_HD void rng_gauss2(u32x4 w, double& g0, double& g1) {
  const double u1 = rng_u01_open0(w.x, w.y);
  const double u2 = rng_u01(w.z, w.w);
  const double r = sqrt(-2. * log(u1));
  const double t = 6.283185307179586 * u2;
  g0 = r * cos(t);
  g1 = r * sin(t);
}

}  // namespace mag
