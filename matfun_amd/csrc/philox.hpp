// philox.hpp -- the counter-based generator Philox4x32-10 (J. K. Salmon, M. A. Moraes, R. O. Dror, D. E. Shaw,
// "Parallel random numbers: as easy as 1, 2, 3", SC'11): ten rounds of two 32x32->64 multiplications on a 128-bit
// counter with a 64-bit key bumped by the Weyl constants between rounds; and standard normals from it by the
// Box-Muller transform.  Host and device (the perturbation kernel of the Newton driver; the CPU lane check).
#pragma once
#include "mf_device.hpp"

namespace mf {

struct U4 {
  uint32_t x[4];
};

MF_DEV U4 philox4x32_10(U4 c, uint32_t k0, uint32_t k1) {
  const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u, W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
#pragma unroll
  for (int r = 0; r < 10; r++) {
    const uint64_t p0 = (uint64_t)M0 * c.x[0], p1 = (uint64_t)M1 * c.x[2];
    const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0, hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    U4 n;
    n.x[0] = hi1 ^ c.x[1] ^ k0;
    n.x[1] = lo1;
    n.x[2] = hi0 ^ c.x[3] ^ k1;
    n.x[3] = lo0;
    c = n;
    k0 += W0;
    k1 += W1;
  }
  return c;
}

// two standard normals from two 32-bit words: u = (w + 1/2) 2^-32 in (0, 1), r = sqrt(-2 log u1),
// (r cos 2 pi u2, r sin 2 pi u2)
MF_DEV void box_muller(uint32_t w1, uint32_t w2, double& z0, double& z1) {
  const double TWO_PI = 6.283185307179586;
  const double u1 = ((double)w1 + 0.5) * 2.3283064365386963e-10, u2 = ((double)w2 + 0.5) * 2.3283064365386963e-10;
  const double r = sqrt(-2.0 * log(u1));
  double sn, cs;
  sincos(TWO_PI * u2, &sn, &cs);                 // (one shared argument reduction)
  z0 = r * cs;
  z1 = r * sn;
}

// the relative perturbation factors 1 + sigma z of the 9 entries of matrix m at a Newton iteration, z from
// counter (m low, m high, iteration, block) under key (seed low, seed high), blocks 0..2 (four words -> four
// normals per block, numbered 4 b ..; the first 9 are used, and only the Box-Muller pairs that hold them are
// computed).  sym: the entries (i, j) and (j, i) share one normal (number k of the 6 of the upper triangle, row-major),
// so a symmetric matrix stays symmetric (normals 0..5: blocks 0 and 1 only).
MF_DEV void perturb_factors(uint64_t seed, uint32_t iteration, uint64_t m, bool sym, double sigma, double* f9) {
  double z[12];
  const int pairs = sym ? 3 : 5;                 // (normals used: 6 or 9; the pairs past them are not computed)
#pragma unroll
  for (int b = 0; b < 3; b++) {
    if (2 * b >= pairs) break;
    U4 c;
    c.x[0] = (uint32_t)m;
    c.x[1] = (uint32_t)(m >> 32);
    c.x[2] = iteration;
    c.x[3] = (uint32_t)b;
    const U4 w = philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
    box_muller(w.x[0], w.x[1], z[4 * b], z[4 * b + 1]);
    if (2 * b + 1 < pairs) box_muller(w.x[2], w.x[3], z[4 * b + 2], z[4 * b + 3]);
  }
  // upper-triangle position of (i, j), i <= j: (0,0) 0, (0,1) 1, (0,2) 2, (1,1) 3, (1,2) 4, (2,2) 5
  constexpr int UP[9] = {0, 1, 2, 1, 3, 4, 2, 4, 5};
#pragma unroll
  for (int k = 0; k < 9; k++) f9[k] = 1.0 + sigma * z[sym ? UP[k] : k];
}

}  // namespace mf
