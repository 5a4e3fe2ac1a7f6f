// mf_device.hpp -- device-side helpers shared by the smf and gmf kernels (gfx950, FP64).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

// (host and device: the per-lane math is also compiled for the CPU to check it there, tests/lanecheck)
#define MF_DEV __host__ __device__ __forceinline__

namespace mf {

enum Kind { EXP = 0, LOG = 1, POW = 2, SQRT = 3 };

// largest power of two <= x for a positive normal x: the exponent bits of x alone (exact)
MF_DEV double pow2_floor(double x) {
  return __builtin_bit_cast(double, __builtin_bit_cast(long long, x) & 0x7ff0000000000000LL);
}
MF_DEV bool sign_bit(double x) { return __builtin_bit_cast(long long, x) < 0; }
MF_DEV double dinf() { return __builtin_bit_cast(double, 0x7ff0000000000000LL); }
MF_DEV double dnan() { return __builtin_bit_cast(double, 0x7ff8000000000000LL); }
// a power of two within a factor 2 of x; 1 outside (1e-300, 1e300) and for NaN
MF_DEV double pow2_near(double x) { return (x > 1e-300 && x < 1e300) ? pow2_floor(x) : 1.0; }

// 0^e: inf (e < 0), 1 (e = 0), 0 (e > 0)
MF_DEV double pow_of_zero(double e) { return e < 0 ? dinf() : (e == 0 ? 1.0 : 0.0); }

// x^(al - k), x > 0, with the exponent al - k = bh + bl split exactly (TwoSum): the rounding of al - k alone would
// cost |log x| times its error; x^(bh + bl) = x^bh (1 + bl log x)
MF_DEV double powk(double x, double al, double k) {
  const double bh = al - k, bb = bh - al, bl = (al - (bh - bb)) + (-k - bb);
  const double p = pow(x, bh);
  return bl == 0 ? p : p * (1.0 + bl * log(x));
}

MF_DEV double dmax(double a, double b) { return a > b ? a : b; }
MF_DEV double dmin(double a, double b) { return a < b ? a : b; }
MF_DEV double copysign_of(double x, double y) { return y < 0 ? -fabs(x) : fabs(x); }

// ---- complex numbers as (re, im)
struct cpx {
  double re, im;
};
MF_DEV cpx mk(double re, double im) { return cpx{re, im}; }
MF_DEV cpx add(cpx a, cpx b) { return cpx{a.re + b.re, a.im + b.im}; }
MF_DEV cpx sub(cpx a, cpx b) { return cpx{a.re - b.re, a.im - b.im}; }
MF_DEV cpx mul(cpx a, cpx b) { return cpx{a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
MF_DEV cpx scale(cpx a, double s) { return cpx{a.re * s, a.im * s}; }
MF_DEV cpx pick(bool m, cpx a, cpx b) { return cpx{m ? a.re : b.re, m ? a.im : b.im}; }
// |a| with the squares scaled (no overflow / underflow); 0 for a NaN modulus
MF_DEV double cabs_(cpx a) {
  const double ar = fabs(a.re), ai = fabs(a.im), big = dmax(ar, ai), sml = dmin(ar, ai);
  const double q = sml / (big > 0 ? big : 1.0);
  return big > 0 ? big * sqrt(1.0 + q * q) : 0.0;
}
// 1 / b by Smith's method
MF_DEV cpx cinv(cpx b) {
  const bool sw = fabs(b.im) > fabs(b.re);
  const double p = sw ? b.im : b.re, q = sw ? b.re : b.im;
  const double r = q / (p != 0 ? p : 1.0), den = p + q * r;
  const double re = 1.0 / den, im = -r / den;
  return sw ? cpx{-im, -re} : cpx{re, im};
}

// ---- index maps of the symmetric layouts
// Voigt index of the row-major 3x3 position m = 3a + b: diagonal 0,1,2; (0,1) 3; (0,2) 4; (1,2) 5
MF_DEV constexpr int voigt_of_m(int m) { return m % 4 == 0 ? m / 4 : ((m == 1 || m == 3) ? 3 : ((m == 2 || m == 6) ? 4 : 5)); }
// d v_voigt / d A_m: 1 on the diagonal, 1/2 off it
MF_DEV constexpr double dv_dA(int m) { return m % 4 == 0 ? 1.0 : 0.5; }
// position of the symmetric pair (p, q) of Voigt indices in a packed upper triangle of 21 entries
MF_DEV constexpr int tri6(int p, int q) { return p <= q ? p * (11 - p) / 2 + q : q * (11 - q) / 2 + p; }

}  // namespace mf
