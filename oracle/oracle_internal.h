/* oracle_internal.h -- shared helpers of the CPU oracle (test infrastructure; see oracle.h). */
#ifndef MATFUN_ORACLE_INTERNAL_H
#define MATFUN_ORACLE_INTERNAL_H
#include <math.h>
#include <stdint.h>

enum { K_EXP = 0, K_LOG = 1, K_POW = 2, K_SQRT = 3 };

/* argument check and exponent normalisation of the C API (c/src/matfun.cc:23-27): returns 0 if invalid */
static inline int oracle_prepare(int *kind, double *alpha, int64_t n, const double *J, const double *H, int general) {
  if (*kind < K_EXP || *kind > K_SQRT || n < 0 || (H && !J) || (*kind == K_POW && !(*alpha > 0))) return 0;
  if (general && *kind == K_SQRT) { *kind = K_POW; *alpha = 0.5; }
  if (*kind != K_POW) *alpha = *kind == K_SQRT ? 0.5 : 0.0;
  return 1;
}

/* largest power of two <= x (x > 0 and normal), as smf_kernel.hpp:54 */
static inline double pow2_floor(double x) { return ldexp(1.0, ilogb(x)); }

/* x^(al - k) with the exponent split exactly into bh + bl (TwoSum), smf_kernel.hpp:66-73 */
static inline double powk(double x, double al, double k) {
  const double bh = al - k, bb = bh - al, bl = (al - (bh - bb)) + (-k - bb);
  const double p = pow(x, bh);
  return bl == 0 ? p : p * (1.0 + bl * log(x));
}

/* 0^e: inf (e < 0), 1 (e = 0), 0 (e > 0), smf_kernel.hpp:58 */
static inline double pow_of_zero(double e) { return e < 0 ? INFINITY : (e == 0 ? 1.0 : 0.0); }

static inline double dmax(double a, double b) { return a > b ? a : b; }   /* (a > b ? a : b: NaN in a -> b) */
static inline double dmin(double a, double b) { return a < b ? a : b; }

/* ---- complex numbers (re, im) */
typedef struct { double re, im; } cplx;
static inline cplx cp(double re, double im) { cplx z = {re, im}; return z; }
static inline cplx cp_add(cplx a, cplx b) { return cp(a.re + b.re, a.im + b.im); }
static inline cplx cp_sub(cplx a, cplx b) { return cp(a.re - b.re, a.im - b.im); }
static inline cplx cp_mul(cplx a, cplx b) { return cp(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re); }
static inline cplx cp_scale(cplx a, double s) { return cp(a.re * s, a.im * s); }
/* |a| scaled so the squares neither overflow nor underflow (gmf_kernel.hpp:112-117) */
static inline double cp_abs(cplx a) {
  const double ar = fabs(a.re), ai = fabs(a.im), big = dmax(ar, ai), sml = dmin(ar, ai);
  if (!(big > 0)) return 0.0;
  const double q = sml / big;
  return big * sqrt(1.0 + q * q);
}
/* 1 / b by Smith's method (gmf_kernel.hpp:118-125) */
static inline cplx cp_inv(cplx b) {
  const int sw = fabs(b.im) > fabs(b.re);
  const double p = sw ? b.im : b.re, q = sw ? b.re : b.im;
  const double r = q / (p != 0 ? p : 1.0), den = p + q * r;
  const double re = 1.0 / den, im = -r / den;
  return sw ? cp(-im, -re) : cp(re, im);
}

/* smf (smf_oracle.c): one matrix, Voigt layout in and out; J (36) and H (216) may be NULL (H needs J);
 * c2: the 30 second-divided-difference series coefficients of smf_dd2_coefs */
#define SMF_NTERMS 30
void smf_dd2_coefs(int kind, double alpha, double *c2);
void smf_eval_one(int kind, double alpha, const double *v, double *F, double *J, double *H, const double *c2);
/* Voigt layout -> matrix layout (weights 1 / 0.5), and the symmetry test of full matrices */
void smf_voigt_to_matrix(const double *Fv, const double *Jv, const double *Hv, double *F, double *J, double *H);
int smf_matrix_to_voigt(const double *A, double *v);

/* gmf (gmf_oracle.c): one matrix; order 0..2 (J, H used for order >= 1, >= 2); bt: 200 series coefficients */
#define GMF_KMAX 200
void gmf_series_coefs(int kind, double alpha, double *bt);
int gmf_eval_one(int kind, int order, double alpha, const double *A, double *F, double *J, double *H,
                 const double *bt);

#endif
