/* gmf_oracle.c -- general real 3x3 matrices: F = f(A), J = dF/dA, H = d2F/dA2 for f = exp, log, x^alpha (principal
 * branch) by the Newton-Hermite form over the eigenvalues with confluent divided differences (exact for defective A
 * and complex pairs).  Scalar FP64 restatement of the reference kernel c/include/matfun/gmf_kernel.hpp (roots3,
 * newton, the pair refinements, node order, groups, the streamed group series, single-node residues, real basis,
 * F / J / H assembly) with its constants and branch choices.  TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Divided differences f[x1^a x2^b x3^c] of the Newton form (a >= b >= c >= 0, a = order + 1 at most 3) are kept in
 * a list of 19 entries (3 / 9 / 19 for orders 0 / 1 / 2): position dd_pos(a, b, c); each entry has one node more
 * than its parent (the node added is the first whose multiplicity exceeds the next one's).  The tables are built
 * from that rule once (gmf_kernel.hpp:34-46, :254-258 list the same structure).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle_internal.h"

static const double RHO = 0.7, R_EXP = 4.0, RHO_POLE = 0.6;   /* gmf_kernel.hpp:30 */
static const double TWO_PI_3 = 2.0943951023931957;

/* position of f[x1^a x2^b x3^c] in the list: orders by the multiplicity of node 1, then 2, then 3 */
static int dd_pos(int a, int b, int c) {
  if (a == 1) return b + c;
  if (a == 2) return b == 0 ? 3 : (b == 1 ? 4 + c : 6 + c);
  return b == 0 ? 9 : (b == 1 ? 10 + c : (b == 2 ? 12 + c : 15 + c));
}
/* list positions of the divided differences over N_j, N_j + N_k, N_j + N_k + N_l (N_j = {x_1 .. x_(j+1)}) */
static int pos1(int j) { return dd_pos(1, j >= 1, j >= 2); }
static int pos2(int j, int k) { return dd_pos(2, (j >= 1) + (k >= 1), (j >= 2) + (k >= 2)); }
static int pos3(int j, int k, int l) { return dd_pos(3, (j >= 1) + (k >= 1) + (l >= 1), (j >= 2) + (k >= 2) + (l >= 2)); }

static int MULT[19][3], PARENT[19], ADDED[19];
static int tables_ready = 0;
static void build_tables(void) {
  if (tables_ready) return;
  for (int a = 1; a <= 3; a++)
    for (int b = 0; b <= a; b++)
      for (int c = 0; c <= b; c++) {
        const int t = dd_pos(a, b, c);
        MULT[t][0] = a; MULT[t][1] = b; MULT[t][2] = c;
      }
  for (int t = 0; t < 19; t++) {
    int m[3] = {MULT[t][0], MULT[t][1], MULT[t][2]};
    const int node = m[0] > m[1] ? 0 : (m[1] > m[2] ? 1 : 2);
    ADDED[t] = node;
    m[node]--;
    PARENT[t] = m[0] == 0 ? -1 : dd_pos(m[0], m[1], m[2]);
  }
  tables_ready = 1;
}

/* series coefficients b_j of f(c (1 + u)) / scale (gmf_kernel.hpp:48-53): exp 1/j!, log (-1)^(j+1)/j, pow binom(al, j) */
void gmf_series_coefs(int kind, double al, double *b) {
  build_tables();
  b[0] = 1.0;
  for (int j = 1; j < GMF_KMAX; j++)
    b[j] = kind == K_EXP ? b[j - 1] / j : (kind == K_LOG ? ((j & 1) ? 1.0 : -1.0) / j : b[j - 1] * (al - j + 1) / j);
}

/* terms of the group series (gmf_kernel.hpp:73-97): sum_k b_(n+k) h_k, |h_k| <= C(k+n, n) r^k, and the pole series
 * C(k+2, 2) q^k, until a bound below 1e-17 x the largest term */
static int series_terms(const double *b, int n, double r, double q) {
  int M = 1;
  if (r > 0) {
    double best = 0.0, c = 1.0, rk = 1.0;
    for (int k = 0; n + k < GMF_KMAX; k++) {
      const double t = fabs(b[n + k]) * c * rk;
      best = t > best ? t : best;
      M = k;
      if ((k > 2 && t <= 1e-17 * best) || !(rk > 0)) break;
      c = c * (k + n + 1) / (k + 1);
      rk *= r;
    }
  }
  if (q > 0) {
    double qk = 1.0;
    for (int k = 0; n + k < GMF_KMAX; k++) {
      if (k > M && (k + 1.0) * (k + 2.0) / 2.0 * qk <= 1e-17) break;
      M = k > M ? k : M;
      qk *= q;
    }
  }
  return n + M + 1 < GMF_KMAX ? n + M + 1 : GMF_KMAX;
}

static double det3(const double *M) {
  return M[0] * (M[4] * M[8] - M[5] * M[7]) - M[1] * (M[3] * M[8] - M[5] * M[6]) + M[2] * (M[3] * M[7] - M[4] * M[6]);
}

/* det by elimination with partial pivoting (exact pivots for diagonal M), gmf_kernel.hpp:138-159 */
static double det_pivot(const double *M) {
  const double a0 = fabs(M[0]), a1 = fabs(M[3]), a2 = fabs(M[6]);
  const int p1 = (a1 > a0) && (a1 >= a2), p2 = (a2 > a0) && (a2 > a1);
  double P[3], R1[3], R2[3];
  for (int c = 0; c < 3; c++) {               /* row order (p, rest): 0,1,2 / 1,0,2 / 2,0,1 */
    P[c] = p1 ? M[3 + c] : (p2 ? M[6 + c] : M[c]);
    R1[c] = (p1 || p2) ? M[c] : M[3 + c];
    R2[c] = p2 ? M[3 + c] : M[6 + c];
  }
  double sgn = p1 ? -1.0 : 1.0;
  const double piv = P[0] != 0 ? P[0] : 1.0;
  const double l1 = R1[0] / piv, l2 = R2[0] / piv;
  const double b = R1[1] - l1 * P[1], c = R1[2] - l1 * P[2], d = R2[1] - l2 * P[1], e = R2[2] - l2 * P[2];
  const int sw = fabs(d) > fabs(b);
  const double bb = sw ? d : b, cc = sw ? e : c, dd = sw ? b : d, ee = sw ? c : e;
  if (sw) sgn = -sgn;
  const double det = sgn * P[0] * bb * (ee - dd / (bb != 0 ? bb : 1.0) * cc);
  return (P[0] != 0 && bb != 0) ? det : 0.0;
}

static double minors2(const double *M) {
  return M[0] * M[4] - M[1] * M[3] + M[0] * M[8] - M[2] * M[6] + M[4] * M[8] - M[5] * M[7];
}

/* a power of two within a factor 2 of x > 0; 1 for 0, tiny, huge, NaN (gmf_kernel.hpp:126-131) */
static double pow2_near(double x) { return (x > 1e-300 && x < 1e300) ? pow2_floor(x) : 1.0; }
static double copysign_of(double x, double y) { return y < 0 ? -fabs(x) : fabs(x); }

/* eigenvalues of S (gmf_kernel.hpp:166-208): tau + the roots of the traceless, twice-normalised cubic
 * x^3 - P x - Q; x[0] the trigonometric (three real) or Cardano root, x[1], x[2] from the deflated quadratic */
static void cubic_roots(const double *S, cplx *x) {
  const double tau = (S[0] + S[4] + S[8]) / 3.0;
  double Sp[9];
  for (int k = 0; k < 9; k++) Sp[k] = k % 4 == 0 ? S[k] - tau : S[k];
  double s = fabs(Sp[0]);
  for (int k = 1; k < 9; k++) s = dmax(s, fabs(Sp[k]));
  double t = pow2_near(s);
  const double it = 1.0 / t;
  for (int k = 0; k < 9; k++) Sp[k] = Sp[k] * it;
  double P = 0.5 * (Sp[0] * Sp[0] + Sp[4] * Sp[4] + Sp[8] * Sp[8]) + Sp[1] * Sp[3] + Sp[2] * Sp[6] + Sp[5] * Sp[7];
  double Q = det3(Sp);
  const double u = pow2_near(dmax(sqrt(fabs(P)), cbrt(fabs(Q)))), iu = 1.0 / u;
  P = P * iu * iu;
  Q = Q * iu * iu * iu;
  t = t * u;
  const int trig = (P > 0) && (4.0 * P * P * P >= 27.0 * Q * Q);
  double s1;
  if (trig) {
    const double p3 = P / 3.0;
    double r = 0.5 * Q / (p3 * sqrt(p3));
    r = r > 1.0 ? 1.0 : (r < -1.0 ? -1.0 : r);
    const double ph = acos(r) / 3.0;
    s1 = 2.0 * sqrt(p3) * cos(r >= 0 ? ph : ph + TWO_PI_3);
  } else {
    const double D = 0.25 * Q * Q - P * P * P / 27.0;
    const double uc = cbrt(0.5 * Q + copysign_of(sqrt(D > 0 ? D : 0.0), Q));
    s1 = uc + (uc != 0 ? P / (3.0 * uc) : 0.0);
  }
  const double d = P - 0.75 * s1 * s1;                  /* x^2 + s1 x + (s1^2 - P) = 0 */
  const int real = d >= 0;
  const double sq = sqrt(fabs(d));
  const double sa = -0.5 * s1 - copysign_of(sq, s1);
  const double prod = s1 != 0 ? Q / s1 : -P;
  const double sb = sa != 0 ? prod / sa : 0.0;
  const double im = (real ? 0.0 : sq) * t;
  x[0] = cp(s1 * t + tau, 0.0 * t);
  x[1] = cp((real ? sa : -0.5 * s1) * t + tau, im);
  x[2] = cp((real ? sb : -0.5 * s1) * t + tau, -im);
}

/* real node j: 3 Newton steps on det(As - mu I), the derivative from the other nodes (gmf_kernel.hpp:210-226) */
static void newton_node(const double *As, cplx *mu, int j) {
  for (int it = 0; it < 3; it++) {
    cplx den = cp(1.0, 0.0);
    for (int k = 0; k < 3; k++)
      if (k != j) den = cp_mul(den, cp_sub(mu[k], mu[j]));
    if (!((den.re != 0) || (den.im != 0))) continue;
    double B[9];
    for (int k = 0; k < 9; k++) B[k] = k % 4 == 0 ? As[k] - mu[j].re : As[k];
    mu[j].re = mu[j].re + det_pivot(B) * cp_inv(den).re;
  }
}

/* Taylor coefficient 0 of f at a node (gmf_kernel.hpp:228-248): exp e^x (relative to the shift); log log(sig mu)
 * (log1p of x/m near m); pow (sig |mu|)^(al-k) e^(i (al-k) arg mu) mu^k, i.e. f^(k) / binomial factors */
static cplx node_value(int kind, double al, double sig, double m, cplx x, cplx mu, int k) {
  if (kind == K_EXP) {
    const double e = exp(x.re);
    return cp(e * cos(x.im), e * sin(x.im));
  }
  const double th = atan2(mu.im, mu.re);
  if (kind == K_LOG) {
    if ((m > 0) && (cp_abs(x) <= 0.5 * m)) {
      const double zr = x.re / m, zi = x.im / m;
      return cp(log(sig * m) + 0.5 * log1p(zr * (2.0 + zr) + zi * zi), atan2(zi, 1.0 + zr));
    }
    return cp(log(sig * cp_abs(mu)), th);
  }
  const double mag = powk(sig * cp_abs(mu), al, (double)k);
  const double ph = (al - k) * th;
  cplx out = cp(mag * cos(ph), mag * sin(ph));
  for (int i = 0; i < k; i++) out = cp_mul(out, mu);
  return out;
}

/* ---- one matrix (gmf_kernel.hpp:251-692); kind EXP, LOG or POW (sqrt: POW 0.5) */
static int gmf_core(int kind, int order, double al, const double *a, double *F, double *J, double *H,
                    const double *bt) {
  build_tables();
  const int ND = order == 0 ? 3 : (order == 1 ? 9 : 19);
  /* scaling by a power of two (exp: none), shift m = tr / 3, deviator S */
  double amax = fabs(a[0]);
  for (int k = 1; k < 9; k++) amax = dmax(amax, fabs(a[k]));
  const double sig = kind == K_EXP ? 1.0 : pow2_near(amax), isig = 1.0 / sig;
  double As[9], S[9];
  for (int k = 0; k < 9; k++) As[k] = a[k] * isig;
  const double m = (As[0] + As[4] + As[8]) / 3.0;
  for (int k = 0; k < 9; k++) S[k] = k % 4 == 0 ? As[k] - m : As[k];
  cplx x[3], mu[3];
  cubic_roots(S, x);
  for (int i = 0; i < 3; i++) mu[i] = cp(m + x[i].re, x[i].im);

  /* the pair from sum and product with A itself after Newton on mu[0], where the roots of S are not the more
   * accurate ones (gmf_kernel.hpp:277-304) */
  double smax = fabs(S[0]);
  for (int k = 1; k < 9; k++) smax = dmax(smax, fabs(S[k]));
  const double mid = 0.5 * (mu[1].re + mu[2].re), Rp = 0.5 * cp_abs(cp_sub(mu[1], mu[2]));
  const int small = fabs(mid) + Rp <= 0.25 * fabs(mu[0].re);
  const int aniso = (fabs(m) <= smax) && (Rp <= 0.25 * fabs(mid - mu[0].re));
  int pc = (mu[0].re != 0) && (small || aniso);
  double pcen = 0.0, ps = 0.0;
  int preal = 1;
  if (pc) {
    newton_node(As, mu, 0);
    const double m0 = mu[0].re;
    const int big = fabs(m0) >= fabs(mid) + Rp;
    double sg, pi;
    if (big) {                                 /* product pi = det A / mu0, sum sigma = (c2 - pi) / mu0 */
      pi = det_pivot(As) / m0;
      sg = (minors2(As) - pi) / m0;
    } else {                                   /* sigma = tr A - mu0, pi = c2 - mu0 sigma */
      sg = As[0] + As[4] + As[8] - m0;
      pi = minors2(As) - m0 * sg;
    }
    const double disc = 0.25 * sg * sg - pi;
    pcen = 0.5 * sg;
    ps = sqrt(fabs(disc));
    preal = disc >= 0;
    const double r1 = pcen + copysign_of(ps, pcen);
    const double r2 = r1 != 0 ? pi / r1 : pcen - copysign_of(ps, pcen);
    mu[1] = preal ? cp(r1, 0.0) : cp(pcen, ps);
    mu[2] = preal ? cp(r2, 0.0) : cp(pcen, -ps);
  }
  /* a conjugate pair near the negative real axis (log, powers): beta^2 from the 2x2 block of the deviator on the
   * complement of the left eigenvector of mu[0] (gmf_kernel.hpp:305-367) */
  const int cut = kind != K_EXP && (mu[1].re + mu[2].re < 0);
  if (cut) {
    if (!pc) newton_node(As, mu, 0);
    const double l0 = mu[0].re;
    double Mc[9];
    for (int k = 0; k < 9; k++) Mc[k] = k % 4 == 0 ? As[k] - l0 : As[k];
    /* u orthogonal to the columns of A - mu0 I: the longest cross product of two columns */
    const double c01[3] = {Mc[3] * Mc[7] - Mc[6] * Mc[4], Mc[6] * Mc[1] - Mc[0] * Mc[7], Mc[0] * Mc[4] - Mc[3] * Mc[1]};
    const double c02[3] = {Mc[3] * Mc[8] - Mc[6] * Mc[5], Mc[6] * Mc[2] - Mc[0] * Mc[8], Mc[0] * Mc[5] - Mc[3] * Mc[2]};
    const double c12[3] = {Mc[4] * Mc[8] - Mc[7] * Mc[5], Mc[7] * Mc[2] - Mc[1] * Mc[8], Mc[1] * Mc[5] - Mc[4] * Mc[2]};
    const double n01 = c01[0] * c01[0] + c01[1] * c01[1] + c01[2] * c01[2];
    const double n02 = c02[0] * c02[0] + c02[1] * c02[1] + c02[2] * c02[2];
    const double n12 = c12[0] * c12[0] + c12[1] * c12[1] + c12[2] * c12[2];
    double u[3] = {c01[0], c01[1], c01[2]}, nn = n01;
    if (n02 > n01) { u[0] = c02[0]; u[1] = c02[1]; u[2] = c02[2]; nn = n02; }
    if (n12 > nn) { u[0] = c12[0]; u[1] = c12[1]; u[2] = c12[2]; nn = n12; }
    if (nn > 0) {
      const double iu = 1.0 / sqrt(nn);
      for (int k = 0; k < 3; k++) u[k] = u[k] * iu;
      /* orthonormal basis q1, q2 of the complement of u */
      const double au0 = fabs(u[0]), au1 = fabs(u[1]), au2 = fabs(u[2]);
      const int k1 = au1 < au0, k2 = au2 < (k1 ? au1 : au0);
      const double ex = (k1 || k2) ? 0.0 : 1.0, ey = (k1 && !k2) ? 1.0 : 0.0, ez = k2 ? 1.0 : 0.0;
      const double eu = ex * u[0] + ey * u[1] + ez * u[2];
      double q1[3] = {ex - eu * u[0], ey - eu * u[1], ez - eu * u[2]};
      const double iq = 1.0 / sqrt(q1[0] * q1[0] + q1[1] * q1[1] + q1[2] * q1[2]);
      for (int k = 0; k < 3; k++) q1[k] = q1[k] * iq;
      const double q2[3] = {u[1] * q1[2] - u[2] * q1[1], u[2] * q1[0] - u[0] * q1[2], u[0] * q1[1] - u[1] * q1[0]};
      double Sq1[3], Sq2[3];
      for (int r = 0; r < 3; r++) {
        Sq1[r] = S[3 * r] * q1[0] + S[3 * r + 1] * q1[1] + S[3 * r + 2] * q1[2];
        Sq2[r] = S[3 * r] * q2[0] + S[3 * r + 1] * q2[1] + S[3 * r + 2] * q2[2];
      }
      const double b11 = q1[0] * Sq1[0] + q1[1] * Sq1[1] + q1[2] * Sq1[2];
      const double b12 = q1[0] * Sq2[0] + q1[1] * Sq2[1] + q1[2] * Sq2[2];
      const double b21 = q2[0] * Sq1[0] + q2[1] * Sq1[1] + q2[2] * Sq1[2];
      const double b22 = q2[0] * Sq2[0] + q2[1] * Sq2[1] + q2[2] * Sq2[2];
      const double dd = 0.5 * (b11 - b22), beta2 = -(dd * dd + b12 * b21), cc = m + 0.5 * (b11 + b22);
      const double bta = sqrt(fabs(beta2));
      const int creal = beta2 <= 0;
      mu[1] = creal ? cp(cc + bta, 0.0) : cp(cc, bta);
      mu[2] = creal ? cp(cc - bta, 0.0) : cp(cc, -bta);
      pcen = cc;
      ps = bta;
      preal = creal;
      pc = 1;
    }
  }
  int ab[3] = {pc, pc, pc};                    /* node refined against A (mu exact, x = mu - m derived) */
  int og[3] = {0, 1, 2};                       /* original position (1, 2: the pair) */
  for (int i = 0; i < 3; i++)
    if (ab[i]) x[i] = cp(mu[i].re - m, mu[i].im);

  /* node order: the node with the largest high derivatives last (exp: largest real part; log, pow: smallest
   * modulus), a conjugate pair adjacent (gmf_kernel.hpp:373-410) */
#define NODE_KEY(v) (kind == K_EXP ? (v).re : -cp_abs(v))
  {
    cplx sx[3] = {x[0], x[1], x[2]}, sm[3] = {mu[0], mu[1], mu[2]};
    int sa[3] = {ab[0], ab[1], ab[2]}, so[3] = {og[0], og[1], og[2]};
    static const int NET[3][2] = {{0, 1}, {1, 2}, {0, 1}};
    for (int s = 0; s < 3; s++) {
      const int i = NET[s][0], j = NET[s][1];
      if (NODE_KEY(sm[i]) > NODE_KEY(sm[j])) {
        cplx tx = sx[i], tm = sm[i];
        int ta = sa[i], to = so[i];
        sx[i] = sx[j]; sx[j] = tx;
        sm[i] = sm[j]; sm[j] = tm;
        sa[i] = sa[j]; sa[j] = ta;
        so[i] = so[j]; so[j] = to;
      }
    }
    const int first = (NODE_KEY(mu[0]) <= NODE_KEY(mu[1])) || cut;
    const int cplxpair = x[1].im != 0;
    cplx nx[3], nm[3];
    int na[3], no[3];
    for (int k = 0; k < 3; k++) {
      const int r = (k + 1) % 3;
      if (cplxpair) {                          /* real node first (or last): rotate (0,1,2) -> (1,2,0) */
        const int src = first ? k : r;
        nx[k] = x[src]; nm[k] = mu[src]; na[k] = ab[src]; no[k] = og[src];
      } else {
        nx[k] = sx[k]; nm[k] = sm[k]; na[k] = sa[k]; no[k] = so[k];
      }
    }
    for (int k = 0; k < 3; k++) { x[k] = nx[k]; mu[k] = nm[k]; ab[k] = na[k]; og[k] = no[k]; }
  }
#undef NODE_KEY

  /* groups: all three nodes, else the closest real or conjugate pair (with the pole of the third node) */
  /* (|al|: the per-matrix exponent extension also takes al <= 0; for al > 0 this is gmf_kernel.hpp:413) */
  const double rho_f = (kind == K_POW && 3.0 / fabs(al) < RHO) ? 3.0 / fabs(al) : RHO;
#define GROUP_OK(c, R) (kind == K_EXP ? ((R) <= R_EXP) : ((m + (c) > 0) && ((R) <= rho_f * (m + (c)))))
  const double cs = (x[0].re + x[1].re + x[2].re) / 3.0;
  const double R3 = dmax(dmax(cp_abs(cp(x[0].re - cs, x[0].im)), cp_abs(cp(x[1].re - cs, x[1].im))),
                         cp_abs(cp(x[2].re - cs, x[2].im)));
  const int triple = GROUP_OK(cs, R3);
  int psel[3] = {0, 0, 0};                     /* pairs (0,1), (0,2), (1,2) */
  double bestR = INFINITY;
  static const int PI_[3] = {0, 0, 1}, PJ_[3] = {1, 2, 2};
  for (int p = 0; p < 3; p++) {
    const int i = PI_[p], j = PJ_[p], k = 3 - i - j;
    const int elig = ((x[i].im == 0) && (x[j].im == 0)) || ((x[i].re == x[j].re) && (x[i].im == -x[j].im));
    const double c = 0.5 * (x[i].re + x[j].re), R = 0.5 * cp_abs(cp_sub(x[i], x[j]));
    const int okp = elig && GROUP_OK(c, R) && (R <= RHO_POLE * cp_abs(cp(c - x[k].re, -x[k].im))) && !triple;
    if (okp && R < bestR) {
      bestR = R;
      for (int q = 0; q < p; q++) psel[q] = 0;
      psel[p] = 1;
    }
  }
#undef GROUP_OK
  const int inG[3] = {triple || psel[0] || psel[1], triple || psel[0] || psel[2], triple || psel[1] || psel[2]};
  const int pk[3] = {psel[2], psel[1], psel[0]};   /* node outside the pair group */

  /* group centre and offsets; the refined pair as a group: symmetric offsets about its centre */
  const double gsum = (inG[0] ? 1.0 : 0.0) + (inG[1] ? 1.0 : 0.0) + (inG[2] ? 1.0 : 0.0);
  double cxg = ((inG[0] ? x[0].re : 0.0) + (inG[1] ? x[1].re : 0.0) + (inG[2] ? x[2].re : 0.0)) / (gsum > 0 ? gsum : 1.0);
  double cmu = m + cxg;
  int is12[3];
  for (int i = 0; i < 3; i++) is12[i] = og[i] > 0;
  const int ovr = pc && (gsum == 2.0) &&
                  ((inG[0] && inG[1] && is12[0] && is12[1]) || (inG[0] && inG[2] && is12[0] && is12[2]) ||
                   (inG[1] && inG[2] && is12[1] && is12[2]));
  if (ovr) {
    cxg = pcen - m;
    cmu = pcen;
  }
  const cplx dpair = preal ? cp(ps, 0.0) : cp(0.0, ps);
  cplx off[3];
  for (int i = 0; i < 3; i++) {
    const int neg = (mu[i].re - pcen < 0) || (mu[i].im < 0);
    const cplx di = neg ? cp(-dpair.re, -dpair.im) : dpair;
    const int ovi = ovr && inG[i];
    if (ovi) {
      mu[i] = cp(pcen + di.re, di.im);
      x[i] = cp(mu[i].re - m, mu[i].im);
    }
    off[i] = inG[i] ? (ovi ? di : cp(x[i].re - cxg, x[i].im)) : cp(0.0, 0.0);
  }

  /* single real nodes far from m: Newton on det(A - mu I) */
  for (int i = 0; i < 3; i++) {
    const int mask = !inG[i] && !ab[i] && (x[i].im == 0) && (fabs(x[i].re) > 0.5 * fabs(m));
    if (mask) {
      newton_node(As, mu, i);
      ab[i] = 1;
      x[i] = cp(mu[i].re - m, mu[i].im);
    }
  }
  const int absc = (inG[0] && ab[0]) || (inG[1] && ab[1]) || (inG[2] && ab[2]);
  double apole = 0.0;                          /* pole distance of the node outside a pair group */
  for (int k = 0; k < 3; k++)
    if (pk[k]) apole = (absc || ab[k]) ? cmu - mu[k].re : cxg - x[k].re;
  const double xshift = kind == K_EXP ? dmax(dmax(x[0].re, x[1].re), x[2].re) : 0.0;
  const int bad = kind != K_EXP && (((x[0].im == 0) && (mu[0].re <= 0)) || ((x[1].im == 0) && (mu[1].re <= 0)) ||
                                    ((x[2].im == 0) && (mu[2].re <= 0)));

  /* ---- group part of the divided differences: the Taylor series of the group, streamed over the series index j:
   * sum_j G_j e_j, G = b / (a + u)^(multiplicity of the outside node) by the pole recursion q1..q3, e the dual
   * vectors of the group nodes (e_child_j = e_parent_(j-1) + u e_child_(j-1) if its node is in the group) */
  const int hasg = inG[0] || inG[1] || inG[2];
  const double L = kind == K_EXP ? 1.0 : ((hasg && cmu > 0) ? cmu : 1.0);
  const double iL = 1.0 / L;
  cplx u[3];
  for (int i = 0; i < 3; i++) u[i] = cp_scale(off[i], iL);
  const double ap = fabs(apole) > 0 ? apole * iL : 1.0, iap = 1.0 / ap;
  double rr = dmax(cp_abs(u[0]), dmax(cp_abs(u[1]), cp_abs(u[2])));
  const double rq = ((pk[0] || pk[1] || pk[2]) && (rr * fabs(iap) <= 1.0)) ? rr * fabs(iap) : 0.0;
  rr = (hasg && rr <= R_EXP + 0.5) ? rr : 0.0;
  const int K = series_terms(bt, 3 * (order + 1) - 1, rr, rq);
  double b0 = 1.0;
  if (kind == K_LOG) b0 = node_value(K_LOG, al, sig, m, cp(cxg, 0.0), cp(cmu, 0.0), 0).re;
  cplx Ep[19], Ec[19], acc[19];
  for (int c = 0; c < ND; c++) {
    Ep[c] = inG[ADDED[c]] ? cp(0.0, 0.0) : (PARENT[c] < 0 ? cp(1.0, 0.0) : Ep[PARENT[c]]);
    acc[c] = cp(0.0, 0.0);
  }
  double q1 = 0.0, q2 = 0.0, q3 = 0.0;
  for (int j = 0; j < K; j++) {
    const double bj = j == 0 ? b0 : bt[j];
    const double g1 = (bj - q1) * iap, g2 = (g1 - q2) * iap, g3 = (g2 - q3) * iap;
    q1 = g1; q2 = g2; q3 = g3;
    const double qv[4] = {bj, g1, g2, g3};
    for (int c = 0; c < ND; c++) {
      const int par = PARENT[c], i = ADDED[c];
      const cplx epp = par < 0 ? cp(j == 0 ? 1.0 : 0.0, 0.0) : Ep[par];
      const cplx epc = par < 0 ? cp(0.0, 0.0) : Ec[par];
      Ec[c] = inG[i] ? cp_add(epp, cp_mul(u[i], Ep[c])) : epc;
      const double g = pk[0] ? qv[MULT[c][0]] : (pk[1] ? qv[MULT[c][1]] : (pk[2] ? qv[MULT[c][2]] : qv[0]));
      acc[c] = cp_add(acc[c], cp_scale(Ec[c], g));
    }
    for (int c = 0; c < ND; c++) Ep[c] = Ec[c];
  }
  cplx D[19];
  for (int c = 0; c < ND; c++) {
    const int n = MULT[c][0] + MULT[c][1] + MULT[c][2] - 1, k = MULT[c][0] - 1;
    double fac;
    if (kind == K_EXP) {
      fac = exp(cxg - xshift);
    } else {
      double iLn = 1.0;
      for (int r = 0; r < n - (kind == K_POW ? k : 0); r++) iLn = iLn * iL;
      fac = kind == K_LOG ? iLn : powk(sig * L, al, (double)k) * iLn;
    }
    D[c] = hasg ? cp_scale(acc[c], fac) : cp(0.0, 0.0);
  }
  /* ---- residues of the single nodes i: coefficient of (t - mu_i)^(m_i - 1) in f(t) / prod_(z != i) (t - mu_z)^m_z */
  for (int i = 0; i < 3; i++) {
    if (inG[i]) continue;
    const cplx xs = kind == K_EXP ? cp(x[i].re - xshift, x[i].im) : x[i];
    cplx f0[3];
    f0[0] = node_value(kind, al, sig, m, xs, mu[i], 0);
    if (kind == K_POW && order >= 1) f0[1] = node_value(kind, al, sig, m, xs, mu[i], 1);
    if (kind == K_POW && order >= 2) f0[2] = node_value(kind, al, sig, m, xs, mu[i], 2);
    const cplx imu = cp_inv(mu[i]);
    cplx w[3];
    for (int z = 0; z < 3; z++)
      if (z != i) w[z] = cp_inv((ab[i] || ab[z]) ? cp_sub(mu[i], mu[z]) : cp_sub(x[i], x[z]));
    for (int c = 0; c < ND; c++) {
      const int mi = MULT[c][i];
      if (mi == 0) continue;
      const int k = MULT[c][0] - 1;
      const cplx fb = kind == K_POW ? f0[k] : f0[0];
      cplx ser[3];                             /* Taylor coefficients 0..2 of f at mu_i */
      if (kind == K_EXP) {
        ser[0] = fb; ser[1] = fb; ser[2] = cp_scale(fb, 0.5);
      } else if (kind == K_LOG) {
        ser[0] = fb; ser[1] = imu; ser[2] = cp_scale(cp_mul(imu, imu), -0.5);
      } else {
        ser[0] = fb;
        ser[1] = cp_scale(cp_mul(fb, imu), al);
        ser[2] = cp_scale(cp_mul(ser[1], imu), (al - 1.0) / 2.0);
      }
      for (int z = 0; z < 3; z++) {            /* times the series of 1 / (d + tau)^m_z */
        const int mz = MULT[c][z];
        if (z == i || mz == 0) continue;
        cplx wp = cp(1.0, 0.0);
        for (int r = 0; r < mz; r++) wp = cp_mul(wp, w[z]);
        const cplx p1 = cp_scale(cp_mul(wp, w[z]), -(double)mz);
        const cplx p2 = cp_scale(cp_mul(cp_mul(wp, w[z]), w[z]), mz * (mz + 1) / 2.0);
        const cplx n0 = cp_mul(ser[0], wp), n1 = cp_add(cp_mul(ser[1], wp), cp_mul(ser[0], p1));
        const cplx n2 = cp_add(cp_add(cp_mul(ser[2], wp), cp_mul(ser[1], p1)), cp_mul(ser[0], p2));
        ser[0] = n0; ser[1] = n1; ser[2] = n2;
      }
      D[c] = cp_add(D[c], ser[mi - 1]);
    }
  }

  /* ---- real basis w = T r: r_0 = I, r_1 = B_0, r_2 = B_0 B_1 - b0 b1 I, T = I + i N (gmf_kernel.hpp:578-599) */
  const double b0i = x[0].im, b1i = x[1].im;
  double B0[9], B1[9], r[3][9];
  for (int k = 0; k < 9; k++) {
    const int dg = k % 4 == 0;
    B0[k] = dg ? (ab[0] ? As[k] - mu[0].re : S[k] - x[0].re) : S[k];
    B1[k] = dg ? (ab[1] ? As[k] - mu[1].re : S[k] - x[1].re) : S[k];
    r[0][k] = dg ? 1.0 : 0.0;
    r[1][k] = B0[k];
  }
  for (int i2 = 0; i2 < 3; i2++)
    for (int j2 = 0; j2 < 3; j2++)
      r[2][3 * i2 + j2] = B0[3 * i2] * B1[j2] + B0[3 * i2 + 1] * B1[3 + j2] + B0[3 * i2 + 2] * B1[6 + j2] -
                          (i2 == j2 ? b0i * b1i : 0.0);
  const double a01 = (ab[0] || ab[1]) ? mu[0].re - mu[1].re : x[0].re - x[1].re;
  const double N10 = -b0i, N20 = -b0i * a01, N21 = -(b0i + b1i);
  const double nanf = bad ? NAN : 1.0;
#define NT(v, o) do { (o)[0] = N10 * (v)[1] + N20 * (v)[2]; (o)[1] = N21 * (v)[2]; (o)[2] = 0.0; } while (0)
  {
    double Dr[3], Di[3], Nr[3], Ni[3];
    for (int j = 0; j < 3; j++) { Dr[j] = D[pos1(j)].re; Di[j] = D[pos1(j)].im; }
    NT(Dr, Nr); NT(Di, Ni);
    (void)Nr;
    const double e = kind == K_EXP ? exp(m + xshift) : 1.0;
    for (int k = 0; k < 9; k++)
      F[k] = ((Dr[0] - Ni[0]) * r[0][k] + (Dr[1] - Ni[1]) * r[1][k] + (Dr[2] - Ni[2]) * r[2][k]) * e * nanf;
  }
  if (order < 1) return K;
  const double sc1 = (kind == K_EXP ? exp(m + xshift) : (kind == K_LOG ? isig : 1.0)) * nanf;
  double D2[9];
  {
    double Dr[9], Di[9];
    for (int j = 0; j < 3; j++)
      for (int k = 0; k < 3; k++) { Dr[3 * j + k] = D[pos2(j, k)].re; Di[3 * j + k] = D[pos2(j, k)].im; }
    for (int k = 0; k < 3; k++) {                /* T on index 0: (Dr, Di) <- (Dr - N^T Di, Di + N^T Dr) */
      const double vr[3] = {Dr[k], Dr[3 + k], Dr[6 + k]}, vi[3] = {Di[k], Di[3 + k], Di[6 + k]};
      double nr[3], ni[3];
      NT(vr, nr); NT(vi, ni);
      for (int j = 0; j < 3; j++) { Dr[3 * j + k] = vr[j] - ni[j]; Di[3 * j + k] = vi[j] + nr[j]; }
    }
    for (int j = 0; j < 3; j++) {                /* T on index 1, real part only */
      double ni[3];
      NT(Di + 3 * j, ni);
      for (int k = 0; k < 3; k++) D2[3 * j + k] = Dr[3 * j + k] - ni[k];
    }
  }
  /* J_abcd = sum_mn D2_mn r_m[a][c] r_n[d][b] */
  double Y[3][9];
  for (int mm = 0; mm < 3; mm++)
    for (int k = 0; k < 9; k++) Y[mm][k] = D2[3 * mm] * r[0][k] + D2[3 * mm + 1] * r[1][k] + D2[3 * mm + 2] * r[2][k];
  for (int aa = 0; aa < 3; aa++)
    for (int bb = 0; bb < 3; bb++)
      for (int cc = 0; cc < 3; cc++)
        for (int dd = 0; dd < 3; dd++)
          J[((3 * aa + bb) * 3 + cc) * 3 + dd] = (r[0][3 * aa + cc] * Y[0][3 * dd + bb] + r[1][3 * aa + cc] * Y[1][3 * dd + bb] +
                                                  r[2][3 * aa + cc] * Y[2][3 * dd + bb]) * sc1;
  if (order < 2) return K;
  const double sc2 = (kind == K_EXP ? exp(m + xshift) : (kind == K_LOG ? isig * isig : 1.0)) * nanf;
  double D3[27];
  {
    double Dr[27], Di[27];
    for (int j = 0; j < 3; j++)
      for (int k = 0; k < 3; k++)
        for (int l = 0; l < 3; l++) {
          Dr[9 * j + 3 * k + l] = D[pos3(j, k, l)].re;
          Di[9 * j + 3 * k + l] = D[pos3(j, k, l)].im;
        }
    static const int STR[3] = {9, 3, 1};
    for (int ax = 0; ax < 3; ax++) {             /* T on each index */
      const int s = STR[ax];
      for (int base = 0; base < 27; base++) {
        if ((base / s) % 3 != 0) continue;
        const double vr[3] = {Dr[base], Dr[base + s], Dr[base + 2 * s]};
        const double vi[3] = {Di[base], Di[base + s], Di[base + 2 * s]};
        double nr[3], ni[3];
        NT(vr, nr); NT(vi, ni);
        for (int j = 0; j < 3; j++) { Dr[base + j * s] = vr[j] - ni[j]; Di[base + j * s] = vi[j] + nr[j]; }
      }
    }
    for (int k = 0; k < 27; k++) D3[k] = Dr[k];
  }
#undef NT
  /* H_abcdef = T6_acdefb + T6_aefcdb, T6_acdefb = sum_mnp D3_mnp r_m[a][c] r_n[d][e] r_p[f][b] */
  double Z[9][9], Yh[3][81];
  for (int mn = 0; mn < 9; mn++)
    for (int k = 0; k < 9; k++) Z[mn][k] = D3[3 * mn] * r[0][k] + D3[3 * mn + 1] * r[1][k] + D3[3 * mn + 2] * r[2][k];
  for (int mm = 0; mm < 3; mm++)
    for (int de = 0; de < 9; de++)
      for (int fb = 0; fb < 9; fb++)
        Yh[mm][9 * de + fb] = r[0][de] * Z[3 * mm][fb] + r[1][de] * Z[3 * mm + 1][fb] + r[2][de] * Z[3 * mm + 2][fb];
  for (int aa = 0; aa < 3; aa++)
    for (int bb = 0; bb < 3; bb++)
      for (int cc = 0; cc < 3; cc++)
        for (int dd = 0; dd < 3; dd++)
          for (int ee = 0; ee < 3; ee++)
            for (int ff = 0; ff < 3; ff++) {
              const int i1 = 9 * (3 * dd + ee) + 3 * ff + bb, i2 = 9 * (3 * ff + cc) + 3 * dd + bb;
              const double h = r[0][3 * aa + cc] * Yh[0][i1] + r[1][3 * aa + cc] * Yh[1][i1] + r[2][3 * aa + cc] * Yh[2][i1] +
                               r[0][3 * aa + ee] * Yh[0][i2] + r[1][3 * aa + ee] * Yh[1][i2] + r[2][3 * aa + ee] * Yh[2][i2];
              H[((((3 * aa + bb) * 3 + cc) * 3 + dd) * 3 + ee) * 3 + ff] = h * sc2;
            }
  return K;
}

int gmf_eval_one(int kind, int order, double al, const double *A, double *F, double *J, double *H, const double *bt) {
  return gmf_core(kind, order, al, A, F, J, H, bt);
}
