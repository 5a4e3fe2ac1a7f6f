/* smf_oracle.c -- symmetric 3x3 matrices: f(A) = V f(Lambda) V^T and its first and second derivatives by the
 * Daleckii-Krein formulas, for f = exp, log, sqrt, x^alpha.  Scalar FP64 restatement of the reference kernel
 * c/include/matfun/smf_kernel.hpp (eval3, divdiff, deriv, divdiff2, dd2_coef, to_matrix, from_matrix) with its
 * thresholds and branch choices; branches are ordinary `if`s here (the reference evaluates both sides and selects).
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Voigt index of the symmetric pair (a, b): 0,1,2 on the diagonal, (0,1)->3, (0,2)->4, (1,2)->5.
 */
#include <math.h>
#include <string.h>

#include "oracle_internal.h"

static int voigt_of(int a, int b) { return a == b ? a : a + b + 2; }      /* a != b: 3, 4, 5 */
static const int VROW[6] = {0, 1, 2, 0, 0, 1}, VCOL[6] = {0, 1, 2, 1, 2, 2};

/* f(x): exp, log, and the powers with x <= 0 (round-off of a PSD matrix) mapped to 0 -- smf_kernel.hpp:75-82 */
static double fval(int kind, double x, double al) {
  if (kind == K_EXP) return exp(x);
  if (kind == K_LOG) return log(x);
  if (!(x > 0)) return 0.0;
  return kind == K_SQRT ? sqrt(x) : pow(x, al);
}

/* first divided difference f[a, b], stable for a ~ b -- smf_kernel.hpp:84-115 */
static double dd1(int kind, double a, double b, double fa, double fb, double al) {
  if (kind == K_EXP) {                       /* e^max (1 - e^-d) / d, d = |a - b| */
    const double d = fabs(a - b);
    const double r = d > 0 ? -expm1(-d) / d : 1.0;
    return (a > b ? fa : fb) * r;
  }
  if (kind == K_LOG) {                       /* 2 artanh(z) / (z (a + b)), z = (a - b) / (a + b) */
    const double s = 0.5 * (a + b), z = (a - b) / (a + b), az = fabs(z), z2 = z * z;
    if (az < 1e-3) return (1.0 + z2 / 3.0 + z2 * z2 / 5.0 + z2 * z2 * z2 / 7.0) / s;
    if (az >= 0.5) return (log(a) - log(b)) / (a - b);
    return atanh(z) / z / s;
  }
  a = a > 0 ? a : 0.0;                       /* powers: clamp round-off below 0 (NaN -> 0 as well) */
  b = b > 0 ? b : 0.0;
  if (kind == K_SQRT) return 1.0 / (sqrt(a) + sqrt(b));
  const double s = 0.5 * (a + b), ss = s > 0 ? s : 1.0;
  const double z = (a - b) / (2.0 * ss), az = fabs(z);
  if (az == 0) return (s > 0 ? powk(ss, al, 1.0) : pow_of_zero(al - 1.0)) * al;
  if (az < 0.5) {                            /* s^(al-1) ((1+z)^al - (1-z)^al) / (2 z) via expm1 / log1p */
    const double g = (expm1(al * log1p(z)) - expm1(al * log1p(-z))) / (2.0 * z);
    return (s > 0 ? powk(ss, al, 1.0) : pow_of_zero(al - 1.0)) * g;
  }
  return (fa - fb) / (a - b);
}

/* f'(x) = f[x, x] -- smf_kernel.hpp:117-126 */
static double dfun(int kind, double x, double fx, double al) {
  if (kind == K_EXP) return fx;
  if (kind == K_LOG) return 1.0 / x;
  x = x > 0 ? x : 0.0;
  if (kind == K_SQRT) return 1.0 / (2.0 * sqrt(x));
  return x > 0 ? al * powk(x, al, 1.0) : al * pow_of_zero(al - 1.0);
}

/* Taylor coefficients f^(k+2)(m) / (k+2)! without the m-dependent factor (e^m, 1/m^2, m^(al-2)) --
 * smf_kernel.hpp:135-146: exp 1/(k+2)!, log (-1)^(k+1)/(k+2), powers binom(al, k+2) */
void smf_dd2_coefs(int kind, double al, double *c) {
  if (kind == K_EXP) {
    double f = 0.5;
    for (int k = 0; k < SMF_NTERMS; k++) { c[k] = f; f /= (k + 3.0); }
  } else if (kind == K_LOG) {
    for (int k = 0; k < SMF_NTERMS; k++) c[k] = ((k & 1) ? 1.0 : -1.0) / (k + 2.0);
  } else {
    double b = al * (al - 1.0) / 2.0;
    for (int k = 0; k < SMF_NTERMS; k++) { c[k] = b; b *= (al - k - 2.0) / (k + 3.0); }
  }
}

/* second divided difference f[li, lj, lk] -- smf_kernel.hpp:148-192: all three points in a cluster around the
 * middle one -> 30-term Taylor series sum_k c_k h_k(lo - m, hi - m); otherwise the difference formula over the
 * most distant pair, from the first divided differences */
static double dd2(int kind, double li, double lj, double lk, double fi, double fj, double fk, double Dij, double Dik,
                  double Djk, double al, const double *c) {
  const int powlike = kind == K_POW || kind == K_SQRT;
  if (powlike) {
    li = li > 0 ? li : 0.0;
    lj = lj > 0 ? lj : 0.0;
    lk = lk > 0 ? lk : 0.0;
  }
  const double dA = lk - li, dB = lj - li, dC = lk - lj;
  const double aA = fabs(dA), aB = fabs(dB), aC = fabs(dC);
  double num, den;
  if (aA >= aB && aA >= aC) { num = Djk - Dij; den = dA; }
  else if (aB >= aC) { num = Djk - Dik; den = dB; }
  else { num = Dik - Dij; den = dC; }
  const double direct = num / (den != 0 ? den : 1.0);
  /* order statistics (ties keep the reference's choice of which f to take for the middle point) */
  const double mn = li < lj ? li : lj, mx = li < lj ? lj : li;
  const double lo = mn < lk ? mn : lk, hi = mx > lk ? mx : lk;
  const double mxk = mx < lk ? mx : lk;
  const double mid = mn > mxk ? mn : mxk;
  const double fm = mid == li ? fi : (mid == lj ? fj : fk);
  int cluster;
  double u, w, scale;
  if (kind == K_EXP) {
    cluster = (hi - lo) <= 1.0;
    u = lo - mid;
    w = hi - mid;
    scale = fm;
  } else {
    /* (|al|: the per-matrix exponent extension also takes al <= 0; for al > 0 this is smf_kernel.hpp:173) */
    const double tau = (kind == K_POW && 3.0 / fabs(al) < 0.25) ? 3.0 / fabs(al) : 0.25;
    cluster = (mid > 0) && (mid - lo <= tau * mid) && (hi - mid <= tau * mid);
    const double ms = cluster ? mid : 1.0;
    u = (lo - mid) / ms;
    w = (hi - mid) / ms;
    scale = (kind == K_LOG ? 1.0 : (cluster ? fm : 1.0)) / (ms * ms);
  }
  double out = direct;
  if (cluster) {                  /* h_k(u, w) = sum_i u^i w^(k-i) by h_k = u h_(k-1) + w^k */
    double h = 1.0, wk = 1.0, s = c[0];
    for (int k = 1; k < SMF_NTERMS; k++) {
      wk = wk * w;
      h = u * h + wk;
      s = s + c[k] * h;
    }
    out = scale * s;
  }
  if (powlike && !(hi > 0)) out = al == 1.0 ? 0.0 : al * (al - 1.0) / 2.0 * pow_of_zero(al - 2.0);
  return out;
}

/* ---- one matrix: eigen-decomposition (smf_kernel.hpp:202-280), F (:282-295), J (:296-327), H (:328-381) */
void smf_eval_one(int kind, double al, const double *v, double *F, double *J, double *H, const double *c2) {
  const double SQRT6 = 2.449489742783178, TWO_PI_3 = 2.0943951023931957;
  /* exact power-of-two scaling sig ~ max |A_ij| */
  /* (pairwise max tree as in the reference: a NaN entry drops out of the max and shows up in the results) */
  const double ab0 = fabs(v[0]), ab1 = fabs(v[1]), ab2 = fabs(v[2]), ab3 = fabs(v[3]), ab4 = fabs(v[4]),
               ab5 = fabs(v[5]);
  const double s01 = ab0 > ab1 ? ab0 : ab1, s23 = ab2 > ab3 ? ab2 : ab3, s45 = ab4 > ab5 ? ab4 : ab5;
  const double s03 = s01 > s23 ? s01 : s23;
  const double sig = pow2_floor((s03 > s45 ? s03 : s45) + 1e-300), isig = 1.0 / sig;
  const double a0 = v[0] * isig, a1 = v[1] * isig, a2 = v[2] * isig, a3 = v[3] * isig, a4 = v[4] * isig,
               a5 = v[5] * isig;
  /* deviator A - m I without its trace residue, normalised by a power of two */
  const double m = (a0 + a1 + a2) / 3.0, d0 = a0 - m, d1 = a1 - m, d2 = a2 - m;
  const double tr3 = (d0 + d1 + d2) / 3.0;
  const double e0 = d0 - tr3, e1 = d1 - tr3, e2 = d2 - tr3;
  const double z01 = fabs(e0) > fabs(e1) ? fabs(e0) : fabs(e1), z23 = fabs(e2) > fabs(a3) ? fabs(e2) : fabs(a3);
  const double z45 = fabs(a4) > fabs(a5) ? fabs(a4) : fabs(a5), z03 = z01 > z23 ? z01 : z23;
  const double dmx = z03 > z45 ? z03 : z45;
  const int iso = dmx <= 1e-300;
  const double ids = 1.0 / (iso ? 1.0 : pow2_floor(dmx));
  const double g0 = e0 * ids, g1 = e1 * ids, g2 = e2 * ids, g3 = a3 * ids, g4 = a4 * ids, g5 = a5 * ids;
  /* simple eigenvalue sh of the normalised deviator: trigonometric root of x^3 - (p/2) x - q = 0 */
  const double p = g0 * g0 + g1 * g1 + g2 * g2 + 2.0 * (g3 * g3 + g4 * g4 + g5 * g5);
  const double q = g0 * (g1 * g2 - g5 * g5) - g3 * (g3 * g2 - g5 * g4) + g4 * (g3 * g5 - g1 * g4);
  const double pg = iso ? 1.0 : p, qg = iso ? 0.0 : q;
  double r = 3.0 * SQRT6 * qg / (pg * sqrt(pg));
  r = r > 1.0 ? 1.0 : (r < -1.0 ? -1.0 : r);
  const double t = sqrt(2.0 * pg / 3.0), ph = acos(r) / 3.0;
  const double sh = t * cos(r >= 0 ? ph : ph + TWO_PI_3);
  /* its eigenvector: the longest cross product of two rows of (G - sh I) */
  const double m00 = g0 - sh, m11 = g1 - sh, m22 = g2 - sh;
  const double c0x = g3 * g5 - g4 * m11, c0y = g4 * g3 - m00 * g5, c0z = m00 * m11 - g3 * g3;
  const double c1x = g3 * m22 - g4 * g5, c1y = g4 * g4 - m00 * m22, c1z = m00 * g5 - g3 * g4;
  const double c2x = m11 * m22 - g5 * g5, c2y = g5 * g4 - g3 * m22, c2z = g3 * g5 - m11 * g4;
  const double n0 = c0x * c0x + c0y * c0y + c0z * c0z, n1 = c1x * c1x + c1y * c1y + c1z * c1z,
               n2 = c2x * c2x + c2y * c2y + c2z * c2z;
  double cx = c0x, cy = c0y, cz = c0z, nn = n0;
  if (n1 > nn) { cx = c1x; cy = c1y; cz = c1z; nn = n1; }
  if (n2 > nn) { cx = c2x; cy = c2y; cz = c2z; nn = n2; }
  const double nrm = sqrt(nn);
  const int ok = (nrm > 1e-300) && !iso;
  double vx = 1.0, vy = 0.0, vz = 0.0;
  if (ok) {
    const double inv = 1.0 / nrm;
    vx = cx * inv; vy = cy * inv; vz = cz * inv;
  }
  /* orthonormal complement (u, w): u from the coordinate axis least aligned with v */
  const double ax = fabs(vx), ay = fabs(vy), az = fabs(vz);
  const int k1 = ay < ax, k2 = az < (k1 ? ay : ax);
  const double ex = (k1 || k2) ? 0.0 : 1.0, ey = (k1 && !k2) ? 1.0 : 0.0, ez = k2 ? 1.0 : 0.0;
  const double vk = ex * vx + ey * vy + ez * vz;
  double ux = ex - vk * vx, uy = ey - vk * vy, uz = ez - vk * vz;
  const double nu = sqrt(ux * ux + uy * uy + uz * uz);
  if (nu > 1e-300) {
    const double iu = 1.0 / nu;
    ux = ux * iu; uy = uy * iu; uz = uz * iu;
  } else {
    ux = uy = uz = 0.0;
  }
  const double wx = vy * uz - vz * uy, wy = vz * ux - vx * uz, wz = vx * uy - vy * ux;
  /* Rayleigh quotients of the scaled A (eigenvalues) and of the deviator (its angle and x = lambda - m) */
  const double Aux = a0 * ux + a3 * uy + a4 * uz, Auy = a3 * ux + a1 * uy + a5 * uz, Auz = a4 * ux + a5 * uy + a2 * uz;
  const double Awx = a0 * wx + a3 * wy + a4 * wz, Awy = a3 * wx + a1 * wy + a5 * wz, Awz = a4 * wx + a5 * wy + a2 * wz;
  const double Avx = a0 * vx + a3 * vy + a4 * vz, Avy = a3 * vx + a1 * vy + a5 * vz, Avz = a4 * vx + a5 * vy + a2 * vz;
  const double a11 = ux * Aux + uy * Auy + uz * Auz, a12 = ux * Awx + uy * Awy + uz * Awz;
  const double a22 = wx * Awx + wy * Awy + wz * Awz, lv = vx * Avx + vy * Avy + vz * Avz;
  const double Dux = d0 * ux + a3 * uy + a4 * uz, Duy = a3 * ux + d1 * uy + a5 * uz, Duz = a4 * ux + a5 * uy + d2 * uz;
  const double Dwx = d0 * wx + a3 * wy + a4 * wz, Dwy = a3 * wx + d1 * wy + a5 * wz, Dwz = a4 * wx + a5 * wy + d2 * wz;
  const double b11 = ux * Dux + uy * Duy + uz * Duz, b12 = ux * Dwx + uy * Dwy + uz * Dwz,
               b22 = wx * Dwx + wy * Dwy + wz * Dwz;
  /* 2x2 Jacobi rotation of the deviator block, cos / sin of half the angle from half-angle formulas */
  const double jx = b11 - b22, jy = 2.0 * b12, jr = sqrt(jx * jx + jy * jy);
  double cs = 1.0, sn = 0.0;
  if (!(jr == 0)) {
    const double big = sqrt((jr + fabs(jx)) / (2.0 * jr));
    const double sml = jy / (2.0 * jr * big);
    const double sgn = ((jy > 0) || ((jy == 0) && !signbit(jy))) ? 1.0 : -1.0;
    if (jx >= 0) { cs = big; sn = sml; }
    else { cs = sgn * sml; sn = sgn * big; }
  }
  const double lam[3] = {lv * sig, (cs * cs * a11 + 2.0 * cs * sn * a12 + sn * sn * a22) * sig,
                         (sn * sn * a11 - 2.0 * cs * sn * a12 + cs * cs * a22) * sig};
  const double V[3][3] = {{vx, cs * ux + sn * wx, cs * wx - sn * ux},
                          {vy, cs * uy + sn * wy, cs * wy - sn * uy},
                          {vz, cs * uz + sn * wz, cs * wz - sn * uz}};
  /* F = V diag f(lambda) V^T; log near m: log(m sig) + log1p(x / m) with x from the deviator */
  double fl[3];
  for (int i = 0; i < 3; i++) fl[i] = fval(kind, lam[i], al);
  if (kind == K_LOG) {
    const double bv = vx * (d0 * vx + a3 * vy + a4 * vz) + vy * (a3 * vx + d1 * vy + a5 * vz) +
                      vz * (a4 * vx + a5 * vy + d2 * vz);
    const double xd[3] = {bv, cs * cs * b11 + 2.0 * cs * sn * b12 + sn * sn * b22,
                          sn * sn * b11 - 2.0 * cs * sn * b12 + cs * cs * b22};
    const double lms = log(m * sig);
    for (int i = 0; i < 3; i++)
      if (fabs(xd[i]) <= 0.5 * m) fl[i] = lms + log1p(xd[i] / m);
  }
  for (int x = 0; x < 3; x++)
    for (int y = x; y < 3; y++)
      F[voigt_of(x, y)] = V[x][0] * fl[0] * V[y][0] + V[x][1] * fl[1] * V[y][1] + V[x][2] * fl[2] * V[y][2];
  if (!J) return;
  /* first order: D_ij = f[lam_i, lam_j]; J_(xy),(pq) = sum_i P_i[x][p] Q_i[y][q] (+ the (q, p) term for p != q),
   * P_i = v_i v_i^T, Q_i = sum_j D_ij P_j */
  const double D01 = dd1(kind, lam[0], lam[1], fl[0], fl[1], al), D02 = dd1(kind, lam[0], lam[2], fl[0], fl[2], al);
  const double D12 = dd1(kind, lam[1], lam[2], fl[1], fl[2], al);
  const double D[3][3] = {{dfun(kind, lam[0], fl[0], al), D01, D02},
                          {D01, dfun(kind, lam[1], fl[1], al), D12},
                          {D02, D12, dfun(kind, lam[2], fl[2], al)}};
  double P[3][3][3], Q[3][3][3];
  for (int i = 0; i < 3; i++)
    for (int x = 0; x < 3; x++)
      for (int y = 0; y < 3; y++) P[i][x][y] = V[x][i] * V[y][i];
  for (int i = 0; i < 3; i++)
    for (int x = 0; x < 3; x++)
      for (int y = 0; y < 3; y++) Q[i][x][y] = D[i][0] * P[0][x][y] + D[i][1] * P[1][x][y] + D[i][2] * P[2][x][y];
  for (int x = 0; x < 3; x++)
    for (int y = x; y < 3; y++)
      for (int k = 0; k < 6; k++) {
        const int pp = VROW[k], qq = VCOL[k];
        double s = P[0][x][pp] * Q[0][y][qq] + P[1][x][pp] * Q[1][y][qq] + P[2][x][pp] * Q[2][y][qq];
        if (pp != qq) s += P[0][x][qq] * Q[0][y][pp] + P[1][x][qq] * Q[1][y][pp] + P[2][x][qq] * Q[2][y][pp];
        J[6 * voigt_of(x, y) + k] = s;
      }
  if (!H) return;
  /* second order: H^(pq) = V S V^T, S = X + X^T, X_ab = sum_m D2_amb Et^p_am Et^q_mb, Et^p = V^T E_p V */
  double D2[3][3][3];
  for (int i = 0; i < 3; i++)
    for (int j = i; j < 3; j++)
      for (int k = j; k < 3; k++) {
        const double d = dd2(kind, lam[i], lam[j], lam[k], fl[i], fl[j], fl[k], D[i][j], D[i][k], D[j][k], al, c2);
        D2[i][j][k] = D2[i][k][j] = D2[j][i][k] = D2[j][k][i] = D2[k][i][j] = D2[k][j][i] = d;
      }
  double Et[6][3][3];
  for (int pv = 0; pv < 6; pv++)
    for (int a = 0; a < 3; a++)
      for (int mm = a; mm < 3; mm++) {
        const int k = VROW[pv], l = VCOL[pv];
        double e = V[k][a] * V[l][mm];
        if (k != l) e += V[l][a] * V[k][mm];
        Et[pv][a][mm] = Et[pv][mm][a] = e;
      }
  for (int pv = 0; pv < 6; pv++)
    for (int qv = pv; qv < 6; qv++) {
      double S[3][3], W[3][3];
      for (int a = 0; a < 3; a++)
        for (int b = a; b < 3; b++) {
          const double xab = D2[a][0][b] * Et[pv][a][0] * Et[qv][0][b] + D2[a][1][b] * Et[pv][a][1] * Et[qv][1][b] +
                             D2[a][2][b] * Et[pv][a][2] * Et[qv][2][b];
          const double xba = D2[b][0][a] * Et[pv][b][0] * Et[qv][0][a] + D2[b][1][a] * Et[pv][b][1] * Et[qv][1][a] +
                             D2[b][2][a] * Et[pv][b][2] * Et[qv][2][a];
          S[a][b] = S[b][a] = xab + xba;
        }
      for (int a = 0; a < 3; a++)
        for (int j = 0; j < 3; j++) W[a][j] = S[a][0] * V[j][0] + S[a][1] * V[j][1] + S[a][2] * V[j][2];
      for (int i = 0; i < 3; i++)
        for (int j = i; j < 3; j++) {
          const double h = V[i][0] * W[0][j] + V[i][1] * W[1][j] + V[i][2] * W[2][j];
          H[(6 * voigt_of(i, j) + pv) * 6 + qv] = H[(6 * voigt_of(i, j) + qv) * 6 + pv] = h;
        }
    }
}

/* ---- full-matrix layout (smf_kernel.hpp:384-421): m = 3a + b; Voigt index of m, and d v / d A_m (1 or 1/2) */
static int vmat(int mm) { return mm % 4 == 0 ? mm / 4 : ((mm == 1 || mm == 3) ? 3 : ((mm == 2 || mm == 6) ? 4 : 5)); }
static double wmat(int mm) { return mm % 4 == 0 ? 1.0 : 0.5; }

void smf_voigt_to_matrix(const double *Fv, const double *Jv, const double *Hv, double *F, double *J, double *H) {
  for (int mm = 0; mm < 9; mm++) F[mm] = Fv[vmat(mm)];
  if (J)
    for (int mm = 0; mm < 9; mm++)
      for (int c = 0; c < 9; c++) J[9 * mm + c] = Jv[6 * vmat(mm) + vmat(c)] * wmat(c);
  if (H)
    for (int mm = 0; mm < 9; mm++)
      for (int c = 0; c < 9; c++)
        for (int e = 0; e < 9; e++) H[(9 * mm + c) * 9 + e] = Hv[(6 * vmat(mm) + vmat(c)) * 6 + vmat(e)] * (wmat(c) * wmat(e));
}

/* Voigt vector of (A + A^T) / 2; 0 and NaN if |a_ij - a_ji| > 1e-8 max |a_kl| (NaN / Inf entries pass) */
int smf_matrix_to_voigt(const double *a, double *v) {
  double mx = 0.0;
  for (int k = 0; k < 9; k++) mx = fabs(a[k]) > mx ? fabs(a[k]) : mx;
  const double tol = 1e-8 * mx;
  const int sym = !(fabs(a[1] - a[3]) > tol) && !(fabs(a[2] - a[6]) > tol) && !(fabs(a[5] - a[7]) > tol);
  v[0] = a[0]; v[1] = a[4]; v[2] = a[8];
  v[3] = 0.5 * (a[1] + a[3]); v[4] = 0.5 * (a[2] + a[6]); v[5] = 0.5 * (a[5] + a[7]);
  if (!sym)
    for (int k = 0; k < 6; k++) v[k] = NAN;
  return sym;
}
