// smf_device.hpp -- per-lane math of the symmetric kernels: closed-form eigen-decomposition of a symmetric 3x3
// matrix, f(A), the Daleckii-Krein first derivative as its 21 distinct symmetric entries, and the second derivative
// one Voigt row at a time.  Same method, thresholds and branch structure as the reference kernel
// c/include/matfun/smf_kernel.hpp (eval3 :200-382, divdiff :84-115, deriv :117-126, divdiff2 :148-192); the
// second derivative is assembled one output row at a time for the GPU (SmfHess::h_row).
#pragma once
#include "mf_device.hpp"

namespace mf {

constexpr int SMF_NTERMS = 30;   // terms of the second-divided-difference Taylor series (tau^30 * 31 < 1e-16)
struct SmfCoef {                  // f^(k+2)(m) / (k+2)! without the m-dependent factor, built on the host per call
  double c[SMF_NTERMS];
};
// exp 1/(k+2)!, log (-1)^(k+1)/(k+2), powers binom(al, k+2)  (smf_kernel.hpp:135-146)
inline void smf_coefs(int kind, double al, SmfCoef& c) {
  if (kind == EXP) {
    double f = 0.5;
    for (int k = 0; k < SMF_NTERMS; k++) { c.c[k] = f; f /= (k + 3.0); }
  } else if (kind == LOG) {
    for (int k = 0; k < SMF_NTERMS; k++) c.c[k] = ((k & 1) ? 1.0 : -1.0) / (k + 2.0);
  } else {
    double b = al * (al - 1.0) / 2.0;
    for (int k = 0; k < SMF_NTERMS; k++) { c.c[k] = b; b *= (al - k - 2.0) / (k + 3.0); }
  }
}

// Voigt index pairs: k -> (row, col) of the symmetric 3x3 matrix
MF_DEV constexpr int vrow(int k) { return k < 3 ? k : (k == 5 ? 1 : 0); }
MF_DEV constexpr int vcol(int k) { return k < 3 ? k : (k == 3 ? 1 : 2); }
MF_DEV constexpr int vidx(int a, int b) { return a == b ? a : a + b + 2; }
MF_DEV constexpr int tri3(int a, int b) { return a <= b ? vidx(a, b) : vidx(b, a); }   // packed symmetric 3x3

// f(x); the powers map round-off below 0 (and NaN) to 0 (smf_kernel.hpp:75-82)
template <int KIND>
MF_DEV double fval(double x, double al) {
  if (KIND == EXP) return exp(x);
  if (KIND == LOG) return log(x);
  if (!(x > 0)) return 0.0;
  return KIND == SQRT ? sqrt(x) : pow(x, al);
}

// f[a, b], stable for a ~ b (smf_kernel.hpp:84-115)
template <int KIND>
MF_DEV double dd1(double a, double b, double fa, double fb, double al) {
  if (KIND == EXP) {
    const double d = fabs(a - b);
    return (a > b ? fa : fb) * (d > 0 ? -expm1(-d) / d : 1.0);
  }
  if (KIND == LOG) {
    const double s = 0.5 * (a + b), z = (a - b) / (a + b), az = fabs(z), z2 = z * z;
    if (az < 1e-3) return (1.0 + z2 / 3.0 + z2 * z2 / 5.0 + z2 * z2 * z2 / 7.0) / s;
    if (az >= 0.5) return (log(a) - log(b)) / (a - b);
    return atanh(z) / z / s;
  }
  a = a > 0 ? a : 0.0;
  b = b > 0 ? b : 0.0;
  if (KIND == SQRT) return 1.0 / (sqrt(a) + sqrt(b));
  const double s = 0.5 * (a + b), ss = s > 0 ? s : 1.0;
  const double z = (a - b) / (2.0 * ss), az = fabs(z);
  if (az >= 0.5 || !(az == az)) return (fa - fb) / (a - b);
  const double g = az == 0 ? al : (expm1(al * log1p(z)) - expm1(al * log1p(-z))) / (2.0 * z);
  return (s > 0 ? powk(ss, al, 1.0) : pow_of_zero(al - 1.0)) * g;
}

// f'(x) (smf_kernel.hpp:117-126)
template <int KIND>
MF_DEV double dfun(double x, double fx, double al) {
  if (KIND == EXP) return fx;
  if (KIND == LOG) return 1.0 / x;
  x = x > 0 ? x : 0.0;
  if (KIND == SQRT) return 1.0 / (2.0 * sqrt(x));
  return x > 0 ? al * powk(x, al, 1.0) : al * pow_of_zero(al - 1.0);
}

// f[x, y, z] (smf_kernel.hpp:148-192): Taylor series around the middle point inside a cluster, else the difference
// formula over the most distant pair.  PA (powers with a per-matrix exponent): the series coefficients
// binom(al, k + 2) are formed in the loop from binom(al, 2) = al (al - 1) / 2 by c_k = c_(k-1) (al - k - 1) / (k + 2),
// the recurrence of smf_coefs, and c is not read.  NT: series terms (30 as the reference for FP64 results; the mixed
// form's FP32 results take 20: the remainder is below 21 (1/4)^20 < 2e-11 of the sum for tau = 1/4, below
// 3^20 / 20! < 2e-9 for the powers' tau = 3/|al| < 1/4, and below 1/22! for exp)
template <int KIND, bool PA = false, int NT = SMF_NTERMS>
MF_DEV double dd2(double li, double lj, double lk, double fi, double fj, double fk, double Dij, double Dik, double Djk,
                  double al, const double* __restrict__ c) {
  constexpr bool powlike = KIND == POW || KIND == SQRT;
  if (powlike) {
    li = li > 0 ? li : 0.0;
    lj = lj > 0 ? lj : 0.0;
    lk = lk > 0 ? lk : 0.0;
  }
  const double dA = lk - li, dB = lj - li, dC = lk - lj;
  const double aA = fabs(dA), aB = fabs(dB), aC = fabs(dC);
  const bool useA = (aA >= aB) && (aA >= aC), useB = !useA && (aB >= aC);
  const double num = useA ? Djk - Dij : (useB ? Djk - Dik : Dik - Dij);
  const double den = useA ? dA : (useB ? dB : dC);
  double out = num / (den != 0 ? den : 1.0);
  const double mn = li < lj ? li : lj, mx = li < lj ? lj : li;
  const double lo = mn < lk ? mn : lk, hi = mx > lk ? mx : lk;
  const double mxk = mx < lk ? mx : lk;
  const double mid = mn > mxk ? mn : mxk;
  const double fm = mid == li ? fi : (mid == lj ? fj : fk);
  bool cluster;
  double u, w, scl;
  if (KIND == EXP) {
    cluster = (hi - lo) <= 1.0;
    u = lo - mid;
    w = hi - mid;
    scl = fm;
  } else {
    // (|al|: per-matrix exponents may be <= 0 -- an extension; for al > 0 the reference's radius)
    const double tau = (KIND == POW && 3.0 / fabs(al) < 0.25) ? 3.0 / fabs(al) : 0.25;
    cluster = (mid > 0) && (mid - lo <= tau * mid) && (hi - mid <= tau * mid);
    const double ms = cluster ? mid : 1.0;
    u = (lo - mid) / ms;
    w = (hi - mid) / ms;
    scl = (KIND == LOG ? 1.0 : (cluster ? fm : 1.0)) / (ms * ms);
  }
  if (cluster) {                       // h_k(u, w) = u h_(k-1) + w^k
    double h = 1.0, wk = 1.0, ck = PA ? al * (al - 1.0) * 0.5 : c[0], s = ck;
#pragma unroll
    for (int k = 1; k < NT; k++) {
      wk = wk * w;
      h = u * h + wk;
      ck = PA ? ck * ((al - (k + 1.0)) * (1.0 / (k + 2))) : c[k];
      s = s + ck * h;
    }
    out = scl * s;
  }
  if (powlike && !(hi > 0)) out = al == 1.0 ? 0.0 : al * (al - 1.0) / 2.0 * pow_of_zero(al - 2.0);
  return out;
}

// packed position of the sorted triple (i <= j <= k) of eigenvalue indices among the 10 second divided differences
MF_DEV constexpr int d2idx(int i, int j, int k) {
  return i == 0 ? (j == 0 ? k : (j == 1 ? 2 + k : 5)) : (i == 1 ? (j == 1 ? 5 + k : 8) : 9);
}
// ... for any order of the three indices (constant after unrolling)
MF_DEV constexpr int d2pos(int i, int j, int k) {
  return d2idx(i < j ? (i < k ? i : k) : (j < k ? j : k), i + j + k - (i < j ? (i < k ? i : k) : (j < k ? j : k)) -
                   (i > j ? (i > k ? i : k) : (j > k ? j : k)), i > j ? (i > k ? i : k) : (j > k ? j : k));
}

// state of one matrix after the eigen-decomposition.  Eigenvalues, eigenvectors, f and the divided differences are
// FP64; the first-derivative assembly (P, Q) runs in AT: double for the FP64 form, float for the mixed form
template <int KIND, class AT = double>
struct SmfLane {
  double lam[3], fl[3], V[3][3];
  double D[3][3];                 // f[lam_i, lam_j] (diagonal: f'(lam_i))
  double D2[10];                  // f[lam_i, lam_j, lam_k], i <= j <= k, in the order of d2idx

  // ---- eigenvalues lam and vectors V (columns) of the Voigt vector v, and F (Voigt) -- smf_kernel.hpp:202-295
  MF_DEV void eigen(const double* v, double al, double* Fv) {
    const double SQRT6 = 2.449489742783178, TWO_PI_3 = 2.0943951023931957;
    const double ab0 = fabs(v[0]), ab1 = fabs(v[1]), ab2 = fabs(v[2]), ab3 = fabs(v[3]), ab4 = fabs(v[4]),
                 ab5 = fabs(v[5]);
    const double s01 = ab0 > ab1 ? ab0 : ab1, s23 = ab2 > ab3 ? ab2 : ab3, s45 = ab4 > ab5 ? ab4 : ab5;
    const double s03 = s01 > s23 ? s01 : s23;
    const double sig = pow2_floor((s03 > s45 ? s03 : s45) + 1e-300), isig = 1.0 / sig;
    const double a0 = v[0] * isig, a1 = v[1] * isig, a2 = v[2] * isig, a3 = v[3] * isig, a4 = v[4] * isig,
                 a5 = v[5] * isig;
    const double m = (a0 + a1 + a2) / 3.0, d0 = a0 - m, d1 = a1 - m, d2_ = a2 - m;
    const double tr3 = (d0 + d1 + d2_) / 3.0;
    const double e0 = d0 - tr3, e1 = d1 - tr3, e2 = d2_ - tr3;
    const double z01 = dmax(fabs(e0), fabs(e1)), z23 = fabs(e2) > fabs(a3) ? fabs(e2) : fabs(a3);
    const double z45 = fabs(a4) > fabs(a5) ? fabs(a4) : fabs(a5), z03 = z01 > z23 ? z01 : z23;
    const double dmx = z03 > z45 ? z03 : z45;
    const bool iso = dmx <= 1e-300;
    const double ids = 1.0 / (iso ? 1.0 : pow2_floor(dmx));
    const double g0 = e0 * ids, g1 = e1 * ids, g2 = e2 * ids, g3 = a3 * ids, g4 = a4 * ids, g5 = a5 * ids;
    const double p = g0 * g0 + g1 * g1 + g2 * g2 + 2.0 * (g3 * g3 + g4 * g4 + g5 * g5);
    const double q = g0 * (g1 * g2 - g5 * g5) - g3 * (g3 * g2 - g5 * g4) + g4 * (g3 * g5 - g1 * g4);
    const double pg = iso ? 1.0 : p, qg = iso ? 0.0 : q;
    double r = 3.0 * SQRT6 * qg / (pg * sqrt(pg));
    r = r > 1.0 ? 1.0 : (r < -1.0 ? -1.0 : r);
    const double t = sqrt(2.0 * pg / 3.0), ph = acos(r) / 3.0;
    const double sh = t * cos(r >= 0 ? ph : ph + TWO_PI_3);
    const double m00 = g0 - sh, m11 = g1 - sh, m22 = g2 - sh;
    const double c0x = g3 * g5 - g4 * m11, c0y = g4 * g3 - m00 * g5, c0z = m00 * m11 - g3 * g3;
    const double c1x = g3 * m22 - g4 * g5, c1y = g4 * g4 - m00 * m22, c1z = m00 * g5 - g3 * g4;
    const double c2x = m11 * m22 - g5 * g5, c2y = g5 * g4 - g3 * m22, c2z = g3 * g5 - m11 * g4;
    const double n0 = c0x * c0x + c0y * c0y + c0z * c0z, n1 = c1x * c1x + c1y * c1y + c1z * c1z,
                 n2 = c2x * c2x + c2y * c2y + c2z * c2z;
    const bool s1 = n1 > n0;
    double cx = s1 ? c1x : c0x, cy = s1 ? c1y : c0y, cz = s1 ? c1z : c0z, nn = s1 ? n1 : n0;
    const bool s2 = n2 > nn;
    cx = s2 ? c2x : cx; cy = s2 ? c2y : cy; cz = s2 ? c2z : cz; nn = s2 ? n2 : nn;
    const double nrm = sqrt(nn);
    const bool ok = (nrm > 1e-300) && !iso;
    const double inv = 1.0 / (ok ? nrm : 1.0);
    const double vx = ok ? cx * inv : 1.0, vy = ok ? cy * inv : 0.0, vz = ok ? cz * inv : 0.0;
    const double ax = fabs(vx), ay = fabs(vy), az = fabs(vz);
    const bool k1 = ay < ax, k2 = az < (k1 ? ay : ax);
    const double ex = (k1 || k2) ? 0.0 : 1.0, ey = (k1 && !k2) ? 1.0 : 0.0, ez = k2 ? 1.0 : 0.0;
    const double vk = ex * vx + ey * vy + ez * vz;
    double ux = ex - vk * vx, uy = ey - vk * vy, uz = ez - vk * vz;
    const double nu = sqrt(ux * ux + uy * uy + uz * uz);
    const bool oku = nu > 1e-300;
    const double iu = 1.0 / (oku ? nu : 1.0);
    ux = oku ? ux * iu : 0.0; uy = oku ? uy * iu : 0.0; uz = oku ? uz * iu : 0.0;
    const double wx = vy * uz - vz * uy, wy = vz * ux - vx * uz, wz = vx * uy - vy * ux;
    const double Aux = a0 * ux + a3 * uy + a4 * uz, Auy = a3 * ux + a1 * uy + a5 * uz, Auz = a4 * ux + a5 * uy + a2 * uz;
    const double Awx = a0 * wx + a3 * wy + a4 * wz, Awy = a3 * wx + a1 * wy + a5 * wz, Awz = a4 * wx + a5 * wy + a2 * wz;
    const double Avx = a0 * vx + a3 * vy + a4 * vz, Avy = a3 * vx + a1 * vy + a5 * vz, Avz = a4 * vx + a5 * vy + a2 * vz;
    const double a11 = ux * Aux + uy * Auy + uz * Auz, a12 = ux * Awx + uy * Awy + uz * Awz;
    const double a22 = wx * Awx + wy * Awy + wz * Awz, lv = vx * Avx + vy * Avy + vz * Avz;
    const double Dux = d0 * ux + a3 * uy + a4 * uz, Duy = a3 * ux + d1 * uy + a5 * uz, Duz = a4 * ux + a5 * uy + d2_ * uz;
    const double Dwx = d0 * wx + a3 * wy + a4 * wz, Dwy = a3 * wx + d1 * wy + a5 * wz, Dwz = a4 * wx + a5 * wy + d2_ * wz;
    const double b11 = ux * Dux + uy * Duy + uz * Duz, b12 = ux * Dwx + uy * Dwy + uz * Dwz,
                 b22 = wx * Dwx + wy * Dwy + wz * Dwz;
    const double jx = b11 - b22, jy = 2.0 * b12, jr = sqrt(jx * jx + jy * jy);
    const bool jz = jr == 0;
    const double jrs = jz ? 1.0 : jr;
    const double big = sqrt((jrs + fabs(jx)) / (2.0 * jrs));
    const double sml = jy / (2.0 * jrs * big);
    const double sgn = ((jy > 0) || ((jy == 0) && !sign_bit(jy))) ? 1.0 : -1.0;
    const double cs = jz ? 1.0 : (jx >= 0 ? big : sgn * sml);
    const double sn = jz ? 0.0 : (jx >= 0 ? sml : sgn * big);
    lam[0] = lv * sig;
    lam[1] = (cs * cs * a11 + 2.0 * cs * sn * a12 + sn * sn * a22) * sig;
    lam[2] = (sn * sn * a11 - 2.0 * cs * sn * a12 + cs * cs * a22) * sig;
    V[0][0] = vx; V[0][1] = cs * ux + sn * wx; V[0][2] = cs * wx - sn * ux;
    V[1][0] = vy; V[1][1] = cs * uy + sn * wy; V[1][2] = cs * wy - sn * uy;
    V[2][0] = vz; V[2][1] = cs * uz + sn * wz; V[2][2] = cs * wz - sn * uz;
#pragma unroll
    for (int i = 0; i < 3; i++) fl[i] = fval<KIND>(lam[i], al);
    if (KIND == LOG) {            // near m: log(m sig) + log1p(x / m), x from the deviator
      const double bv = vx * (d0 * vx + a3 * vy + a4 * vz) + vy * (a3 * vx + d1 * vy + a5 * vz) +
                        vz * (a4 * vx + a5 * vy + d2_ * vz);
      const double xd[3] = {bv, cs * cs * b11 + 2.0 * cs * sn * b12 + sn * sn * b22,
                            sn * sn * b11 - 2.0 * cs * sn * b12 + cs * cs * b22};
      const double lms = log(m * sig);
#pragma unroll
      for (int i = 0; i < 3; i++)
        if (fabs(xd[i]) <= 0.5 * m) fl[i] = lms + log1p(xd[i] / m);
    }
#pragma unroll
    for (int k = 0; k < 6; k++) {
      const int x = vrow(k), y = vcol(k);
      Fv[k] = V[x][0] * fl[0] * V[y][0] + V[x][1] * fl[1] * V[y][1] + V[x][2] * fl[2] * V[y][2];
    }
  }

  // ---- first derivative: the 21 distinct values Js[tri6(i,k)] = J_voigt[i][k] * w(k), w = 1 (k < 3) or 1/2,
  //      i.e. dF_ab / dA_cd of the matrix layout (symmetric in (ab) <-> (cd)) -- Daleckii-Krein, smf_kernel.hpp:296-327
  MF_DEV void jacobian(double al, AT* Js) {
    const double D01 = dd1<KIND>(lam[0], lam[1], fl[0], fl[1], al);
    const double D02 = dd1<KIND>(lam[0], lam[2], fl[0], fl[2], al);
    const double D12 = dd1<KIND>(lam[1], lam[2], fl[1], fl[2], al);
    D[0][0] = dfun<KIND>(lam[0], fl[0], al); D[1][1] = dfun<KIND>(lam[1], fl[1], al); D[2][2] = dfun<KIND>(lam[2], fl[2], al);
    D[0][1] = D[1][0] = D01; D[0][2] = D[2][0] = D02; D[1][2] = D[2][1] = D12;
    AT P[3][6], Q[3][6];            // P_j = v_j v_j^T, Q_j = sum_l D_jl P_l (packed symmetric)
#pragma unroll
    for (int j = 0; j < 3; j++)
#pragma unroll
      for (int k = 0; k < 6; k++) P[j][k] = (AT)V[vrow(k)][j] * (AT)V[vcol(k)][j];
#pragma unroll
    for (int j = 0; j < 3; j++)
#pragma unroll
      for (int k = 0; k < 6; k++) Q[j][k] = (AT)D[j][0] * P[0][k] + (AT)D[j][1] * P[1][k] + (AT)D[j][2] * P[2][k];
#pragma unroll
    for (int i = 0; i < 6; i++)
#pragma unroll
      for (int k = i; k < 6; k++) {
        const int x = vrow(i), y = vcol(i), p = vrow(k), q = vcol(k);
        AT s = P[0][tri3(x, p)] * Q[0][tri3(y, q)] + P[1][tri3(x, p)] * Q[1][tri3(y, q)] +
               P[2][tri3(x, p)] * Q[2][tri3(y, q)];
        if (p != q)
          s += P[0][tri3(x, q)] * Q[0][tri3(y, p)] + P[1][tri3(x, q)] * Q[1][tri3(y, p)] +
               P[2][tri3(x, q)] * Q[2][tri3(y, p)];
        Js[tri6(i, k)] = k < 3 ? s : (AT)0.5 * s;
      }
  }

  // ---- the 10 second divided differences (smf_kernel.hpp:330-338)
  template <bool PA = false, int NT = SMF_NTERMS>
  MF_DEV void second_dd(double al, const double* __restrict__ c2) {
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
      for (int j = i; j < 3; j++)
#pragma unroll
        for (int k = j; k < 3; k++)
          D2[d2idx(i, j, k)] =
              dd2<KIND, PA, NT>(lam[i], lam[j], lam[k], fl[i], fl[j], fl[k], D[i][j], D[i][k], D[j][k], al, c2);
  }
};

// ---- second derivative from V (eigenvectors as columns) and the second divided differences D2 of one matrix
//      (KIND no longer matters).  With Et^p = V^T E_p V for the Voigt basis E_p, S^pq = X + X^T,
//      X_kl = sum_m D2_kml Et^p_km Et^q_ml and H_i^pq = (V S^pq V^T)_ab (i = (a, b)) (smf_kernel.hpp:339-381):
//      H_i^pq = sum_m Et^p_m. M^m Et^q_m. with the symmetric M^m_kl = D2_kml G_kl, G_kl = V_ak V_bl + V_al V_bk,
//      the same sum associated per output row i, so that a row of 21 values is produced on its own
template <class AT>
struct SmfHess {
  AT V[3][3], D2[10];
  AT Et[6][6];                    // Et^p packed by tri3

  MF_DEV SmfHess(const AT* v9, const AT* d10) {
#pragma unroll
    for (int k = 0; k < 9; k++) V[k / 3][k % 3] = v9[k];
#pragma unroll
    for (int k = 0; k < 10; k++) D2[k] = d10[k];
#pragma unroll
    for (int p = 0; p < 6; p++)
#pragma unroll
      for (int t = 0; t < 6; t++) {
        const int k = vrow(p), l = vcol(p), a = vrow(t), b = vcol(t);
        AT e = V[k][a] * V[l][b];
        if (k != l) e += V[l][a] * V[k][b];
        Et[p][t] = e;
      }
  }

  // row i of the Voigt H: h[tri6(p, q)] = d2F_i / dv_p dv_q  (i may differ between lanes: the rows a, b of V are
  // selected by value, so that V stays in registers)
  MF_DEV void h_row(int i, AT* h) const {
    const int a = vrow(i), b = vcol(i);
    AT va[3], vb[3];
#pragma unroll
    for (int k = 0; k < 3; k++) {
      const AT v0 = V[0][k], v1 = V[1][k], v2 = V[2][k];
      va[k] = a == 0 ? v0 : (a == 1 ? v1 : v2);
      vb[k] = b == 0 ? v0 : (b == 1 ? v1 : v2);
    }
    AT G[6];
#pragma unroll
    for (int t = 0; t < 6; t++) {
      const int k = vrow(t), l = vcol(t);
      G[t] = va[k] * vb[l] + va[l] * vb[k];
    }
    AT M[3][6];
#pragma unroll
    for (int m = 0; m < 3; m++)
#pragma unroll
      for (int t = 0; t < 6; t++) M[m][t] = D2[d2pos(vrow(t), m, vcol(t))] * G[t];
#pragma unroll
    for (int p = 0; p < 6; p++) {
      AT y[3][3];                   // y^(p,m)_l = sum_k M^m_lk Et^p_mk
#pragma unroll
      for (int m = 0; m < 3; m++)
#pragma unroll
        for (int l = 0; l < 3; l++)
          y[m][l] = M[m][tri3(l, 0)] * Et[p][tri3(m, 0)] + M[m][tri3(l, 1)] * Et[p][tri3(m, 1)] +
                    M[m][tri3(l, 2)] * Et[p][tri3(m, 2)];
#pragma unroll
      for (int q = p; q < 6; q++) {
        AT s = 0;
#pragma unroll
        for (int m = 0; m < 3; m++)
#pragma unroll
          for (int l = 0; l < 3; l++) s += y[m][l] * Et[q][tri3(m, l)];
        h[tri6(p, q)] = s;
      }
    }
  }
};

// ---- output index maps: element e of one matrix's F / J / H in the API layout -> position among the distinct
//      values (F: Fv[6]; J: Js[21], J_voigt[i][k] w(k); H: 6 rows of 21, Hv[i][tri6(p,q)]) and its weight
//      (matrix layout: d v / d A_cd = 1 or 1/2 per derivative index, smf_kernel.hpp:384-406; Voigt: J_voigt = Js / w)
MF_DEV int f_src(bool mat, int e) { return mat ? voigt_of_m(e) : e; }
MF_DEV int j_src(bool mat, int e, double& w) {
  if (mat) {                          // J[9 m + c] = Js[(midx(m), midx(c))]
    w = 1.0;
    return tri6(voigt_of_m(e / 9), voigt_of_m(e % 9));
  }
  w = e % 6 < 3 ? 1.0 : 2.0;          // J[6 i + k] = Js[(i, k)] / w(k)
  return tri6(e / 6, e % 6);
}
MF_DEV int h_src(bool mat, int e, double& w) {
  if (mat) {                          // H[(9 m + c) 9 + d] = Hv[midx(m)][(midx(c), midx(d))] w(c) w(d)
    const int m = e / 81, c = (e / 9) % 9, d = e % 9;
    w = dv_dA(c) * dv_dA(d);
    return 21 * voigt_of_m(m) + tri6(voigt_of_m(c), voigt_of_m(d));
  }
  w = 1.0;                            // H[(6 i + p) 6 + q] = Hv[i][(p, q)]
  return 21 * (e / 36) + tri6((e / 6) % 6, e % 6);
}

// Voigt vector of (A + A^T) / 2 and the symmetry test |a_ij - a_ji| <= 1e-8 max |a_kl| (NaN / Inf pass); NaN if
// not symmetric (smf_kernel.hpp:408-421; the GPU API reports no -3, the results are NaN)
MF_DEV void matrix_to_voigt(const double* a, double* v) {
  double mx = 0.0;
#pragma unroll
  for (int k = 0; k < 9; k++) mx = fabs(a[k]) > mx ? fabs(a[k]) : mx;
  const double tol = 1e-8 * mx;
  const bool sym = !(fabs(a[1] - a[3]) > tol) && !(fabs(a[2] - a[6]) > tol) && !(fabs(a[5] - a[7]) > tol);
  const double nan = dnan();
  v[0] = sym ? a[0] : nan;
  v[1] = sym ? a[4] : nan;
  v[2] = sym ? a[8] : nan;
  v[3] = sym ? 0.5 * (a[1] + a[3]) : nan;
  v[4] = sym ? 0.5 * (a[2] + a[6]) : nan;
  v[5] = sym ? 0.5 * (a[5] + a[7]) : nan;
}

}  // namespace mf
