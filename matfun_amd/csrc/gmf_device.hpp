// gmf_device.hpp -- per-lane math of the general-matrix kernels: eigenvalues of a real 3x3 matrix in closed form with
// refinements, the confluent divided differences of the Newton-Hermite form (group Taylor series streamed over the
// series index, single-node residues), the real basis r_0..r_2 and the contracted difference tensors D1 / D2 / D3
// from which F, J and H are assembled.  Same method, constants and branch structure as the reference kernel
// c/include/matfun/gmf_kernel.hpp (:166-692); the assembly of J and H is left to the kernel (lane groups).
#pragma once
#include "mf_device.hpp"

namespace mf {

constexpr int GMF_KMAX = 200;     // series terms at most
struct GmfCoef {                  // b_j of f(c (1 + u)): exp 1/j!, log (-1)^(j+1)/j, pow binom(al, j); host-built
  double b[GMF_KMAX];
};
// (gmf_kernel.hpp:48-53)
inline void gmf_coefs(int kind, double al, GmfCoef& c) {
  c.b[0] = 1.0;
  for (int j = 1; j < GMF_KMAX; j++)
    c.b[j] = kind == EXP ? c.b[j - 1] / j : (kind == LOG ? ((j & 1) ? 1.0 : -1.0) / j : c.b[j - 1] * (al - j + 1) / j);
}
// powers with one exponent per matrix: the table holds 1 / j, from which each lane forms its own
// b_j = b_(j-1) (al - j + 1) / j
inline void gmf_recips(GmfCoef& c) {
  c.b[0] = 1.0;
  for (int j = 1; j < GMF_KMAX; j++) c.b[j] = 1.0 / j;
}

// the list of divided differences f[x1^a x2^b x3^c], a >= b >= c, ordered by a, then b, then c (19 entries for
// a <= 3; the first 3 / 9 / 19 serve orders 0 / 1 / 2).  Each entry adds one node to its parent: the first node
// whose multiplicity exceeds the next one's.
MF_DEV constexpr int dd_pos(int a, int b, int c) { return (a == 1 ? 0 : (a == 2 ? 3 : 9)) + b * (b + 1) / 2 + c; }
MF_DEV constexpr int dd_a(int t) { return t < 3 ? 1 : (t < 9 ? 2 : 3); }
MF_DEV constexpr int dd_o(int t) { return t < 3 ? t : (t < 9 ? t - 3 : t - 9); }
MF_DEV constexpr int dd_b(int t) { return dd_o(t) < 1 ? 0 : (dd_o(t) < 3 ? 1 : (dd_o(t) < 6 ? 2 : 3)); }
MF_DEV constexpr int dd_c(int t) { return dd_o(t) - dd_b(t) * (dd_b(t) + 1) / 2; }
MF_DEV constexpr int dd_mult(int t, int node) { return node == 0 ? dd_a(t) : (node == 1 ? dd_b(t) : dd_c(t)); }
MF_DEV constexpr int dd_added(int t) { return dd_a(t) > dd_b(t) ? 0 : (dd_b(t) > dd_c(t) ? 1 : 2); }
MF_DEV constexpr int dd_parent(int t) {
  return dd_added(t) == 0 ? (dd_a(t) == 1 ? -1 : dd_pos(dd_a(t) - 1, dd_b(t), dd_c(t)))
                          : (dd_added(t) == 1 ? dd_pos(dd_a(t), dd_b(t) - 1, dd_c(t)) : dd_pos(dd_a(t), dd_b(t), dd_c(t) - 1));
}
// positions of f[N_j], f[N_j + N_k], f[N_j + N_k + N_l] (N_j = {x_1 .. x_(j+1)})
MF_DEV constexpr int pos1(int j) { return dd_pos(1, j >= 1, j >= 2); }
MF_DEV constexpr int pos2(int j, int k) { return dd_pos(2, (j >= 1) + (k >= 1), (j >= 2) + (k >= 2)); }
MF_DEV constexpr int pos3(int j, int k, int l) {
  return dd_pos(3, (j >= 1) + (k >= 1) + (l >= 1), (j >= 2) + (k >= 2) + (l >= 2));
}

MF_DEV double det3(const double* M) {
  return M[0] * (M[4] * M[8] - M[5] * M[7]) - M[1] * (M[3] * M[8] - M[5] * M[6]) + M[2] * (M[3] * M[7] - M[4] * M[6]);
}
// det by elimination with partial pivoting (exact pivots for a diagonal matrix); the rows are taken as values so
// that the pivot choice selects values, not addresses (a selected address would put the matrix in scratch memory)
MF_DEV double det_pivot(const double* M) {
  const double m0 = M[0], m1 = M[1], m2 = M[2], m3 = M[3], m4 = M[4], m5 = M[5], m6 = M[6], m7 = M[7], m8 = M[8];
  const double a0 = fabs(m0), a1 = fabs(m3), a2 = fabs(m6);
  const bool p1 = (a1 > a0) && (a1 >= a2), p2 = (a2 > a0) && (a2 > a1);
  const double P0 = p1 ? m3 : (p2 ? m6 : m0), P1 = p1 ? m4 : (p2 ? m7 : m1), P2 = p1 ? m5 : (p2 ? m8 : m2);
  const double R10 = (p1 || p2) ? m0 : m3, R11 = (p1 || p2) ? m1 : m4, R12 = (p1 || p2) ? m2 : m5;
  const double R20 = p2 ? m3 : m6, R21 = p2 ? m4 : m7, R22 = p2 ? m5 : m8;
  double sgn = p1 ? -1.0 : 1.0;
  const double piv = P0 != 0 ? P0 : 1.0;
  const double l1 = R10 / piv, l2 = R20 / piv;
  const double b = R11 - l1 * P1, c = R12 - l1 * P2, d = R21 - l2 * P1, e = R22 - l2 * P2;
  const bool sw = fabs(d) > fabs(b);
  const double bb = sw ? d : b, cc = sw ? e : c, dd = sw ? b : d, ee = sw ? c : e;
  sgn = sw ? -sgn : sgn;
  const double det = sgn * P0 * bb * (ee - dd / (bb != 0 ? bb : 1.0) * cc);
  return (P0 != 0 && bb != 0) ? det : 0.0;
}
MF_DEV double minors2(const double* M) {
  return M[0] * M[4] - M[1] * M[3] + M[0] * M[8] - M[2] * M[6] + M[4] * M[8] - M[5] * M[7];
}

// eigenvalues of S (gmf_kernel.hpp:166-208): its trace residue tau plus the roots of the traceless, twice normalised
// x^3 - P x - Q; x[0] by the trigonometric formula (three real roots) or Cardano, x[1], x[2] from the deflated quadratic
MF_DEV void cubic_roots(const double* S, cpx* x) {
  const double TWO_PI_3 = 2.0943951023931957;
  const double tau = (S[0] + S[4] + S[8]) / 3.0;
  double Sp[9];
#pragma unroll
  for (int k = 0; k < 9; k++) Sp[k] = k % 4 == 0 ? S[k] - tau : S[k];
  double s = fabs(Sp[0]);
#pragma unroll
  for (int k = 1; k < 9; k++) s = dmax(s, fabs(Sp[k]));
  double t = pow2_near(s);
  const double it = 1.0 / t;
#pragma unroll
  for (int k = 0; k < 9; k++) Sp[k] = Sp[k] * it;
  double P = 0.5 * (Sp[0] * Sp[0] + Sp[4] * Sp[4] + Sp[8] * Sp[8]) + Sp[1] * Sp[3] + Sp[2] * Sp[6] + Sp[5] * Sp[7];
  double Q = det3(Sp);
  const double u = pow2_near(dmax(sqrt(fabs(P)), cbrt(fabs(Q)))), iu = 1.0 / u;
  P = P * iu * iu;
  Q = Q * iu * iu * iu;
  t = t * u;
  const bool trig = (P > 0) && (4.0 * P * P * P >= 27.0 * Q * Q);
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
  const double d = P - 0.75 * s1 * s1;
  const bool real = d >= 0;
  const double sq = sqrt(fabs(d));
  const double sa = -0.5 * s1 - copysign_of(sq, s1);
  const double prod = s1 != 0 ? Q / s1 : -P;
  const double sb = sa != 0 ? prod / sa : 0.0;
  const double im = (real ? 0.0 : sq) * t;
  x[0] = mk(s1 * t + tau, 0.0 * t);
  x[1] = mk((real ? sa : -0.5 * s1) * t + tau, im);
  x[2] = mk((real ? sb : -0.5 * s1) * t + tau, -im);
}

// real node j: 3 Newton steps on det(As - mu I), the derivative from the other nodes
MF_DEV void newton_node(const double* As, cpx* mu, int j) {
#pragma unroll
  for (int it = 0; it < 3; it++) {
    cpx den = mk(1.0, 0.0);
#pragma unroll
    for (int k = 0; k < 3; k++)
      if (k != j) den = mul(den, sub(mu[k], mu[j]));
    if ((den.re != 0) || (den.im != 0)) {
      double B[9];
#pragma unroll
      for (int k = 0; k < 9; k++) B[k] = k % 4 == 0 ? As[k] - mu[j].re : As[k];
      mu[j].re = mu[j].re + det_pivot(B) * cinv(den).re;
    }
  }
}

// Taylor coefficient 0 of f at a node: exp e^x (x relative to the exp shift); log log(sig mu) (log1p of x/m near m);
// pow (sig |mu|)^(al-k) e^(i (al-k) arg mu) mu^k  (gmf_kernel.hpp:228-248)
template <int KIND>
MF_DEV cpx node_value(double al, double sig, double m, cpx x, cpx mu, int k) {
  if (KIND == EXP) {
    const double e = exp(x.re);
    return mk(e * cos(x.im), e * sin(x.im));
  }
  if (KIND == LOG) {
    if ((m > 0) && (cabs_(x) <= 0.5 * m)) {
      const double zr = x.re / m, zi = x.im / m;
      return mk(log(sig * m) + 0.5 * log1p(zr * (2.0 + zr) + zi * zi), atan2(zi, 1.0 + zr));
    }
    return mk(log(sig * cabs_(mu)), atan2(mu.im, mu.re));
  }
  const double th = atan2(mu.im, mu.re);
  const double mag = powk(sig * cabs_(mu), al, (double)k);
  const double ph = (al - k) * th;
  cpx out = mk(mag * cos(ph), mag * sin(ph));
  for (int i = 0; i < k; i++) out = mul(out, mu);
  return out;
}

// 1 / j for j = 1 .. GMF_KMAX (entry 0: 1), computed at compile time (correctly rounded quotients)
struct GmfRecip {
  double r[GMF_KMAX + 1];
};
constexpr GmfRecip gmf_recip_table() {
  GmfRecip t{};
  t.r[0] = 1.0;
  for (int j = 1; j <= GMF_KMAX; j++) t.r[j] = 1.0 / j;
  return t;
}

// number of group-series terms (gmf_kernel.hpp:73-97); the bound's binomial C(k + n, n) advances by a product with
// 1 / (k + 1) from rinv (gmf_recip_table) instead of a division (the count can differ from the reference's only for a
// term within rounding of the 1e-17 cut, which changes the sum by less than that)
// PA: b holds 1 / j and the coefficients binom(al, j) are formed on the way (gmf_recips)
template <bool PA = false>
MF_DEV int series_terms(const double* __restrict__ b, double al, int n, double r, double q,
                        const double* __restrict__ rinv) {
  int M = 1;
  if (r > 0) {
    double best = 0.0, c = 1.0, rk = 1.0, bn = 1.0;
    if (PA)
      for (int j = 1; j <= n; j++) bn = bn * ((al - (j - 1)) * b[j]);
    for (int k = 0; n + k < GMF_KMAX; k++) {
      const double t = fabs(PA ? bn : b[n + k]) * c * rk;
      best = t > best ? t : best;
      M = k;
      if ((k > 2 && t <= 1e-17 * best) || !(rk > 0)) break;
      c = c * (k + n + 1) * rinv[k + 1];
      rk *= r;
      if (PA) bn = bn * ((al - (n + k)) * b[n + k + 1 < GMF_KMAX ? n + k + 1 : GMF_KMAX - 1]);
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

// the group series of one matrix (gmf_kernel.hpp:478-533): K terms (0: no group) of sum_j b_j E_j with the
// offsets u of the nodes from the group centre, the pole factor 1 / ap (iap) of a node outside a pair group, the
// first coefficient b0 (log: its value at the centre), the exponent al (per-matrix exponents), and the group and
// pole flags (bits 0..2: node in the group; 3..5: the pole node)
struct GmfSeriesIn {
  cpx u[3];
  double iap, b0, al;
  int K, flags;
};

// acc[c] = the series sums of the ND = 3 / 9 / 19 entries of the divided-difference list (order 0 / 1 / 2), streamed
// over the series index j; bt: the coefficients b_j (PA: 1 / j, the coefficients formed per lane).
// The dual vectors E follow one recurrence for every entry, E_j[c] = E_(j-1)[parent] + u_i E_(j-1)[c] (i the node
// the entry adds; a descending pass reads the parents' old values), with the offset u of a node outside a pair group
// zero.  An entry that adds the outside node then repeats its parent one term late, so every entry c runs
// m(c) terms behind the group series it stands for (m = the multiplicity in c of the outside node, which is also
// the pole node): term j of entry c is taken with the pole factor G^(m) of term j - m (kept for the last three
// terms), and the loop runs ORDER + 1 terms past K with the factors of terms >= K zero.  Same values, sums and order
// of operations as copying the parent's value into such entries at every term (up to the sign of exact zeros),
// without the per-entry selects of the copy.
template <int KIND, int ORDER, bool PA>
MF_DEV void gmf_series(const GmfSeriesIn& s, const double* __restrict__ bt, cpx* acc) {
  constexpr int ND = ORDER == 0 ? 3 : (ORDER == 1 ? 9 : 19), LAG = ORDER + 1;
  const bool pk[3] = {(s.flags & 8) != 0, (s.flags & 16) != 0, (s.flags & 32) != 0};
  const bool nopole = !(pk[0] || pk[1] || pk[2]);
  const double iap = s.iap, al = s.al;
  cpx E[ND];
#pragma unroll
  for (int c = 0; c < ND; c++) {
    E[c] = mk(0.0, 0.0);
    acc[c] = mk(0.0, 0.0);
  }
  double q1 = 0.0, q2 = 0.0, q3 = 0.0;        // G^(1..3) of the previous term
  double r2 = 0.0, r3a = 0.0, r3b = 0.0;        // G^(2) of term j - 2, G^(3) of terms j - 2 and j - 3
  double bnext = 0.0;                           // (b_j loaded one iteration ahead of its use)
  const int KT = s.K > 0 ? s.K + LAG : 0;
  for (int j = 0; j < KT; j++) {
    const double bl = j == 0 ? s.b0 : bnext;
    bnext = PA ? bl * ((al - j) * bt[j + 1 < GMF_KMAX ? j + 1 : GMF_KMAX - 1])
               : bt[j + 1 < GMF_KMAX ? j + 1 : GMF_KMAX - 1];
    // factors by the multiplicity m of the pole node: b_j, G^(1) of term j - 1, G^(2) of j - 2, G^(3) of j - 3
    // (zero for terms past K - 1; without a pole every entry takes b_j: two selects per entry below)
    const double bj = j < s.K ? bl : 0.0;
    const double f1 = nopole ? bj : (j - 1 < s.K ? q1 : 0.0);
    const double f2 = nopole ? bj : (j - 2 < s.K ? r2 : 0.0);
    const double f3 = nopole ? bj : (j - 3 < s.K ? r3b : 0.0);
    r3b = r3a; r3a = q3; r2 = q2;
    const double g1 = (bl - q1) * iap, g2 = (g1 - q2) * iap, g3 = (g2 - q3) * iap;
    q1 = g1; q2 = g2; q3 = g3;
#pragma unroll
    for (int c = ND - 1; c >= 0; c--) {
      const int par = dd_parent(c), i = dd_added(c);
      const cpx epp = par < 0 ? mk(j == 0 ? 1.0 : 0.0, 0.0) : E[par];
      E[c] = add(epp, mul(s.u[i], E[c]));
    }
#pragma unroll
    for (int c = 0; c < ND; c++) {
      // (masks loop-invariant, indices constant)
      const double qa = dd_a(c) == 0 ? bj : (dd_a(c) == 1 ? f1 : (dd_a(c) == 2 ? f2 : f3));
      const double qb = dd_b(c) == 0 ? bj : (dd_b(c) == 1 ? f1 : (dd_b(c) == 2 ? f2 : f3));
      const double qc = dd_c(c) == 0 ? bj : (dd_c(c) == 1 ? f1 : (dd_c(c) == 2 ? f2 : f3));
      const double g = pk[0] ? qa : (pk[1] ? qb : qc);
      acc[c] = add(acc[c], scale(E[c], g));
    }
  }
}

// the class of a lane's group series: 0 real offsets and no pole node (a group of all three nodes, or none), 1 real
// offsets with a pole node (a pair group of real nodes), 2 complex offsets (a conjugate pair, or a triple with one)
MF_DEV int gmf_series_class(const GmfSeriesIn& s) {
  const bool real = s.u[0].im == 0.0 && s.u[1].im == 0.0 && s.u[2].im == 0.0;
  return !real ? 2 : ((s.flags & 56) != 0 ? 1 : 0);
}

// gmf_series for a lane whose group offsets u are real: the dual vectors E and the sums stay real (their imaginary
// parts are zero in gmf_series), so the recurrence runs in real arithmetic -- two FP64 operations per entry and term
// instead of six.  POLE = false: for a lane without a pole node (gmf_series_class 0), whose entries all take b_j
// (none of the factor selects; the lagged terms past K are not needed either); POLE = true: class 1.  The sums equal
// gmf_series' up to rounding (the loops contract into FMAs differently); a lane whose offsets are not real (class 2)
// must not call it.
template <int KIND, int ORDER, bool PA, bool POLE>
MF_DEV void gmf_series_real(const GmfSeriesIn& s, const double* __restrict__ bt, cpx* acc) {
  constexpr int ND = ORDER == 0 ? 3 : (ORDER == 1 ? 9 : 19), LAG = ORDER + 1;
  const bool pk[3] = {(s.flags & 8) != 0, (s.flags & 16) != 0, (s.flags & 32) != 0};
  const bool nopole = !(pk[0] || pk[1] || pk[2]);
  const double iap = s.iap, al = s.al;
  const double u[3] = {s.u[0].re, s.u[1].re, s.u[2].re};
  double E[ND], a[ND];
#pragma unroll
  for (int c = 0; c < ND; c++) E[c] = a[c] = 0.0;
  double q1 = 0.0, q2 = 0.0, q3 = 0.0, r2 = 0.0, r3a = 0.0, r3b = 0.0, bnext = 0.0;
  const int KT = s.K > 0 ? s.K + (POLE ? LAG : 0) : 0;
  for (int j = 0; j < KT; j++) {
    const double bl = j == 0 ? s.b0 : bnext;
    bnext = PA ? bl * ((al - j) * bt[j + 1 < GMF_KMAX ? j + 1 : GMF_KMAX - 1])
               : bt[j + 1 < GMF_KMAX ? j + 1 : GMF_KMAX - 1];
    const double bj = POLE ? (j < s.K ? bl : 0.0) : bl;
    double f1 = bj, f2 = bj, f3 = bj;
    if constexpr (POLE) {
      f1 = nopole ? bj : (j - 1 < s.K ? q1 : 0.0);
      f2 = nopole ? bj : (j - 2 < s.K ? r2 : 0.0);
      f3 = nopole ? bj : (j - 3 < s.K ? r3b : 0.0);
      r3b = r3a; r3a = q3; r2 = q2;
      const double g1 = (bl - q1) * iap, g2 = (g1 - q2) * iap, g3 = (g2 - q3) * iap;
      q1 = g1; q2 = g2; q3 = g3;
    }
#pragma unroll
    for (int c = ND - 1; c >= 0; c--) {
      const int par = dd_parent(c), i = dd_added(c);
      const double epp = par < 0 ? (j == 0 ? 1.0 : 0.0) : E[par];
      E[c] = epp + u[i] * E[c];
    }
#pragma unroll
    for (int c = 0; c < ND; c++) {
      double g = bj;
      if constexpr (POLE) {
        const double qa = dd_a(c) == 0 ? bj : (dd_a(c) == 1 ? f1 : (dd_a(c) == 2 ? f2 : f3));
        const double qb = dd_b(c) == 0 ? bj : (dd_b(c) == 1 ? f1 : (dd_b(c) == 2 ? f2 : f3));
        const double qc = dd_c(c) == 0 ? bj : (dd_c(c) == 1 ? f1 : (dd_c(c) == 2 ? f2 : f3));
        g = pk[0] ? qa : (pk[1] ? qb : qc);
      }
      a[c] += E[c] * g;
    }
  }
#pragma unroll
  for (int c = 0; c < ND; c++) acc[c] = mk(a[c], 0.0);
}

// the series computed by the lane itself
template <int KIND, int ORDER, bool PA>
struct GmfSeriesLocal {
  static constexpr bool K_ONLY = false;
  const double* bt;
  MF_DEV void operator()(const GmfSeriesIn& s, cpx* acc) const { gmf_series<KIND, ORDER, PA>(s, bt, acc); }
};

// no series: gmf_lane stops once the series inputs are known, returns the length in GmfOut::K and hands the inputs
// and the nodes to post() (the first pass of the length-sorted evaluation, gmf_sorted.hip; NULL: not stored): the
// nodes staged as gmf_stage_nodes writes them
struct GmfNodesRe;
template <class N>
MF_DEV void gmf_stage_nodes(const N& st, int cls, GmfNodesRe& r, double* im);
struct GmfSeriesLength {
  static constexpr bool K_ONLY = true;
  GmfSeriesIn* out = nullptr;
  GmfNodesRe* nre = nullptr;
  double* nim = nullptr;
  template <class N>
  MF_DEV void post(const GmfSeriesIn& s, const N& st) const {
    if (out) *out = s;
    if (nre) gmf_stage_nodes(st, gmf_series_class(s), *nre, nim);
  }
  MF_DEV void operator()(const GmfSeriesIn&, cpx*) const {}
};

// results of one matrix needed for the output assembly
struct GmfOut {
  double F[9];
  double r[3][9];     // real basis r_0 = I, r_1, r_2
  double D2[9];       // Re(T^T D T) on two indices (order >= 1)
  double D3[27];      // ... on three indices (order >= 2)
  double sc1, sc2;    // factors of J and H (NaN for a matrix outside the domain)
  int K;              // series length (set only by a GmfSeriesLength pass)
  int cls;            // gmf_series_class of the series (set only by a GmfSeriesLength pass)
};

// what the rest of a matrix's evaluation needs from its nodes, once the group series is set up (gmf_nodes ->
// gmf_finish): the node offsets x from m and the nodes mu, the scaled mean m, the scale sig, the exp shift, the group
// centre cxg and the series scale L; flags: bits 0..2 the refined nodes (ab), 3..5 the group (inG), 6 outside the
// domain (bad), 7 a group exists (hasg)
struct GmfNodes {
  cpx x[3], mu[3];
  double m, sig, xshift, cxg, L;
  unsigned flags, pad;
};

// the nodes as the staged evaluation keeps them between its passes, in two parts so that the passes that read them
// move fewer bytes (reads amid the finishing pass's stores cost it several times their share of the traffic): the
// real record (96 B) of every matrix, and the imaginary parts of x and mu (6 doubles) only for a matrix that has one
// other than +0 (flags bit 8); flags bits 0..7 as GmfNodes, bits 9..10 the series class (gmf_series_class: sums with
// imaginary parts only for class 2)
struct GmfNodesRe {
  double xr[3], mur[3];
  double m, sig, xshift, cxg, L;
  unsigned flags, pad;
};
constexpr unsigned GMF_NODES_CPLX = 256u;
template <class N>
MF_DEV void gmf_stage_nodes(const N& st, int cls, GmfNodesRe& r, double* im) {
  bool cn = false;
#pragma unroll
  for (int i = 0; i < 3; i++) {
    cn = cn || st.x[i].im != 0.0 || __builtin_signbit(st.x[i].im) || st.mu[i].im != 0.0 ||
         __builtin_signbit(st.mu[i].im);
    r.xr[i] = st.x[i].re;
    r.mur[i] = st.mu[i].re;
  }
  r.m = st.m;
  r.sig = st.sig;
  r.xshift = st.xshift;
  r.cxg = st.cxg;
  r.L = st.L;
  r.flags = st.flags | (cn ? GMF_NODES_CPLX : 0u) | ((unsigned)cls << 9);
  r.pad = 0;
  if (cn) {
#pragma unroll
    for (int i = 0; i < 3; i++) {
      im[i] = st.x[i].im;
      im[3 + i] = st.mu[i].im;
    }
  }
}
// the nodes back from the two parts (imaginary parts +0 unless stored)
MF_DEV GmfNodes gmf_unstage_nodes(const GmfNodesRe& r, const double* im) {
  GmfNodes st;
  double xi[3] = {0.0, 0.0, 0.0}, mi[3] = {0.0, 0.0, 0.0};
  if (r.flags & GMF_NODES_CPLX) {
#pragma unroll
    for (int i = 0; i < 3; i++) {
      xi[i] = im[i];
      mi[i] = im[3 + i];
    }
  }
#pragma unroll
  for (int i = 0; i < 3; i++) {
    st.x[i] = mk(r.xr[i], xi[i]);
    st.mu[i] = mk(r.mur[i], mi[i]);
  }
  st.m = r.m;
  st.sig = r.sig;
  st.xshift = r.xshift;
  st.cxg = r.cxg;
  st.L = r.L;
  st.flags = r.flags & 255u;
  st.pad = 0;
  return st;
}

// PA (powers with a per-matrix exponent al): bt holds 1 / j instead of the coefficients b_j (gmf_recips)
// rinv: gmf_recip_table().  The nodes and the group series' inputs of one matrix (gmf_kernel.hpp:251-476)
template <int KIND, int ORDER, bool PA>
MF_DEV void gmf_nodes(const double* a, double al, const double* __restrict__ bt, const double* __restrict__ rinv,
                      GmfNodes& st, GmfSeriesIn& sin) {
  const double RHO = 0.7, R_EXP = 4.0, RHO_POLE = 0.6;
  double amax = fabs(a[0]);
#pragma unroll
  for (int k = 1; k < 9; k++) amax = dmax(amax, fabs(a[k]));
  const double sig = KIND == EXP ? 1.0 : pow2_near(amax), isig = 1.0 / sig;
  double As[9], S[9];
#pragma unroll
  for (int k = 0; k < 9; k++) As[k] = a[k] * isig;
  const double m = (As[0] + As[4] + As[8]) / 3.0;
#pragma unroll
  for (int k = 0; k < 9; k++) S[k] = k % 4 == 0 ? As[k] - m : As[k];
  cpx x[3], mu[3];
  cubic_roots(S, x);
#pragma unroll
  for (int i = 0; i < 3; i++) mu[i] = mk(m + x[i].re, x[i].im);

  // ---- pair from sum and product with A itself after Newton on mu[0] (gmf_kernel.hpp:277-304)
  double smax = fabs(S[0]);
#pragma unroll
  for (int k = 1; k < 9; k++) smax = dmax(smax, fabs(S[k]));
  const double mid = 0.5 * (mu[1].re + mu[2].re), Rp = 0.5 * cabs_(sub(mu[1], mu[2]));
  const bool small = fabs(mid) + Rp <= 0.25 * fabs(mu[0].re);
  const bool aniso = (fabs(m) <= smax) && (Rp <= 0.25 * fabs(mid - mu[0].re));
  bool pc = (mu[0].re != 0) && (small || aniso);
  double pcen = 0.0, ps = 0.0;
  bool preal = true;
  if (pc) {
    newton_node(As, mu, 0);
    const double m0 = mu[0].re;
    const bool big = fabs(m0) >= fabs(mid) + Rp;
    double sg, pi;
    if (big) {
      pi = det_pivot(As) / m0;
      sg = (minors2(As) - pi) / m0;
    } else {
      sg = As[0] + As[4] + As[8] - m0;
      pi = minors2(As) - m0 * sg;
    }
    const double disc = 0.25 * sg * sg - pi;
    pcen = 0.5 * sg;
    ps = sqrt(fabs(disc));
    preal = disc >= 0;
    const double r1 = pcen + copysign_of(ps, pcen);
    const double r2 = r1 != 0 ? pi / r1 : pcen - copysign_of(ps, pcen);
    mu[1] = preal ? mk(r1, 0.0) : mk(pcen, ps);
    mu[2] = preal ? mk(r2, 0.0) : mk(pcen, -ps);
  }
  // ---- conjugate pair near the negative real axis (log, powers): beta^2 from the deviator's 2x2 block on the
  //      complement of the left eigenvector of mu[0] (gmf_kernel.hpp:305-367)
  const bool cut = KIND != EXP && (mu[1].re + mu[2].re < 0);
  if (cut) {
    if (!pc) newton_node(As, mu, 0);
    const double l0 = mu[0].re;
    double Mc[9];
#pragma unroll
    for (int k = 0; k < 9; k++) Mc[k] = k % 4 == 0 ? As[k] - l0 : As[k];
    const double c01[3] = {Mc[3] * Mc[7] - Mc[6] * Mc[4], Mc[6] * Mc[1] - Mc[0] * Mc[7], Mc[0] * Mc[4] - Mc[3] * Mc[1]};
    const double c02[3] = {Mc[3] * Mc[8] - Mc[6] * Mc[5], Mc[6] * Mc[2] - Mc[0] * Mc[8], Mc[0] * Mc[5] - Mc[3] * Mc[2]};
    const double c12[3] = {Mc[4] * Mc[8] - Mc[7] * Mc[5], Mc[7] * Mc[2] - Mc[1] * Mc[8], Mc[1] * Mc[5] - Mc[4] * Mc[2]};
    const double n01 = c01[0] * c01[0] + c01[1] * c01[1] + c01[2] * c01[2];
    const double n02 = c02[0] * c02[0] + c02[1] * c02[1] + c02[2] * c02[2];
    const double n12 = c12[0] * c12[0] + c12[1] * c12[1] + c12[2] * c12[2];
    const bool t1 = n02 > n01;
    const double c010 = c01[0], c011 = c01[1], c012 = c01[2], c020 = c02[0], c021 = c02[1], c022 = c02[2];
    const double c120 = c12[0], c121 = c12[1], c122 = c12[2];
    double u0 = t1 ? c020 : c010, u1 = t1 ? c021 : c011, u2 = t1 ? c022 : c012, nn = t1 ? n02 : n01;
    const bool t2 = n12 > nn;
    u0 = t2 ? c120 : u0; u1 = t2 ? c121 : u1; u2 = t2 ? c122 : u2; nn = t2 ? n12 : nn;
    if (nn > 0) {
      const double iu = 1.0 / sqrt(nn);
      u0 = u0 * iu; u1 = u1 * iu; u2 = u2 * iu;
      const double au0 = fabs(u0), au1 = fabs(u1), au2 = fabs(u2);
      const bool k1 = au1 < au0, k2 = au2 < (k1 ? au1 : au0);
      const double ex = (k1 || k2) ? 0.0 : 1.0, ey = (k1 && !k2) ? 1.0 : 0.0, ez = k2 ? 1.0 : 0.0;
      const double eu = ex * u0 + ey * u1 + ez * u2;
      double q1[3] = {ex - eu * u0, ey - eu * u1, ez - eu * u2};
      const double iq = 1.0 / sqrt(q1[0] * q1[0] + q1[1] * q1[1] + q1[2] * q1[2]);
#pragma unroll
      for (int k = 0; k < 3; k++) q1[k] = q1[k] * iq;
      const double q2[3] = {u1 * q1[2] - u2 * q1[1], u2 * q1[0] - u0 * q1[2], u0 * q1[1] - u1 * q1[0]};
      double Sq1[3], Sq2[3];
#pragma unroll
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
      const bool creal = beta2 <= 0;
      mu[1] = creal ? mk(cc + bta, 0.0) : mk(cc, bta);
      mu[2] = creal ? mk(cc - bta, 0.0) : mk(cc, -bta);
      pcen = cc;
      ps = bta;
      preal = creal;
      pc = true;
    }
  }
  bool ab[3] = {pc, pc, pc};
  int og[3] = {0, 1, 2};
#pragma unroll
  for (int i = 0; i < 3; i++) x[i] = ab[i] ? mk(mu[i].re - m, mu[i].im) : x[i];

  // ---- node order: largest high derivatives last; a conjugate pair adjacent (gmf_kernel.hpp:373-410)
  auto key = [&](cpx v) { return KIND == EXP ? v.re : -cabs_(v); };
  {
    cpx sx[3] = {x[0], x[1], x[2]}, sm[3] = {mu[0], mu[1], mu[2]};
    bool sa[3] = {ab[0], ab[1], ab[2]};
    int so[3] = {og[0], og[1], og[2]};
    constexpr int NET[3][2] = {{0, 1}, {1, 2}, {0, 1}};
#pragma unroll
    for (int s = 0; s < 3; s++) {
      const int i = NET[s][0], j = NET[s][1];
      const bool sw = key(sm[i]) > key(sm[j]);
      const cpx tx = sx[i], tm = sm[i];
      const bool ta = sa[i];
      const int to = so[i];
      sx[i] = pick(sw, sx[j], sx[i]); sx[j] = pick(sw, tx, sx[j]);
      sm[i] = pick(sw, sm[j], sm[i]); sm[j] = pick(sw, tm, sm[j]);
      const bool saj = sa[j], sai = sa[i];
      const int soj = so[j], soi = so[i];
      sa[i] = sw ? saj : sai; sa[j] = sw ? ta : saj;
      so[i] = sw ? soj : soi; so[j] = sw ? to : soj;
    }
    const bool first = (key(mu[0]) <= key(mu[1])) || cut;
    const bool cplx_pair = x[1].im != 0;
    cpx nx[3], nm[3];
    bool na[3];
    int no[3];
#pragma unroll
    for (int k = 0; k < 3; k++) {
      const int r = (k + 1) % 3;
      // (values, not lvalues, in the selects: a selected address would move the arrays to scratch memory)
      nx[k] = pick(cplx_pair, pick(first, x[k], x[r]), sx[k]);
      nm[k] = pick(cplx_pair, pick(first, mu[k], mu[r]), sm[k]);
      const bool abk = ab[k], abr = ab[r], sak = sa[k];
      const int ogk = og[k], ogr = og[r], sok = so[k];
      na[k] = cplx_pair ? (first ? abk : abr) : sak;
      no[k] = cplx_pair ? (first ? ogk : ogr) : sok;
    }
#pragma unroll
    for (int k = 0; k < 3; k++) { x[k] = nx[k]; mu[k] = nm[k]; ab[k] = na[k]; og[k] = no[k]; }
  }

  // ---- groups: all three, else the closest real / conjugate pair with the pole of the third node
  // (|al|: per-matrix exponents may be <= 0 -- an extension; for al > 0 the reference's radius)
  const double rho_f = (KIND == POW && 3.0 / fabs(al) < RHO) ? 3.0 / fabs(al) : RHO;
  auto group_ok = [&](double c, double R) -> bool {
    if (KIND == EXP) return R <= R_EXP;
    return (m + c > 0) && (R <= rho_f * (m + c));
  };
  const double cs = (x[0].re + x[1].re + x[2].re) / 3.0;
  const double R3 = dmax(dmax(cabs_(mk(x[0].re - cs, x[0].im)), cabs_(mk(x[1].re - cs, x[1].im))),
                         cabs_(mk(x[2].re - cs, x[2].im)));
  const bool triple = group_ok(cs, R3);
  bool psel[3] = {false, false, false};
  double bestR = dinf();
  constexpr int PI_[3] = {0, 0, 1}, PJ_[3] = {1, 2, 2};
#pragma unroll
  for (int p = 0; p < 3; p++) {
    const int i = PI_[p], j = PJ_[p], k = 3 - i - j;
    const bool elig = ((x[i].im == 0) && (x[j].im == 0)) || ((x[i].re == x[j].re) && (x[i].im == -x[j].im));
    const double c = 0.5 * (x[i].re + x[j].re), R = 0.5 * cabs_(sub(x[i], x[j]));
    const bool okp = elig && group_ok(c, R) && (R <= RHO_POLE * cabs_(mk(c - x[k].re, -x[k].im))) && !triple;
    const bool better = okp && (R < bestR);
    bestR = better ? R : bestR;
#pragma unroll
    for (int q = 0; q < p; q++) psel[q] = psel[q] && !better;
    psel[p] = better;
  }
  const bool inG[3] = {triple || psel[0] || psel[1], triple || psel[0] || psel[2], triple || psel[1] || psel[2]};
  const bool pk[3] = {psel[2], psel[1], psel[0]};

  const double gsum = (inG[0] ? 1.0 : 0.0) + (inG[1] ? 1.0 : 0.0) + (inG[2] ? 1.0 : 0.0);
  double cxg = ((inG[0] ? x[0].re : 0.0) + (inG[1] ? x[1].re : 0.0) + (inG[2] ? x[2].re : 0.0)) / (gsum > 0 ? gsum : 1.0);
  double cmu = m + cxg;
  const bool ovr = pc && (gsum == 2.0) &&
                   ((inG[0] && inG[1] && og[0] > 0 && og[1] > 0) || (inG[0] && inG[2] && og[0] > 0 && og[2] > 0) ||
                    (inG[1] && inG[2] && og[1] > 0 && og[2] > 0));
  cxg = ovr ? pcen - m : cxg;
  cmu = ovr ? pcen : cmu;
  const cpx dpair = preal ? mk(ps, 0.0) : mk(0.0, ps);
  cpx off[3];
#pragma unroll
  for (int i = 0; i < 3; i++) {
    const bool neg = (mu[i].re - pcen < 0) || (mu[i].im < 0);
    const cpx di = neg ? mk(-dpair.re, -dpair.im) : dpair;
    const bool ovi = ovr && inG[i];
    mu[i] = ovi ? mk(pcen + di.re, di.im) : mu[i];
    x[i] = ovi ? mk(mu[i].re - m, mu[i].im) : x[i];
    off[i] = inG[i] ? (ovi ? di : mk(x[i].re - cxg, x[i].im)) : mk(0.0, 0.0);
  }
  // ---- single real nodes far from m: Newton
#pragma unroll
  for (int i = 0; i < 3; i++) {
    const bool mask = !inG[i] && !ab[i] && (x[i].im == 0) && (fabs(x[i].re) > 0.5 * fabs(m));
    if (mask) {
      newton_node(As, mu, i);
      ab[i] = true;
      x[i] = mk(mu[i].re - m, mu[i].im);
    }
  }
  const bool absc = (inG[0] && ab[0]) || (inG[1] && ab[1]) || (inG[2] && ab[2]);
  double apole = 0.0;
#pragma unroll
  for (int k = 0; k < 3; k++)
    apole = pk[k] ? ((absc || ab[k]) ? cmu - mu[k].re : cxg - x[k].re) : apole;
  const double xshift = KIND == EXP ? dmax(dmax(x[0].re, x[1].re), x[2].re) : 0.0;
  const bool bad = KIND != EXP && (((x[0].im == 0) && (mu[0].re <= 0)) || ((x[1].im == 0) && (mu[1].re <= 0)) ||
                                   ((x[2].im == 0) && (mu[2].re <= 0)));

  // ---- group series, streamed over the series index j (gmf_kernel.hpp:478-533)
  const bool hasg = inG[0] || inG[1] || inG[2];
  const double L = KIND == EXP ? 1.0 : ((hasg && cmu > 0) ? cmu : 1.0);
  const double iL = 1.0 / L;
  cpx u[3];
#pragma unroll
  for (int i = 0; i < 3; i++) u[i] = scale(off[i], iL);
  const double ap = fabs(apole) > 0 ? apole * iL : 1.0, iap = 1.0 / ap;
  double rr = dmax(cabs_(u[0]), dmax(cabs_(u[1]), cabs_(u[2])));
  const double rq = ((pk[0] || pk[1] || pk[2]) && (rr * fabs(iap) <= 1.0)) ? rr * fabs(iap) : 0.0;
  rr = (hasg && rr <= R_EXP + 0.5) ? rr : 0.0;
  {
    sin.K = hasg ? series_terms<PA>(bt, al, 3 * (ORDER + 1) - 1, rr, rq, rinv) : 0;
    sin.b0 = (hasg && KIND == LOG) ? node_value<LOG>(al, sig, m, mk(cxg, 0.0), mk(cmu, 0.0), 0).re : 1.0;
#pragma unroll
    for (int i = 0; i < 3; i++) sin.u[i] = u[i];
    sin.iap = iap;
    sin.al = al;
    sin.flags = (inG[0] ? 1 : 0) | (inG[1] ? 2 : 0) | (inG[2] ? 4 : 0) | (pk[0] ? 8 : 0) | (pk[1] ? 16 : 0) |
                (pk[2] ? 32 : 0);
  }
#pragma unroll
  for (int i = 0; i < 3; i++) { st.x[i] = x[i]; st.mu[i] = mu[i]; }
  st.m = m;
  st.sig = sig;
  st.xshift = xshift;
  st.cxg = cxg;
  st.L = L;
  st.flags = (ab[0] ? 1u : 0u) | (ab[1] ? 2u : 0u) | (ab[2] ? 4u : 0u) | (inG[0] ? 8u : 0u) | (inG[1] ? 16u : 0u) |
             (inG[2] ? 32u : 0u) | (bad ? 64u : 0u) | (hasg ? 128u : 0u);
  st.pad = 0;
}

// the rest of one matrix's evaluation from its nodes and group series sums acc (ND of them; gmf_kernel.hpp:478-671):
// the group's divided differences, the single-node residues, the real basis and the contracted tensors
template <int KIND, int ORDER, bool PA>
MF_DEV void gmf_finish(const double* a, double al, const GmfNodes& st, const cpx* acc, GmfOut& o) {
  constexpr int ND = ORDER == 0 ? 3 : (ORDER == 1 ? 9 : 19);
  const double sig = st.sig, isig = 1.0 / sig, m = st.m, xshift = st.xshift, cxg = st.cxg, L = st.L, iL = 1.0 / L;
  double As[9], S[9];
#pragma unroll
  for (int k = 0; k < 9; k++) As[k] = a[k] * isig;
#pragma unroll
  for (int k = 0; k < 9; k++) S[k] = k % 4 == 0 ? As[k] - m : As[k];
  cpx x[3], mu[3];
  bool ab[3], inG[3];
#pragma unroll
  for (int i = 0; i < 3; i++) {
    x[i] = st.x[i];
    mu[i] = st.mu[i];
    ab[i] = (st.flags >> i) & 1u;
    inG[i] = (st.flags >> (3 + i)) & 1u;
  }
  const bool bad = (st.flags & 64u) != 0, hasg = (st.flags & 128u) != 0;
  cpx D[ND];
#pragma unroll
  for (int c = 0; c < ND; c++) D[c] = mk(0.0, 0.0);
  {
    if (hasg) {
#pragma unroll
      for (int c = 0; c < ND; c++) {
        const int n = dd_a(c) + dd_b(c) + dd_c(c) - 1, k = dd_a(c) - 1;
        double fac;
        if (KIND == EXP) {
          fac = exp(cxg - xshift);
        } else {
          double iLn = 1.0;
#pragma unroll
          for (int r = 0; r < n - (KIND == POW ? k : 0); r++) iLn = iLn * iL;
          fac = KIND == LOG ? iLn : powk(sig * L, al, (double)k) * iLn;
        }
        D[c] = scale(acc[c], fac);
      }
    }
  }
  // ---- residues of the single nodes (gmf_kernel.hpp:534-576)
#pragma unroll
  for (int i = 0; i < 3; i++) {
    if (inG[i]) continue;
    const cpx xs = KIND == EXP ? mk(x[i].re - xshift, x[i].im) : x[i];
    cpx f0[3];
    f0[0] = node_value<KIND>(al, sig, m, xs, mu[i], 0);
    if (KIND == POW && ORDER >= 1) f0[1] = node_value<KIND>(al, sig, m, xs, mu[i], 1);
    if (KIND == POW && ORDER >= 2) f0[2] = node_value<KIND>(al, sig, m, xs, mu[i], 2);
    const cpx imu = cinv(mu[i]);
    cpx w[3];
#pragma unroll
    for (int z = 0; z < 3; z++)
      if (z != i) w[z] = cinv((ab[i] || ab[z]) ? sub(mu[i], mu[z]) : sub(x[i], x[z]));
#pragma unroll
    for (int c = 0; c < ND; c++) {
      const int mi = dd_mult(c, i);
      if (mi == 0) continue;
      const int k = dd_a(c) - 1;
      const cpx fb = KIND == POW ? f0[k] : f0[0];
      cpx s0, s1, s2;
      if (KIND == EXP) {
        s0 = fb; s1 = fb; s2 = scale(fb, 0.5);
      } else if (KIND == LOG) {
        s0 = fb; s1 = imu; s2 = scale(mul(imu, imu), -0.5);
      } else {
        s0 = fb;
        s1 = scale(mul(fb, imu), al);
        s2 = scale(mul(s1, imu), (al - 1.0) / 2.0);
      }
#pragma unroll
      for (int z = 0; z < 3; z++) {
        const int mz = dd_mult(c, z);
        if (z == i || mz == 0) continue;
        cpx wp = mk(1.0, 0.0);
#pragma unroll
        for (int r = 0; r < mz; r++) wp = mul(wp, w[z]);
        const cpx p1 = scale(mul(wp, w[z]), -(double)mz);
        const cpx p2 = scale(mul(mul(wp, w[z]), w[z]), mz * (mz + 1) / 2.0);
        const cpx n0 = mul(s0, wp), n1 = add(mul(s1, wp), mul(s0, p1));
        const cpx n2 = add(add(mul(s2, wp), mul(s1, p1)), mul(s0, p2));
        s0 = n0; s1 = n1; s2 = n2;
      }
      D[c] = add(D[c], mi == 1 ? s0 : (mi == 2 ? s1 : s2));
    }
  }

  // ---- real basis r_0 = I, r_1 = B_0, r_2 = B_0 B_1 - b0 b1 I and T = I + i N (gmf_kernel.hpp:578-599)
  const double b0i = x[0].im, b1i = x[1].im;
  double B0[9], B1[9];
#pragma unroll
  for (int k = 0; k < 9; k++) {
    const bool dg = k % 4 == 0;
    B0[k] = dg ? (ab[0] ? As[k] - mu[0].re : S[k] - x[0].re) : S[k];
    B1[k] = dg ? (ab[1] ? As[k] - mu[1].re : S[k] - x[1].re) : S[k];
    o.r[0][k] = dg ? 1.0 : 0.0;
    o.r[1][k] = B0[k];
  }
#pragma unroll
  for (int i2 = 0; i2 < 3; i2++)
#pragma unroll
    for (int j2 = 0; j2 < 3; j2++)
      o.r[2][3 * i2 + j2] = B0[3 * i2] * B1[j2] + B0[3 * i2 + 1] * B1[3 + j2] + B0[3 * i2 + 2] * B1[6 + j2] -
                            (i2 == j2 ? b0i * b1i : 0.0);
  const double a01 = (ab[0] || ab[1]) ? mu[0].re - mu[1].re : x[0].re - x[1].re;
  const double N10 = -b0i, N20 = -b0i * a01, N21 = -(b0i + b1i);
  const double nanf = bad ? dnan() : 1.0;
  {
    // F = sum_j Re(T^T d)_j r_j, (T^T v)_0 = v_0 + i (N10 v_1 + N20 v_2), (T^T v)_1 = v_1 + i N21 v_2
    const double Dr0 = D[pos1(0)].re, Dr1 = D[pos1(1)].re, Dr2 = D[pos1(2)].re;
    const double Di1 = D[pos1(1)].im, Di2 = D[pos1(2)].im;
    const double c0 = Dr0 - (N10 * Di1 + N20 * Di2), c1 = Dr1 - N21 * Di2, c2 = Dr2 - 0.0;
    const double e = KIND == EXP ? exp(m + xshift) : 1.0;
#pragma unroll
    for (int k = 0; k < 9; k++) o.F[k] = (c0 * o.r[0][k] + c1 * o.r[1][k] + c2 * o.r[2][k]) * e * nanf;
  }
  if (ORDER < 1) return;
  o.sc1 = (KIND == EXP ? exp(m + xshift) : (KIND == LOG ? isig : 1.0)) * nanf;
  {
    double Dr[9], Di[9];
#pragma unroll
    for (int j = 0; j < 3; j++)
#pragma unroll
      for (int k = 0; k < 3; k++) { Dr[3 * j + k] = D[pos2(j, k)].re; Di[3 * j + k] = D[pos2(j, k)].im; }
#pragma unroll
    for (int k = 0; k < 3; k++) {          // T on index 0
      const double vr0 = Dr[k], vr1 = Dr[3 + k], vr2 = Dr[6 + k], vi0 = Di[k], vi1 = Di[3 + k], vi2 = Di[6 + k];
      Dr[k] = vr0 - (N10 * vi1 + N20 * vi2); Di[k] = vi0 + (N10 * vr1 + N20 * vr2);
      Dr[3 + k] = vr1 - N21 * vi2;           Di[3 + k] = vi1 + N21 * vr2;
      Dr[6 + k] = vr2 - 0.0;                 Di[6 + k] = vi2 + 0.0;
    }
#pragma unroll
    for (int j = 0; j < 3; j++) {          // T on index 1, real part
      o.D2[3 * j + 0] = Dr[3 * j + 0] - (N10 * Di[3 * j + 1] + N20 * Di[3 * j + 2]);
      o.D2[3 * j + 1] = Dr[3 * j + 1] - N21 * Di[3 * j + 2];
      o.D2[3 * j + 2] = Dr[3 * j + 2] - 0.0;
    }
  }
  if (ORDER < 2) return;
  o.sc2 = (KIND == EXP ? exp(m + xshift) : (KIND == LOG ? isig * isig : 1.0)) * nanf;
  {
    double Dr[27], Di[27];
#pragma unroll
    for (int j = 0; j < 3; j++)
#pragma unroll
      for (int k = 0; k < 3; k++)
#pragma unroll
        for (int l = 0; l < 3; l++) {
          Dr[9 * j + 3 * k + l] = D[pos3(j, k, l)].re;
          Di[9 * j + 3 * k + l] = D[pos3(j, k, l)].im;
        }
    constexpr int STR[3] = {9, 3, 1};
#pragma unroll
    for (int ax = 0; ax < 3; ax++) {
      const int s = STR[ax];
#pragma unroll
      for (int base = 0; base < 27; base++) {
        if ((base / s) % 3 != 0) continue;
        const double vr0 = Dr[base], vr1 = Dr[base + s], vr2 = Dr[base + 2 * s];
        const double vi0 = Di[base], vi1 = Di[base + s], vi2 = Di[base + 2 * s];
        Dr[base] = vr0 - (N10 * vi1 + N20 * vi2);  Di[base] = vi0 + (N10 * vr1 + N20 * vr2);
        Dr[base + s] = vr1 - N21 * vi2;            Di[base + s] = vi1 + N21 * vr2;
        Dr[base + 2 * s] = vr2 - 0.0;              Di[base + 2 * s] = vi2 + 0.0;
      }
    }
#pragma unroll
    for (int k = 0; k < 27; k++) o.D3[k] = Dr[k];
  }
}

// one matrix: its nodes, the group series (series(in, acc): GmfSeriesLocal, one shared by the lanes of a workgroup,
// or sums computed elsewhere; called by every lane, K = 0 without a group), the rest.  With a K_ONLY functor (the
// first pass of the length-sorted evaluation) it stops after the nodes: the series length in o.K, the series inputs
// and the nodes handed to series.post
template <int KIND, int ORDER, bool PA, class SER>
MF_DEV void gmf_lane(const double* a, double al, const double* __restrict__ bt, GmfOut& o,
                     const double* __restrict__ rinv, const SER& series) {
  constexpr int ND = ORDER == 0 ? 3 : (ORDER == 1 ? 9 : 19);
  GmfNodes st;
  GmfSeriesIn sin;
  gmf_nodes<KIND, ORDER, PA>(a, al, bt, rinv, st, sin);
  if constexpr (SER::K_ONLY) {
    o.K = sin.K;
    o.cls = gmf_series_class(sin);
    series.post(sin, st);
    return;
  }
  cpx acc[ND];
  series(sin, acc);
  gmf_finish<KIND, ORDER, PA>(a, al, st, acc, o);
}

// ---- assembly of J and H from the staged per-matrix results of one matrix: r1, r2 (9 each, row-major; r_0 = I is
// implied), the distinct entries of the symmetric D2 (6, packed by sym3) and D3 (10, packed by sym33), and the scale
// factors.  D2 and D3 are symmetric: T acts alike on every index of the symmetric divided-difference tensors.
MF_DEV constexpr int sym3(int a, int b) { return a == b ? a : a + b + 2; }                  // (0,1) 3, (0,2) 4, (1,2) 5
MF_DEV constexpr int sym33s(int i, int j, int k) {                                          // i <= j <= k
  return i == 0 ? (j == 0 ? k : (j == 1 ? 2 + k : 5)) : (i == 1 ? (j == 1 ? 5 + k : 8) : 9);
}
MF_DEV constexpr int sym33(int i, int j, int k) {
  return sym33s(i < j ? (i < k ? i : k) : (j < k ? j : k),
                i + j + k - (i < j ? (i < k ? i : k) : (j < k ? j : k)) - (i > j ? (i > k ? i : k) : (j > k ? j : k)),
                i > j ? (i > k ? i : k) : (j > k ? j : k));
}

// element (row, col) of r_m (row, col may differ between lanes: r_0 = I by comparison)
template <class T>
MF_DEV T rget(const T* r1, const T* r2, int m, int row, int col) {
  return m == 0 ? (row == col ? (T)1 : (T)0) : (m == 1 ? r1[3 * row + col] : r2[3 * row + col]);
}

// J_abcd = sc1 sum_mn D2_mn r_m[a][c] r_n[d][b]  (gmf_kernel.hpp:636-647); lane (b, c, d): out[a], a = 0..2
template <class T>
MF_DEV void gmf_j_lane(const T* r1, const T* r2, const T* d2s, T sc1, int b, int c, int d, T* out) {
  T rdb[3], Y[3];
#pragma unroll
  for (int n = 0; n < 3; n++) rdb[n] = rget(r1, r2, n, d, b);
#pragma unroll
  for (int mm = 0; mm < 3; mm++) Y[mm] = d2s[sym3(mm, 0)] * rdb[0] + d2s[sym3(mm, 1)] * rdb[1] + d2s[sym3(mm, 2)] * rdb[2];
#pragma unroll
  for (int aa = 0; aa < 3; aa++)
    out[aa] = (rget(r1, r2, 0, aa, c) * Y[0] + rget(r1, r2, 1, aa, c) * Y[1] + rget(r1, r2, 2, aa, c) * Y[2]) * sc1;
}

// H_abcdef = sc2 (T6_acdefb + T6_aefcdb), T6_acdefb = sum_mnp D3_mnp r_m[a][c] r_n[d][e] r_p[f][b]
// (gmf_kernel.hpp:672-691), assembled by 27 lanes (d, e, f) per matrix in two steps.  With Z_mn[fb] = sum_p D3_mnp
// r_p[f][b] and Y_m[(de)(fb)] = sum_n r_n[d][e] Z_mn[fb]: H = sc2 (sum_m r_m[a][c] X1[m][b] + sum_m r_m[a][e]
// X2[m][c][b]) with X1[m][b] = Y_m[(de)(fb)] and X2[m][c][b] = Y_m[(fc)(db)] -- the X1 of lane (f, c, d).
// gmf_h_y: the lane's X1 (y[3 m + b]); gmf_h_out: its 27 outputs out[9a + 3b + c] from its X1 and the 27 lanes' X1
// (ys: lane L's at ys[9 L])
template <class T>
MF_DEV void gmf_h_y(const T* r1, const T* r2, const T* d3s, int d, int e, int f, T* y) {
  T rfb[3][3], rde[3];              // r_p[f][b], r_p[d][e]
#pragma unroll
  for (int p = 0; p < 3; p++) {
    rde[p] = rget(r1, r2, p, d, e);
#pragma unroll
    for (int b = 0; b < 3; b++) rfb[p][b] = rget(r1, r2, p, f, b);
  }
#pragma unroll
  for (int mm = 0; mm < 3; mm++) {
    T Z1[3][3];                     // [n][b]: Z[mn][fb]
#pragma unroll
    for (int nn = 0; nn < 3; nn++)
#pragma unroll
      for (int b = 0; b < 3; b++) {
        const T D0 = d3s[sym33(mm, nn, 0)], D1 = d3s[sym33(mm, nn, 1)], D2 = d3s[sym33(mm, nn, 2)];
        Z1[nn][b] = D0 * rfb[0][b] + D1 * rfb[1][b] + D2 * rfb[2][b];
      }
#pragma unroll
    for (int b = 0; b < 3; b++) y[3 * mm + b] = rde[0] * Z1[0][b] + rde[1] * Z1[1][b] + rde[2] * Z1[2][b];
  }
}
template <class T>
MF_DEV void gmf_h_out(const T* r1, const T* r2, T sc2, int d, int e, int f, const T* x1, const T* ys, T* out) {
#pragma unroll
  for (int aa = 0; aa < 3; aa++) {
    const T re0 = rget(r1, r2, 0, aa, e), re1 = rget(r1, r2, 1, aa, e), re2 = rget(r1, r2, 2, aa, e);
#pragma unroll
    for (int c = 0; c < 3; c++) {
      const T rc0 = rget(r1, r2, 0, aa, c), rc1 = rget(r1, r2, 1, aa, c), rc2 = rget(r1, r2, 2, aa, c);
      const T* x2 = ys + 9 * (9 * f + 3 * c + d);   // X2[m][c][b] = x2[3 m + b]
#pragma unroll
      for (int b = 0; b < 3; b++)
        out[9 * aa + 3 * b + c] = (rc0 * x1[b] + rc1 * x1[3 + b] + rc2 * x1[6 + b] + re0 * x2[b] + re1 * x2[3 + b] +
                                   re2 * x2[6 + b]) * sc2;
    }
  }
}

}  // namespace mf
