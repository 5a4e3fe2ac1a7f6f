// lanecheck.hip -- the HIP kernels' per-lane math (matfun_amd/csrc/*_device.hpp, __host__ __device__) run on the
// CPU, assembled into the API layouts the way the kernels assemble them, so that tests/test_lanecheck.py can compare
// it with the oracle without a GPU.  Test infrastructure: built by __graft_entry__.build(), loaded only by tests.
#include <stdint.h>
#include <string.h>

#include "gmf_device.hpp"
#include "hencky_device.hpp"
#include "philox.hpp"
#include "smf_device.hpp"

namespace {

// a per-matrix exponent as the kernels take it: a non-finite one computes with 1 and stores NaN
double exponent_of(const double* alv, int64_t m, double& nanf) {
  const double a = alv[m];
  const bool ok = fabs(a) <= 1.7976931348623157e308;
  nanf = ok ? 1.0 : mf::dnan();
  return ok ? a : 1.0;
}

// one matrix per call of the lane math; the outputs are gathered through the kernels' index maps (f_src, j_src, h_src).
// PA: powers with the per-matrix exponents alv
template <int KIND, bool PA = false>
void smf_any(bool mat, double al, const double* alv, int64_t n, const double* x, double* F, double* J, double* H) {
  mf::SmfCoef c2;
  mf::smf_coefs(KIND, al, c2);
  const int nf = mat ? 9 : 6, nj = nf * nf, nh = nj * nf;
  for (int64_t m = 0; m < n; m++) {
    double nanf = 1.0;
    if (PA) al = exponent_of(alv, m, nanf);
    double v[6];
    if (mat) mf::matrix_to_voigt(x + 9 * m, v);
    else for (int k = 0; k < 6; k++) v[k] = x[6 * m + k];
    mf::SmfLane<KIND, double> s;
    double Fv[6], Js[21], Hv[126];
    s.eigen(v, al, Fv);
    for (int e = 0; e < nf; e++) F[nf * m + e] = Fv[mf::f_src(mat, e)] * nanf;
    if (!J) continue;
    s.jacobian(al, Js);
    for (int e = 0; e < nj; e++) {
      double w;
      const int src = mf::j_src(mat, e, w);
      J[nj * m + e] = Js[src] * nanf * w;
    }
    if (!H) continue;
    s.template second_dd<PA>(al, c2.c);
    for (int k = 0; k < 10; k++) s.D2[k] *= nanf;
    double V9[9];
    for (int k = 0; k < 9; k++) V9[k] = s.V[k / 3][k % 3];
    const mf::SmfHess<double> hs(V9, s.D2);
    for (int i = 0; i < 6; i++) hs.h_row(i, Hv + 21 * i);
    for (int e = 0; e < nh; e++) {
      double w;
      const int src = mf::h_src(mat, e, w);
      H[nh * m + e] = Hv[src] * w;
    }
  }
}

template <int KIND, int ORDER, bool PA = false>
void gmf_all(double al, const double* alv, int64_t n, const double* A, double* F, double* J, double* H) {
  mf::GmfCoef bt;
  if (PA) mf::gmf_recips(bt);
  else mf::gmf_coefs(KIND, al, bt);
  for (int64_t m = 0; m < n; m++) {
    double nanf = 1.0;
    if (PA) al = exponent_of(alv, m, nanf);
    mf::GmfOut o;
    static constexpr mf::GmfRecip rinv = mf::gmf_recip_table();
    mf::gmf_lane<KIND, ORDER, PA>(A + 9 * m, al, bt.b, o, rinv.r, mf::GmfSeriesLocal<KIND, ORDER, PA>{bt.b});
    for (int k = 0; k < 9; k++) o.F[k] *= nanf;
    o.sc1 *= nanf;
    o.sc2 *= nanf;
    for (int k = 0; k < 9; k++) F[9 * m + k] = o.F[k];
    if (ORDER < 1) continue;
    // the values the kernel stages: r1, r2, the distinct entries of the symmetric D2 / D3
    double r1[9], r2[9], d2s[6], d3s[10], out[27];
    for (int k = 0; k < 9; k++) { r1[k] = o.r[1][k]; r2[k] = o.r[2][k]; }
    for (int i = 0; i < 3; i++)
      for (int j = i; j < 3; j++) d2s[mf::sym3(i, j)] = o.D2[3 * i + j];
    for (int L = 0; L < 27; L++) {
      mf::gmf_j_lane(r1, r2, d2s, o.sc1, L / 9, (L / 3) % 3, L % 3, out);
      for (int a = 0; a < 3; a++) J[81 * m + 27 * a + L] = out[a];
    }
    if (ORDER < 2) continue;
    for (int i = 0; i < 3; i++)
      for (int j = i; j < 3; j++)
        for (int k = j; k < 3; k++) d3s[mf::sym33s(i, j, k)] = o.D3[9 * i + 3 * j + k];
    double ys[27 * 9];
    for (int L = 0; L < 27; L++) mf::gmf_h_y(r1, r2, d3s, L / 9, (L / 3) % 3, L % 3, ys + 9 * L);
    for (int L = 0; L < 27; L++) {
      mf::gmf_h_out(r1, r2, o.sc2, L / 9, (L / 3) % 3, L % 3, ys + 9 * L, ys, out);
      for (int k = 0; k < 27; k++) H[729 * m + 27 * k + L] = out[k];
    }
  }
}

// the group series of each matrix through the complex loop (gmf_series) and, for a lane whose offsets are real,
// through the real loops (gmf_series_real with and, without a pole node, without the pole factors): cls[m] the
// lane's class, err[3 m ..] the normwise relative differences of the real loops' sums (0 where not run) and the
// largest imaginary part of the complex loop's sums for a real lane
template <int KIND, int ORDER>
void gmf_series_check(double al, int64_t n, const double* A, int* cls, double* err) {
  constexpr int ND = ORDER == 0 ? 3 : (ORDER == 1 ? 9 : 19);
  mf::GmfCoef bt;
  mf::gmf_coefs(KIND, al, bt);
  static constexpr mf::GmfRecip rinv = mf::gmf_recip_table();
  for (int64_t m = 0; m < n; m++) {
    mf::GmfNodes st;
    mf::GmfSeriesIn s;
    mf::gmf_nodes<KIND, ORDER, false>(A + 9 * m, al, bt.b, rinv.r, st, s);
    mf::cpx ref[ND], r0[ND], r1[ND];
    mf::gmf_series<KIND, ORDER, false>(s, bt.b, ref);
    const int c = mf::gmf_series_class(s);
    cls[m] = c;
    double nrm = 0.0, d0 = 0.0, d1 = 0.0, imax = 0.0;
    for (int k = 0; k < ND; k++) nrm += ref[k].re * ref[k].re + ref[k].im * ref[k].im;
    if (c <= 1) {
      mf::gmf_series_real<KIND, ORDER, false, true>(s, bt.b, r1);
      for (int k = 0; k < ND; k++) {
        d1 += (r1[k].re - ref[k].re) * (r1[k].re - ref[k].re) + (r1[k].im - ref[k].im) * (r1[k].im - ref[k].im);
        imax = fabs(ref[k].im) > imax ? fabs(ref[k].im) : imax;
      }
    }
    if (c == 0) {
      mf::gmf_series_real<KIND, ORDER, false, false>(s, bt.b, r0);
      for (int k = 0; k < ND; k++)
        d0 += (r0[k].re - ref[k].re) * (r0[k].re - ref[k].re) + (r0[k].im - ref[k].im) * (r0[k].im - ref[k].im);
    }
    err[3 * m] = nrm > 0 ? sqrt(d0 / nrm) : sqrt(d0);
    err[3 * m + 1] = nrm > 0 ? sqrt(d1 / nrm) : sqrt(d1);
    err[3 * m + 2] = imax;
  }
}

// the staged nodes (gmf_stage_nodes -> gmf_unstage_nodes) of each matrix's log order 2 nodes: nbad = matrices whose
// nodes do not come back bitwise (bit patterns, so the signs of zeros count); ncplx = matrices stored with their
// imaginary parts
void gmf_stage_check(int64_t n, const double* A, int64_t* nbad, int64_t* ncplx) {
  mf::GmfCoef bt;
  mf::gmf_coefs(mf::LOG, 0.0, bt);
  static constexpr mf::GmfRecip rinv = mf::gmf_recip_table();
  *nbad = *ncplx = 0;
  for (int64_t m = 0; m < n; m++) {
    mf::GmfNodes st;
    mf::GmfSeriesIn s;
    mf::gmf_nodes<mf::LOG, 2, false>(A + 9 * m, 0.0, bt.b, rinv.r, st, s);
    mf::GmfNodesRe r;
    double im[6] = {7.0, 7.0, 7.0, 7.0, 7.0, 7.0};
    mf::gmf_stage_nodes(st, mf::gmf_series_class(s), r, im);
    const mf::GmfNodes b = mf::gmf_unstage_nodes(r, im);
    *ncplx += (r.flags & mf::GMF_NODES_CPLX) ? 1 : 0;
    bool ok = ((r.flags >> 9) & 3u) == (unsigned)mf::gmf_series_class(s) && b.flags == st.flags;
    const double va[16] = {st.x[0].re, st.x[0].im, st.x[1].re, st.x[1].im, st.x[2].re, st.x[2].im, st.mu[0].re,
                           st.mu[0].im, st.mu[1].re, st.mu[1].im, st.mu[2].re, st.mu[2].im, st.m, st.sig, st.xshift,
                           st.cxg};
    const double vb[16] = {b.x[0].re, b.x[0].im, b.x[1].re, b.x[1].im, b.x[2].re, b.x[2].im, b.mu[0].re,
                           b.mu[0].im, b.mu[1].re, b.mu[1].im, b.mu[2].re, b.mu[2].im, b.m, b.sig, b.xshift, b.cxg};
    ok = ok && memcmp(va, vb, sizeof(va)) == 0 && memcmp(&st.L, &b.L, sizeof(double)) == 0;
    *nbad += ok ? 0 : 1;
  }
}

}  // namespace

extern "C" void lanecheck_gmf_stage(int64_t n, const double* A, int64_t* nbad, int64_t* ncplx) {
  gmf_stage_check(n, A, nbad, ncplx);
}

extern "C" int lanecheck_gmf_series_real(int kind, double al, int order, int64_t n, const double* A, int* cls,
                                         double* err) {
  if (kind == mf::SQRT) kind = mf::POW, al = 0.5;
  if (kind != mf::POW) al = 0.0;
#define MF_SC(K)                                                         \
  if (order == 2) gmf_series_check<K, 2>(al, n, A, cls, err);            \
  else if (order == 1) gmf_series_check<K, 1>(al, n, A, cls, err);       \
  else gmf_series_check<K, 0>(al, n, A, cls, err);
  switch (kind) {
    case mf::EXP: MF_SC(mf::EXP) break;
    case mf::LOG: MF_SC(mf::LOG) break;
    default: MF_SC(mf::POW) break;
  }
#undef MF_SC
  return 0;
}

static int smf_entry(bool mat, int kind, double al, int64_t n, const double* x, double* F, double* J, double* H) {
  if (kind != mf::POW) al = kind == mf::SQRT ? 0.5 : 0.0;
  switch (kind) {
    case mf::EXP: smf_any<mf::EXP>(mat, al, nullptr, n, x, F, J, H); break;
    case mf::LOG: smf_any<mf::LOG>(mat, al, nullptr, n, x, F, J, H); break;
    case mf::SQRT: smf_any<mf::SQRT>(mat, al, nullptr, n, x, F, J, H); break;
    default: smf_any<mf::POW>(mat, al, nullptr, n, x, F, J, H); break;
  }
  return 0;
}
extern "C" int lanecheck_smf(int kind, double al, int64_t n, const double* v, double* F, double* J, double* H) {
  return smf_entry(false, kind, al, n, v, F, J, H);
}
extern "C" int lanecheck_smf_matrix(int kind, double al, int64_t n, const double* A, double* F, double* J, double* H) {
  return smf_entry(true, kind, al, n, A, F, J, H);
}

template <int KIND, bool PA = false>
static void gmf_order(int order, double al, const double* alv, int64_t n, const double* A, double* F, double* J,
                      double* H) {
  if (order == 2) gmf_all<KIND, 2, PA>(al, alv, n, A, F, J, H);
  else if (order == 1) gmf_all<KIND, 1, PA>(al, alv, n, A, F, J, H);
  else gmf_all<KIND, 0, PA>(al, alv, n, A, F, J, H);
}

extern "C" int lanecheck_gmf(int kind, double al, int64_t n, const double* A, double* F, double* J, double* H) {
  if (kind == mf::SQRT) kind = mf::POW, al = 0.5;
  if (kind != mf::POW) al = 0.0;
  const int order = H ? 2 : (J ? 1 : 0);
  switch (kind) {
    case mf::EXP: gmf_order<mf::EXP>(order, al, nullptr, n, A, F, J, H); break;
    case mf::LOG: gmf_order<mf::LOG>(order, al, nullptr, n, A, F, J, H); break;
    default: gmf_order<mf::POW>(order, al, nullptr, n, A, F, J, H); break;
  }
  return 0;
}

// powers with one exponent per matrix (host array alv)
extern "C" int lanecheck_smf_pow(const double* alv, int64_t n, const double* v, double* F, double* J, double* H) {
  smf_any<mf::POW, true>(false, 1.0, alv, n, v, F, J, H);
  return 0;
}
extern "C" int lanecheck_smf_matrix_pow(const double* alv, int64_t n, const double* A, double* F, double* J,
                                        double* H) {
  smf_any<mf::POW, true>(true, 1.0, alv, n, A, F, J, H);
  return 0;
}
extern "C" int lanecheck_gmf_pow(const double* alv, int64_t n, const double* A, double* F, double* J, double* H) {
  gmf_order<mf::POW, true>(H ? 2 : (J ? 1 : 0), 1.0, alv, n, A, F, J, H);
  return 0;
}

// the fused Hencky step of symmetric matrices (hencky_device.hpp) in the matrix layout, as smf_hencky_kernel stores it
extern "C" int lanecheck_smf_hencky(int64_t n, const double* A, double* L, double* psi, double* g, double* K) {
  mf::SmfCoef c2;
  mf::smf_coefs(mf::LOG, 0.0, c2);
  for (int64_t m = 0; m < n; m++) {
    double v[6], Lv[6], gv[6], kv[21], ps;
    mf::matrix_to_voigt(A + 9 * m, v);
    mf::smf_hencky_lane(v, c2.c, Lv, gv, kv, ps);
    psi[m] = ps;
    for (int e = 0; e < 9; e++) {
      L[9 * m + e] = Lv[mf::voigt_of_m(e)];
      g[9 * m + e] = gv[mf::voigt_of_m(e)];
    }
    for (int e = 0; e < 81; e++) K[81 * m + e] = kv[mf::tri6(mf::voigt_of_m(e / 9), mf::voigt_of_m(e % 9))];
  }
  return 0;
}

// Philox4x32-10 and the perturbation factors of the Newton driver (philox.hpp), on the CPU
extern "C" void lanecheck_philox(const uint32_t* ctr, const uint32_t* key, uint32_t* out) {
  mf::U4 c;
  for (int k = 0; k < 4; k++) c.x[k] = ctr[k];
  const mf::U4 r = mf::philox4x32_10(c, key[0], key[1]);
  for (int k = 0; k < 4; k++) out[k] = r.x[k];
}
extern "C" void lanecheck_perturb(uint64_t seed, uint32_t iteration, int64_t n, const int32_t* sym, double sigma,
                                  double* f9) {
  for (int64_t m = 0; m < n; m++) mf::perturb_factors(seed, iteration, (uint64_t)m, sym[m] != 0, sigma, f9 + 9 * m);
}
