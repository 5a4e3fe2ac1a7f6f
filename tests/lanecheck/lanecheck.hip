// lanecheck.hip -- the HIP kernels' per-lane math (matfun_amd/csrc/*_device.hpp, __host__ __device__) run on the
// CPU, assembled into the API layouts the way the kernels assemble them, so that tests/test_lanecheck.py can compare
// it with the oracle without a GPU.  Test infrastructure: built by __graft_entry__.build(), loaded only by tests.
#include <stdint.h>

#include "gmf_device.hpp"
#include "smf_device.hpp"

namespace {

template <int KIND>
void smf_voigt(double al, int64_t n, const double* v, double* F, double* J, double* H) {
  mf::SmfCoef c2;
  mf::smf_coefs(KIND, al, c2);
  for (int64_t m = 0; m < n; m++) {
    mf::SmfLane<KIND> s;
    double Fv[6], Js[21], h[21];
    s.eigen(v + 6 * m, al, Fv);
    for (int k = 0; k < 6; k++) F[6 * m + k] = Fv[k];
    if (!J) continue;
    s.jacobian(al, Js);
    for (int i = 0; i < 6; i++)
      for (int k = 0; k < 6; k++) J[36 * m + 6 * i + k] = Js[mf::tri6(i, k)] * (k < 3 ? 1.0 : 2.0);
    if (!H) continue;
    s.prepare_hessian(al, c2.c);
    for (int i = 0; i < 6; i++) {
      s.h_row(i, h);
      for (int p = 0; p < 6; p++)
        for (int q = 0; q < 6; q++) H[216 * m + 36 * i + 6 * p + q] = h[mf::tri6(p, q)];
    }
  }
}

template <int KIND, int ORDER>
void gmf_all(double al, int64_t n, const double* A, double* F, double* J, double* H) {
  mf::GmfCoef bt;
  mf::gmf_coefs(KIND, al, bt);
  for (int64_t m = 0; m < n; m++) {
    mf::GmfOut o;
    mf::gmf_lane<KIND, ORDER>(A + 9 * m, al, bt.b, o);
    for (int k = 0; k < 9; k++) F[9 * m + k] = o.F[k];
    if (ORDER < 1) continue;
    double r[27], out[27];
    for (int k = 0; k < 27; k++) r[k] = o.r[k / 9][k % 9];
    for (int L = 0; L < 27; L++) {
      mf::gmf_j_lane(r, o.D2, o.sc1, L / 9, (L / 3) % 3, L % 3, out);
      for (int a = 0; a < 3; a++) J[81 * m + 27 * a + L] = out[a];
    }
    if (ORDER < 2) continue;
    for (int L = 0; L < 27; L++) {
      mf::gmf_h_lane(r, o.D3, o.sc2, L / 9, (L / 3) % 3, L % 3, out);
      for (int k = 0; k < 27; k++) H[729 * m + 27 * k + L] = out[k];
    }
  }
}

}  // namespace

extern "C" int lanecheck_smf(int kind, double al, int64_t n, const double* v, double* F, double* J, double* H) {
  if (kind != mf::POW) al = kind == mf::SQRT ? 0.5 : 0.0;
  switch (kind) {
    case mf::EXP: smf_voigt<mf::EXP>(al, n, v, F, J, H); break;
    case mf::LOG: smf_voigt<mf::LOG>(al, n, v, F, J, H); break;
    case mf::SQRT: smf_voigt<mf::SQRT>(al, n, v, F, J, H); break;
    default: smf_voigt<mf::POW>(al, n, v, F, J, H); break;
  }
  return 0;
}

template <int KIND>
static void gmf_order(int order, double al, int64_t n, const double* A, double* F, double* J, double* H) {
  if (order == 2) gmf_all<KIND, 2>(al, n, A, F, J, H);
  else if (order == 1) gmf_all<KIND, 1>(al, n, A, F, J, H);
  else gmf_all<KIND, 0>(al, n, A, F, J, H);
}

extern "C" int lanecheck_gmf(int kind, double al, int64_t n, const double* A, double* F, double* J, double* H) {
  if (kind == mf::SQRT) kind = mf::POW, al = 0.5;
  if (kind != mf::POW) al = 0.0;
  const int order = H ? 2 : (J ? 1 : 0);
  switch (kind) {
    case mf::EXP: gmf_order<mf::EXP>(order, al, n, A, F, J, H); break;
    case mf::LOG: gmf_order<mf::LOG>(order, al, n, A, F, J, H); break;
    default: gmf_order<mf::POW>(order, al, n, A, F, J, H); break;
  }
  return 0;
}
