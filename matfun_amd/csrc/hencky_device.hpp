// hencky_device.hpp -- the Newton step of the Hencky-type potential psi(X) = 1/2 |log X|^2 for one symmetric matrix,
// fused with its logarithm (reference theory.md:190: the Newton method for this potential with the tangent
// J^T J + log X : H).  __host__ __device__: the fused kernel (hencky.hip) and the CPU build of the lane math
// (tests/lanecheck) run the same code.
//
// From the smf lane math (SmfLane: eigen-decomposition, log, the 21 distinct J values Js, the 10 second divided
// differences; SmfHess::h_row: one row of the Voigt second derivatives Hv at a time), in the Voigt basis, with
// m_r = 1 (r < 3) or 2 (the number of matrix entries of Voigt index r) and w_p = dv_p/dX = 1 or 1/2:
//   psi = 1/2 sum_r m_r L_r^2,   g_p = sum_r m_r L_r Js_rp,   K_pq = sum_r m_r (Js_rp Js_rq + L_r Hv_r,pq w_p w_q)
// These are the matrix-layout sums over a = 0..8 of psi = 1/2 L_a L_a, g_c = L_a J_ac, K_ce = J_ac J_ae + L_a H_ace,
// two equal terms at a time, since J_ac = Js[v(a), v(c)] and H_ace = Hv[v(a)][v(c), v(e)] w_c w_e in the matrix
// layout (j_src, h_src); the matrix layout of g and K: g_c = g[v(c)], K_ce = K[tri6(v(c), v(e))].
#pragma once
#include "smf_device.hpp"

namespace mf {

// v: the Voigt vector of X (NaN for a non-symmetric X: NaN results); c2: smf_coefs(LOG, 0).  Lv: log X (Voigt),
// gv: g by Voigt index, kv: K by tri6(p, q), psi
MF_DEV void smf_hencky_lane(const double* v, const double* __restrict__ c2, double* Lv, double* gv, double* kv,
                            double& ps) {
  SmfLane<LOG, double> s;
  double Js[21];
  s.eigen(v, 0.0, Lv);
  s.jacobian(0.0, Js);
  s.template second_dd<false, SMF_NTERMS>(0.0, c2);
  ps = 0.0;
#pragma unroll
  for (int r = 0; r < 6; r++) ps = fma(r < 3 ? Lv[r] : 2.0 * Lv[r], Lv[r], ps);
#pragma unroll
  for (int p = 0; p < 6; p++) {
    double acc = 0.0;
#pragma unroll
    for (int r = 0; r < 6; r++) acc = fma(r < 3 ? Lv[r] : 2.0 * Lv[r], Js[tri6(r, p)], acc);
    gv[p] = acc;
  }
#pragma unroll
  for (int p = 0; p < 6; p++)
#pragma unroll
    for (int q = p; q < 6; q++) {
      double acc = 0.0;
#pragma unroll
      for (int r = 0; r < 6; r++) acc = fma(r < 3 ? Js[tri6(r, p)] : 2.0 * Js[tri6(r, p)], Js[tri6(r, q)], acc);
      kv[tri6(p, q)] = acc;
    }
  {
    double V9[9];
#pragma unroll
    for (int k = 0; k < 9; k++) V9[k] = s.V[k / 3][k % 3];
    const SmfHess<double> hs(V9, s.D2);
#pragma unroll
    for (int r = 0; r < 6; r++) {
      double h[21];
      hs.h_row(r, h);
      const double c = r < 3 ? Lv[r] : 2.0 * Lv[r];
#pragma unroll
      for (int p = 0; p < 6; p++)
#pragma unroll
        for (int q = p; q < 6; q++) {
          const double w = (p < 3 ? 1.0 : 0.5) * (q < 3 ? 1.0 : 0.5);
          kv[tri6(p, q)] = fma(c, h[tri6(p, q)] * w, kv[tri6(p, q)]);
        }
    }
  }
  ps *= 0.5;
}

}  // namespace mf
