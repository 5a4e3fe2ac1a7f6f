// hencky.hip -- the Newton step after the matrix logarithm: value, gradient and tangent of the Hencky-type potential
// psi(X) = 1/2 |log X|^2 from L = log X, J = dL/dX and H = d2L/dX2 in the matrix layout (reference theory.md:190:
// the Newton method for this potential with the tangent J^T J + log X : H), per matrix
//   psi = 1/2 L_a L_a,   g_c = L_a J_ac,   K_ce = J_ac J_ae + L_a H_ace     (a, c, e = 0..8, summed over a)
// so that g = dpsi/dX and K = dg/dX.  No counterpart in the reference's API.
//
// HBM-bound streaming (per matrix: L 9, J 81, H 729 values read, psi 1, g 9, K 81 written): a workgroup of 256
// work-items takes 32 consecutive matrices; their L and J (contiguous blocks) are staged in LDS with coalesced
// loads, then work-item i of the block's 32 x 81 outputs K_(m, c, e) reads H_(m, a, c, e) for a = 0..8 straight
// from HBM (consecutive i: consecutive addresses; every H value is read once) and writes K in the same flat order.
// Sums in FP64 (FMA), also for the FP32 form (FP32 in and out, as the *_mixed evaluations give them).
//
// smf_hencky_kernel: the same step fused with the evaluation for symmetric X (matrix layout, FP64): one lane per
// matrix computes log X, the 21 distinct J values and the 10 second divided differences as the smf kernels do
// (SmfLane), then contracts in the Voigt basis without forming J or H in the API layout (smf_hencky_lane,
// hencky_device.hpp).  Outputs (log X, psi, g, K: 91 + 9 values, 728 B + 72 B per matrix instead
// of 6.6 KB of J and H written and read again) staged per matrix in LDS and written as contiguous blocks.
//
// gmf_hencky_kernel: the same for general real X (FP64): the gmf tile's phase 1 per lane (gmf_lane), then groups of
// six matrices: their J (9 lanes each, gmf_j_lane) into an LDS tile, K = J^T J and g from it, then H passes of two
// matrices in which the H lanes (27 per matrix, lane (d, e, f) holding H_(ab)(cd)(ef) for all (a, b, c): gmf_h_y /
// gmf_h_out) add L_(ab) H_(ab)(cd)(ef) to the three entries K_(cd)(ef), c = 0..2, that no other lane touches.  The
// sums run over a = 0..8 in the order of hencky_kernel on the same J and H values.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/matfun_hip.h"
#include "gmf_tile.hpp"
#include "hencky_device.hpp"
#include "tile_rows.hpp"

namespace mf {
namespace {
constexpr int MB = 32;    // matrices per workgroup
constexpr int NT = 256;   // work-items per workgroup
}  // namespace

template <class T>
__global__ void __launch_bounds__(NT) hencky_kernel(int64_t n, const T* __restrict__ L, const T* __restrict__ J,
                                                    const T* __restrict__ H, T* __restrict__ psi, T* __restrict__ g,
                                                    T* __restrict__ K) {
  __shared__ T sJ[MB * 81];
  __shared__ T sL[MB * 9];
  const int64_t m0 = (int64_t)blockIdx.x * MB;
  const int nb = n - m0 < MB ? (int)(n - m0) : MB;
  const int tid = threadIdx.x;
  if (g || K) {                                  // (J may be NULL when only psi is asked for)
    for (int i = tid; i < nb * 81; i += NT) sJ[i] = J[m0 * 81 + i];
  }
  for (int i = tid; i < nb * 9; i += NT) sL[i] = L[m0 * 9 + i];
  __syncthreads();
  if (K) {
    const T* Hb = H + m0 * 729;
    for (int i = tid; i < nb * 81; i += NT) {
      const int m = i / 81, t = i - 81 * m, c = t / 9, e = t - 9 * c;
      const T* hm = Hb + 729 * m + t;
      double h[9];
#pragma unroll
      for (int a = 0; a < 9; a++) h[a] = (double)hm[81 * a];
      const T* jm = sJ + 81 * m;
      const T* lm = sL + 9 * m;
      double acc = 0.0;
#pragma unroll
      for (int a = 0; a < 9; a++) acc = fma((double)jm[9 * a + c], (double)jm[9 * a + e], acc);
#pragma unroll
      for (int a = 0; a < 9; a++) acc = fma((double)lm[a], h[a], acc);
      K[m0 * 81 + i] = (T)acc;
    }
  }
  if (g) {
    for (int i = tid; i < nb * 9; i += NT) {
      const int m = i / 9, c = i - 9 * m;
      double acc = 0.0;
#pragma unroll
      for (int a = 0; a < 9; a++) acc = fma((double)sL[9 * m + a], (double)sJ[81 * m + 9 * a + c], acc);
      g[m0 * 9 + i] = (T)acc;
    }
  }
  if (psi) {
    for (int i = tid; i < nb; i += NT) {
      double acc = 0.0;
#pragma unroll
      for (int a = 0; a < 9; a++) acc = fma((double)sL[9 * i + a], (double)sL[9 * i + a], acc);
      psi[m0 + i] = (T)(0.5 * acc);
    }
  }
}

// arguments as the entry points take them; batches above 2^30 matrices go in several launches on the stream
template <class T>
int hencky_launch(int64_t n, const T* L, const T* J, const T* H, T* psi, T* g, T* K, void* stream) {
  if (n < 0 || ((psi || g || K) && n > 0 && !L) || ((g || K) && n > 0 && !J) || (K && n > 0 && !H)) return -1;
  if (n == 0 || (!psi && !g && !K)) return 0;
  const hipStream_t st = (hipStream_t)stream;
  constexpr int64_t CHUNK = (int64_t)1 << 30;
  for (int64_t b = 0; b < n; b += CHUNK) {
    const int64_t c = n - b < CHUNK ? n - b : CHUNK;
    hencky_kernel<T><<<dim3((unsigned)((c + MB - 1) / MB)), dim3(NT), 0, st>>>(
        c, L + b * 9, J ? J + b * 81 : nullptr, K ? H + b * 729 : nullptr, psi ? psi + b : nullptr,
        g ? g + b * 9 : nullptr, K ? K + b * 81 : nullptr);
    if (hipGetLastError() != hipSuccess) return -2;
  }
  return 0;
}

constexpr int SK = 35;    // fused kernel: staged values per matrix (K 21, g 6, psi, log X 6; odd stride)

// MAT: X and the outputs in the matrix layout (log X 9, g 9, K 81; g_c = g[v(c)], K_ce = K[tri6(v(c), v(e))]); else
// in Voigt coordinates v (log X 6, g_p = dpsi/dv_p = m_p g[p], K_pq = d2psi/dv_p dv_q = m_p m_q K[tri6(p, q)], 36)
template <bool MAT>
__global__ void __launch_bounds__(WAVE) __attribute__((amdgpu_waves_per_eu(2)))
smf_hencky_kernel(int64_t n, const double* __restrict__ x, double* __restrict__ Lo, double* __restrict__ psi,
                  double* __restrict__ g, double* __restrict__ K, SmfCoef c2) {
  constexpr int NIN = MAT ? 9 : 6;
  __shared__ __attribute__((aligned(16))) double st[WAVE * SK];   // (the 64 x NIN input block first)
  const int64_t base = (int64_t)blockIdx.x * WAVE;
  const int nm = (int)(n - base < WAVE ? n - base : WAVE);
  const int lane = threadIdx.x;
  const RunRows rows{base};
  load_in<NIN>(st, x, rows, nm, lane);
  __syncthreads();
  double v[6];
  if (MAT) {
    double a[9];
#pragma unroll
    for (int k = 0; k < 9; k++) a[k] = lane < nm ? st[9 * lane + k] : (k % 4 == 0 ? 1.0 : 0.0);
    matrix_to_voigt(a, v);                       // (NaN for a non-symmetric matrix)
  } else {
#pragma unroll
    for (int k = 0; k < 6; k++) v[k] = lane < nm ? st[6 * lane + k] : (k < 3 ? 1.0 : 0.0);
  }
  __syncthreads();
  double Lv[6], gv[6], kv[21], ps;
  smf_hencky_lane(v, c2.c, Lv, gv, kv, ps);
  double* sm = st + SK * lane;
#pragma unroll
  for (int k = 0; k < 21; k++) sm[k] = kv[k];
#pragma unroll
  for (int k = 0; k < 6; k++) sm[21 + k] = gv[k];
  sm[27] = ps;
#pragma unroll
  for (int k = 0; k < 6; k++) sm[28 + k] = Lv[k];
  __syncthreads();
  if (psi) copy_out<1, 1, R_IN>(psi, rows, nm, lane, [&](int j, int) { return st[SK * j + 27]; });
  if constexpr (MAT) {
    if (Lo) copy_out<9, 1, R_IN>(Lo, rows, 9 * nm, lane, [&](int j, int e) { return st[SK * j + 28 + voigt_of_m(e)]; });
    if (g) copy_out<9, 1, R_IN>(g, rows, 9 * nm, lane, [&](int j, int e) { return st[SK * j + 21 + voigt_of_m(e)]; });
    if (K)
      copy_out<81, 1, R_IN>(K, rows, 81 * nm, lane, [&](int j, int e) {
        return st[SK * j + tri6(voigt_of_m(e / 9), voigt_of_m(e % 9))];
      });
  } else {                                       // (m_p = 1 or 2: exact)
    if (Lo) copy_out<6, 1, R_IN>(Lo, rows, 6 * nm, lane, [&](int j, int e) { return st[SK * j + 28 + e]; });
    if (g)
      copy_out<6, 1, R_IN>(g, rows, 6 * nm, lane,
                           [&](int j, int e) { return (e < 3 ? 1.0 : 2.0) * st[SK * j + 21 + e]; });
    if (K)
      copy_out<36, 1, R_IN>(K, rows, 36 * nm, lane, [&](int j, int e) {
        const int p = e / 6, q = e - 6 * p;
        return (p < 3 ? 1.0 : 2.0) * (q < 3 ? 1.0 : 2.0) * st[SK * j + tri6(p, q)];
      });
  }
}

// LDS of the fused gmf kernel, in doubles: r1, r2, D3 per matrix [64][28] (as the gmf tile), the J tile of a group
// of six matrices [6][81] (the X1 tables of the two H lane groups [2][243] once K = J^T J is formed): 18.2 KB, eight
// workgroups per CU (K goes to its output and is completed there; log X stays in its lane's registers and reaches
// the J / H lanes by shuffles)
struct GhLds {
  static constexpr int MS = 28, GM = 6, JB = WAVE * MS, TOT = JB + GM * 81;
  static_assert(TOT >= WAVE * 9 && GM * 81 >= 2 * 243, "the input block overlays the start; X1 tables in the J tile");
};

__global__ void __launch_bounds__(WAVE) gmf_hencky_kernel(int64_t n, const double* __restrict__ A,
                                                          double* __restrict__ Lo, double* __restrict__ psi,
                                                          double* __restrict__ g, double* __restrict__ K, GmfCoef bt) {
  __shared__ __attribute__((aligned(16))) double lds[GhLds::TOT];
  constexpr int GM = GhLds::GM;
  const int64_t base = (int64_t)blockIdx.x * WAVE;
  const int nm = (int)(n - base < WAVE ? n - base : WAVE);
  const int lane = threadIdx.x;
  double* mainr = lds;
  double* jt = lds + GhLds::JB;
  load_in<9>(lds, A, RunRows{base}, nm, lane);
  wave_sync();
  double a[9];
#pragma unroll
  for (int k = 0; k < 9; k++) a[k] = lane < nm ? lds[9 * lane + k] : (k % 4 == 0 ? 1.0 : 0.0);
  wave_sync();
  GmfOut o;
  gmf_lane<LOG, 2, false>(a, 0.0, bt.b, o, g_gmf_recip.r, GmfSeriesLocal<LOG, 2, false>{bt.b});
  {
    double* my = mainr + lane * GhLds::MS;
#pragma unroll
    for (int k = 0; k < 9; k++) { my[k] = o.r[1][k]; my[9 + k] = o.r[2][k]; }
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
      for (int j = i; j < 3; j++)
#pragma unroll
        for (int k = j; k < 3; k++) my[18 + sym33s(i, j, k)] = o.D3[9 * i + 3 * j + k];
  }
  double ps = 0.0;
#pragma unroll
  for (int k = 0; k < 9; k++) ps = fma(o.F[k], o.F[k], ps);
  double d2r[7];                                   // D2 (6 distinct, by sym3) and sc1 of this lane's matrix
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = i; j < 3; j++) d2r[sym3(i, j)] = o.D2[3 * i + j];
  d2r[6] = o.sc1;
  const double sc2 = o.sc2;
  wave_sync();
  if (Lo && lane < nm) {                           // (log X from the lanes: 72 B per lane)
#pragma unroll
    for (int k = 0; k < 9; k++) Lo[(base + lane) * 9 + k] = o.F[k];
  }
  if (psi && lane < nm) psi[base + lane] = 0.5 * ps;
  const int gj = lane / 9, Lj = lane - 9 * gj;     // J: lane (gj, Lj = 3 c + d), gj < 6
  const int gh = lane / 27, Lg = lane - 27 * gh;   // H: lane (gh, Lg = 9 d + 3 e + f), gh < 2
  const int d = Lg / 9, e = (Lg / 3) % 3, f = Lg % 3;
  for (int p0 = 0; p0 < nm; p0 += GM) {            // groups of six matrices: J, K = J^T J, g, then K += L : H
    const int cg = nm - p0 < GM ? nm - p0 : GM;
    double dj[7], lj[9];                           // (from the matrix's lane; every lane takes part)
#pragma unroll
    for (int k = 0; k < 7; k++) dj[k] = __shfl(d2r[k], (p0 + (gj < GM ? gj : 0)) & (WAVE - 1));
#pragma unroll
    for (int k = 0; k < 9; k++) lj[k] = __shfl(o.F[k], (p0 + (gj < GM ? gj : 0)) & (WAVE - 1));
    if (gj < cg) {
      const double* s = mainr + (p0 + gj) * GhLds::MS;
      double out[3];
#pragma unroll
      for (int b = 0; b < 3; b++) {
        gmf_j_lane(s, s + 9, dj, dj[6], b, Lj / 3, Lj % 3, out);
#pragma unroll
        for (int k = 0; k < 3; k++) jt[gj * 81 + 27 * k + 9 * b + Lj] = out[k];
      }
    }
    wave_sync();                                   // (J tile staged)
    double* Kg = K ? K + (base + p0) * 81 : nullptr;   // (the group's K block)
    for (int t = lane; K && t < 81 * cg; t += WAVE) {   // K = J^T J of the group, stored (consecutive t)
      const int m = t / 81, ce = t - 81 * m, c = ce / 9, ee = ce - 9 * c;
      const double* jm = jt + 81 * m;
      double acc = 0.0;
#pragma unroll
      for (int q = 0; q < 9; q++) acc = fma(jm[9 * q + c], jm[9 * q + ee], acc);
      Kg[t] = acc;
    }
    if (g && lane < 9 * cg) {                      // g = J^T L, stored at once (consecutive lanes, consecutive g)
      const int m = lane / 9, c = lane - 9 * m;    // (m = gj: lj is this matrix's L)
      double acc = 0.0;
#pragma unroll
      for (int q = 0; q < 9; q++) acc = fma(lj[q], jt[81 * m + 9 * q + c], acc);
      g[(base + p0) * 9 + lane] = acc;
    }
    wave_sync();                                   // (J tile read: it takes the X1 tables; J^T J stored and visible)
    for (int h0 = 0; h0 < cg; h0 += 2) {
      const int cnt = cg - h0 < 2 ? cg - h0 : 2;
      const int mh = h0 + (gh < 2 ? gh : 0);       // the lane group's matrix in the group
      const double s2 = __shfl(sc2, (p0 + mh) & (WAVE - 1));
      double lm[9];                                // (L of the lane group's matrix, from its lane)
#pragma unroll
      for (int k = 0; k < 9; k++) lm[k] = __shfl(o.F[k], (p0 + mh) & (WAVE - 1));
      const double* sh = mainr + (p0 + mh) * GhLds::MS;
      double* ys = jt + (gh < 2 ? gh : 0) * 243;
      double y[9], out[27];
      if (gh < cnt) {
        gmf_h_y(sh, sh + 9, sh + 18, d, e, f, y);
#pragma unroll
        for (int k = 0; k < 9; k++) ys[9 * Lg + k] = y[k];
      }
      wave_sync();
      if (gh < cnt) {
        gmf_h_out(sh, sh + 9, s2, d, e, f, y, ys, out);
        if (K) {                                   // K += L : H on this lane's entries (c, Lg), in the output
#pragma unroll
          for (int c = 0; c < 3; c++) {
            double acc = Kg[mh * 81 + 27 * c + Lg];
#pragma unroll
            for (int q = 0; q < 9; q++) acc = fma(lm[q], out[3 * q + c], acc);
            Kg[mh * 81 + 27 * c + Lg] = acc;
          }
        }
      }
      wave_sync();                                 // (X1 tables read before the next pass writes them)
    }
  }
}

int gmf_hencky_launch(int64_t n, const double* A, double* Lo, double* psi, double* g, double* K, void* stream) {
  if (n < 0 || ((Lo || psi || g || K) && n > 0 && !A)) return -1;
  if (n == 0 || (!Lo && !psi && !g && !K)) return 0;
  GmfCoef bt;
  gmf_coefs(LOG, 0.0, bt);
  const hipStream_t st = (hipStream_t)stream;
  constexpr int64_t CHUNK = (int64_t)1 << 30;
  for (int64_t b = 0; b < n; b += CHUNK) {
    const int64_t c = n - b < CHUNK ? n - b : CHUNK;
    gmf_hencky_kernel<<<dim3((unsigned)((c + WAVE - 1) / WAVE)), dim3(WAVE), 0, st>>>(
        c, A + b * 9, Lo ? Lo + b * 9 : nullptr, psi ? psi + b : nullptr, g ? g + b * 9 : nullptr,
        K ? K + b * 81 : nullptr, bt);
    if (hipGetLastError() != hipSuccess) return -2;
  }
  return 0;
}

int smf_hencky_launch(bool mat, int64_t n, const double* A, double* Lo, double* psi, double* g, double* K,
                      void* stream) {
  if (n < 0 || ((Lo || psi || g || K) && n > 0 && !A)) return -1;
  if (n == 0 || (!Lo && !psi && !g && !K)) return 0;
  SmfCoef c2;
  smf_coefs(LOG, 0.0, c2);
  const hipStream_t st = (hipStream_t)stream;
  constexpr int64_t CHUNK = (int64_t)1 << 30;
  for (int64_t b = 0; b < n; b += CHUNK) {
    const int64_t c = n - b < CHUNK ? n - b : CHUNK;
    const int nin = mat ? 9 : 6, nk = nin * nin;
    const dim3 grid((unsigned)((c + WAVE - 1) / WAVE));
    double *lo = Lo ? Lo + b * nin : nullptr, *ps = psi ? psi + b : nullptr, *gg = g ? g + b * nin : nullptr,
           *kk = K ? K + b * nk : nullptr;
    if (mat) smf_hencky_kernel<true><<<grid, dim3(WAVE), 0, st>>>(c, A + b * nin, lo, ps, gg, kk, c2);
    else smf_hencky_kernel<false><<<grid, dim3(WAVE), 0, st>>>(c, A + b * nin, lo, ps, gg, kk, c2);
    if (hipGetLastError() != hipSuccess) return -2;
  }
  return 0;
}

}  // namespace mf

extern "C" int smf_hip_hencky_matrix(int64_t n, const double* A, double* L, double* psi, double* g, double* K,
                                     void* stream) {
  return mf::smf_hencky_launch(true, n, A, L, psi, g, K, stream);
}

extern "C" int smf_hip_hencky(int64_t n, const double* v, double* L, double* psi, double* g, double* K, void* stream) {
  return mf::smf_hencky_launch(false, n, v, L, psi, g, K, stream);
}

extern "C" int gmf_hip_hencky(int64_t n, const double* A, double* L, double* psi, double* g, double* K,
                              void* stream) {
  return mf::gmf_hencky_launch(n, A, L, psi, g, K, stream);
}

extern "C" int mf_hip_hencky(int64_t n, const double* L, const double* J, const double* H, double* psi, double* g,
                             double* K, void* stream) {
  return mf::hencky_launch<double>(n, L, J, H, psi, g, K, stream);
}

extern "C" int mf_hip_hencky_mixed(int64_t n, const float* L, const float* J, const float* H, float* psi, float* g,
                                   float* K, void* stream) {
  return mf::hencky_launch<float>(n, L, J, H, psi, g, K, stream);
}
