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
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/matfun_hip.h"

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

}  // namespace mf

extern "C" int mf_hip_hencky(int64_t n, const double* L, const double* J, const double* H, double* psi, double* g,
                             double* K, void* stream) {
  return mf::hencky_launch<double>(n, L, J, H, psi, g, K, stream);
}

extern "C" int mf_hip_hencky_mixed(int64_t n, const float* L, const float* J, const float* H, float* psi, float* g,
                                   float* K, void* stream) {
  return mf::hencky_launch<float>(n, L, J, H, psi, g, K, stream);
}
