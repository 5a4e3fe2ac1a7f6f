// smf_kernels.hip -- symmetric 3x3 matrices on gfx950: one matrix per lane, one wavefront (64 matrices) per
// workgroup.  Inputs are read coalesced through LDS; every output block (F, J, each H row) is staged in a 64 x 21
// LDS tile of distinct values and written by lane groups so that consecutive lanes store consecutive doubles
// (the outputs are 96 B .. 5.8 KB per matrix and bound the kernel by HBM write bandwidth).
#include "smf_device.hpp"
#include "smf_launch.hpp"

namespace mf {

constexpr int WAVE = 64;
constexpr int TILE = 21;           // staged doubles per matrix (odd: conflict-free 64-bit LDS access by lane)

// write, for matrices j < nm of this wave, a block of G * KPL doubles at out + j * ostride (+ lane-group offset):
// element e = k * G + L of matrix j is tile[j * TILE + src[k]] * wt[k]; MPP = 64 / G matrices per pass
template <int G, int KPL>
MF_DEV void store_blocks(double* __restrict__ out, int64_t ostride, const double* tile, int nm, int lane,
                         const int (&src)[KPL], const double (&wt)[KPL]) {
  constexpr int MPP = WAVE / G;
  const int g = lane / G, L = lane - g * G;
  if (g >= MPP) return;
  for (int j = g; j < nm; j += MPP) {
    double* o = out + (int64_t)j * ostride + L;
    const double* s = tile + j * TILE;
#pragma unroll
    for (int k = 0; k < KPL; k++) __builtin_nontemporal_store(s[src[k]] * wt[k], o + k * G);
  }
}

template <int KIND, bool MAT, int ORDER>
__global__ void __launch_bounds__(WAVE) smf_kernel(int64_t n, const double* __restrict__ x, double* __restrict__ F,
                                                   double* __restrict__ J, double* __restrict__ H, double al,
                                                   SmfCoef c2) {
  __shared__ double tile[WAVE * TILE];
  constexpr int NIN = MAT ? 9 : 6, NF = MAT ? 9 : 6, NJ = NF * NF, NH = NJ * NF;
  const int lane = threadIdx.x;
  const int64_t base = (int64_t)blockIdx.x * WAVE;
  const int nm = (int)(n - base < WAVE ? n - base : WAVE);

  // ---- input, coalesced through LDS (lanes past the end take the identity)
  for (int t = lane; t < NIN * nm; t += WAVE) tile[t] = x[base * NIN + t];
  __syncthreads();
  double v[6];
  if (MAT) {
    double a[9];
#pragma unroll
    for (int k = 0; k < 9; k++) a[k] = lane < nm ? tile[9 * lane + k] : (k % 4 == 0 ? 1.0 : 0.0);
    matrix_to_voigt(a, v);
  } else {
#pragma unroll
    for (int k = 0; k < 6; k++) v[k] = lane < nm ? tile[6 * lane + k] : (k < 3 ? 1.0 : 0.0);
  }
  __syncthreads();

  SmfLane<KIND> s;
  double Fv[6];
  s.eigen(v, al, Fv);

  // ---- F
#pragma unroll
  for (int k = 0; k < 6; k++) tile[lane * TILE + k] = Fv[k];
  __syncthreads();
  for (int t = lane; t < NF * nm; t += WAVE) {
    const int j = t / NF, e = t - j * NF;
    __builtin_nontemporal_store(tile[j * TILE + (MAT ? voigt_of_m(e) : e)], F + base * NF + t);
  }
  if (ORDER < 1) return;
  __syncthreads();

  // ---- lane-group maps: matrix layout 81-blocks as 27 lanes x 3, Voigt 36-blocks as 12 lanes x 3
  constexpr int G = MAT ? 27 : 12;
  int src[3];
  double wJ[3], wH[3];
  {
    const int L = lane % G;
#pragma unroll
    for (int k = 0; k < 3; k++) {
      if (MAT) {                               // element 27k + L of an 81-block: (r, c) = (3k + L/9, L%9)
        const int r = 3 * k + L / 9, c = L % 9;
        src[k] = tri6(voigt_of_m(r), voigt_of_m(c));
        wJ[k] = 1.0;
        wH[k] = dv_dA(r) * dv_dA(c);
      } else {                                 // element 12k + L of a 36-block: (p, q) = (2k + L/6, L%6)
        const int p = 2 * k + L / 6, q = L % 6;
        src[k] = tri6(p, q);
        wJ[k] = q < 3 ? 1.0 : 2.0;             // J_voigt[p][q] = Js / w(q)
        wH[k] = 1.0;
      }
    }
  }

  // ---- J
  {
    double Js[21];
    s.jacobian(al, Js);
#pragma unroll
    for (int k = 0; k < 21; k++) tile[lane * TILE + k] = Js[k];
  }
  __syncthreads();
  store_blocks<G, 3>(J + base * NJ, NJ, tile, nm, lane, src, wJ);
  if (ORDER < 2) return;

  // ---- H, one Voigt row i at a time; matrix layout: row blocks m with voigt_of_m(m) == i (m and its transpose)
  s.prepare_hessian(al, c2.c);
  constexpr int ORDER_I[6] = {0, 3, 4, 1, 5, 2};
#pragma unroll
  for (int ii = 0; ii < 6; ii++) {
    const int i = ORDER_I[ii];
    double h[21];
    s.h_row(i, h);
    __syncthreads();                           // (previous tile fully read)
#pragma unroll
    for (int k = 0; k < 21; k++) tile[lane * TILE + k] = h[k];
    __syncthreads();
    if (MAT) {
#pragma unroll
      for (int m = 0; m < 9; m++)
        if (voigt_of_m(m) == i) store_blocks<G, 3>(H + base * NH + 81 * m, NH, tile, nm, lane, src, wH);
    } else {
      store_blocks<G, 3>(H + base * NH + 36 * i, NH, tile, nm, lane, src, wH);
    }
  }
}

template <int KIND, bool MAT>
static void launch_kind(hipStream_t st, unsigned blocks, int64_t n, const double* x, double* F, double* J, double* H,
                        double al, const SmfCoef& c2) {
  if (H) hipLaunchKernelGGL((smf_kernel<KIND, MAT, 2>), dim3(blocks), dim3(WAVE), 0, st, n, x, F, J, H, al, c2);
  else if (J) hipLaunchKernelGGL((smf_kernel<KIND, MAT, 1>), dim3(blocks), dim3(WAVE), 0, st, n, x, F, J, H, al, c2);
  else hipLaunchKernelGGL((smf_kernel<KIND, MAT, 0>), dim3(blocks), dim3(WAVE), 0, st, n, x, F, J, H, al, c2);
}

void smf_launch(hipStream_t st, int kind, bool mat, int64_t n, const double* x, double* F, double* J, double* H,
                double al, const SmfCoef& c2) {
  const unsigned blocks = (unsigned)((n + WAVE - 1) / WAVE);
#define MF_CASE(K)                                                             \
  if (mat) launch_kind<K, true>(st, blocks, n, x, F, J, H, al, c2);            \
  else launch_kind<K, false>(st, blocks, n, x, F, J, H, al, c2);
  switch (kind) {
    case EXP: MF_CASE(EXP) break;
    case LOG: MF_CASE(LOG) break;
    case SQRT: MF_CASE(SQRT) break;
    default: MF_CASE(POW) break;
  }
#undef MF_CASE
}

}  // namespace mf
