// gmf_kernels.hip -- general real 3x3 matrices on gfx950, one wavefront (64 matrices) per workgroup.
//   phase 1, one matrix per lane: eigenvalues, divided differences, real basis, contracted tensors (gmf_device.hpp);
//            r_m (27 values) and D3 (27) staged in LDS per matrix, D2 (9) and F (9) in a second region
//   phase 2, lane groups: 27 lanes per matrix, two matrices per pass, assemble J (81: 3 per lane) and H (729: 27 per
//            lane) into an LDS tile in the API layout, which the wave then writes as one contiguous block (2 x 5.8 KB
//            of consecutive output per pass; whole cache lines per store instruction)
// OT = double: the FP64 form; OT = float: the mixed form (FP64 per-lane math, FP32 staged tensors, J / H assembly
// and outputs).  PA: powers with one exponent per matrix (alv[i], read by its lane; bt then holds 1 / j; a matrix
// whose exponent is not finite gets NaN results).
#include "gmf_device.hpp"
#include "smf_launch.hpp"

namespace mf {

constexpr int GWAVE = 64;

// LDS layout (elements of OT):
//   main [64][MS]   r1 (9), r2 (9); order 2: the 10 distinct D3 and sc2 (MS = 29; order 1: 19)
//   x    [XS]       phase 1: F of the 64 matrices [0, 576), then D2 (6 distinct) and sc1 per matrix [162, 610);
//                   J passes: the 2 x 81 output tile [0, 162); H passes: a half-block tile 2 x 378 (or 2 x 351)
template <int ORDER>
struct GmfLds {
  static constexpr int MS = ORDER >= 2 ? 29 : 19;
  static constexpr int XAUX = 162;
  static constexpr int XS = ORDER == 0 ? 576 : (ORDER == 1 ? 610 : 756);
  static constexpr int MAIN = ORDER == 0 ? 0 : 64 * MS;
};

template <int KIND, int ORDER, class OT, bool PA>
__global__ void __launch_bounds__(GWAVE) gmf_kernel(int64_t n, const double* __restrict__ A, OT* __restrict__ F,
                                                    OT* __restrict__ J, OT* __restrict__ H, double al,
                                                    const double* __restrict__ alv, GmfCoef bt) {
  using L = GmfLds<ORDER>;
  constexpr int TOT = L::MAIN + L::XS;
  constexpr int BYTES = (TOT * (int)sizeof(OT) > GWAVE * 9 * 8) ? TOT * (int)sizeof(OT) : GWAVE * 9 * 8;
  __shared__ double raw[(BYTES + 7) / 8];
  double* xin = raw;                                 // (the FP64 input block, before phase 1)
  OT* mainr = reinterpret_cast<OT*>(raw);
  OT* xr = mainr + L::MAIN;
  const int lane = threadIdx.x;
  const int64_t base = (int64_t)blockIdx.x * GWAVE;
  const int nm = (int)(n - base < GWAVE ? n - base : GWAVE);

  for (int t = lane; t < 9 * nm; t += GWAVE) xin[t] = A[base * 9 + t];
  __syncthreads();
  double a[9];
#pragma unroll
  for (int k = 0; k < 9; k++) a[k] = lane < nm ? xin[9 * lane + k] : (k % 4 == 0 ? 1.0 : 0.0);
  __syncthreads();

  bool alok = true;                                  // (PA: a non-finite exponent computes with 1 and stores NaN)
  if (PA) {
    al = lane < nm ? alv[base + lane] : 1.0;
    alok = fabs(al) <= 1.7976931348623157e308;
    al = alok ? al : 1.0;
  }

  // ---- phase 1
  GmfOut o;
  gmf_lane<KIND, ORDER, PA>(a, al, bt.b, o);
  if (PA) {                                          // (J and H carry the factors sc1, sc2)
    const double nanf = alok ? 1.0 : dnan();
#pragma unroll
    for (int k = 0; k < 9; k++) o.F[k] *= nanf;
    o.sc1 *= nanf;
    o.sc2 *= nanf;
  }
  if (ORDER >= 1) {
    OT* my = mainr + lane * L::MS;
#pragma unroll
    for (int k = 0; k < 9; k++) { my[k] = (OT)o.r[1][k]; my[9 + k] = (OT)o.r[2][k]; }
    if (ORDER >= 2) {
#pragma unroll
      for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = i; j < 3; j++)
#pragma unroll
          for (int k = j; k < 3; k++) my[18 + sym33s(i, j, k)] = (OT)o.D3[9 * i + 3 * j + k];
      my[28] = (OT)o.sc2;
    }
  }
#pragma unroll
  for (int k = 0; k < 9; k++) xr[lane * 9 + k] = (OT)o.F[k];
  __syncthreads();
  for (int t = lane; t < 9 * nm; t += GWAVE) F[base * 9 + t] = xr[t];
  if (ORDER < 1) return;
  __syncthreads();
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = i; j < 3; j++) xr[L::XAUX + lane * 7 + sym3(i, j)] = (OT)o.D2[3 * i + j];
  xr[L::XAUX + lane * 7 + 6] = (OT)o.sc1;
  __syncthreads();

  // ---- phase 2: lane (g, Lg) = matrix g of the pass, position Lg = 9 u + 3 v + w of a 27-lane group
  constexpr int G = 27, MPP = GWAVE / G;
  const int g = lane / G, Lg = lane - g * G;
  for (int p0 = 0; p0 < nm; p0 += MPP) {          // J: J_abcd for a = 0..2 at 27 a + (9b + 3c + d)
    const int cnt = nm - p0 < MPP ? nm - p0 : MPP;
    if (g < cnt) {
      const int j = p0 + g;
      const OT* s = mainr + j * L::MS;
      const OT* aux = xr + L::XAUX + j * 7;
      OT out[3];
      gmf_j_lane(s, s + 9, aux, aux[6], Lg / 9, (Lg / 3) % 3, Lg % 3, out);
#pragma unroll
      for (int k = 0; k < 3; k++) xr[g * 81 + 27 * k + Lg] = out[k];
    }
    __syncthreads();
    for (int t = lane; t < 81 * cnt; t += GWAVE) J[(base + p0) * 81 + t] = xr[t];
    __syncthreads();
  }
  if (ORDER < 2) return;
  for (int p0 = 0; p0 < nm; p0 += MPP) {          // H: H_abcdef for (a,b,c) = 0..26 at 27 (9a + 3b + c) + (9d + 3e + f)
    const int cnt = nm - p0 < MPP ? nm - p0 : MPP;
    OT out[27];
    if (g < cnt) {
      const OT* s = mainr + (p0 + g) * L::MS;
      gmf_h_lane(s, s + 9, s + 18, s[28], Lg / 9, (Lg / 3) % 3, Lg % 3, out);
    }
#pragma unroll
    for (int half = 0; half < 2; half++) {        // rows k < 14 (378 elements), then k >= 14 (351)
      constexpr int K0[2] = {0, 14}, SZ[2] = {378, 351};
      if (g < cnt) {
#pragma unroll
        for (int k = 0; k < 14; k++)
          if (K0[half] + k < 27) xr[g * SZ[half] + 27 * k + Lg] = out[K0[half] + k];
      }
      __syncthreads();
      for (int t = lane; t < SZ[half] * cnt; t += GWAVE) {
        const int gg = t >= SZ[half] ? 1 : 0, e = t - gg * SZ[half];
        H[(base + p0 + gg) * 729 + 27 * K0[half] + e] = xr[t];
      }
      __syncthreads();
    }
  }
}

template <int KIND, class OT, bool PA>
static void glaunch_kind(hipStream_t st, unsigned blocks, int64_t n, const double* A, OT* F, OT* J, OT* H, double al,
                         const double* alv, const GmfCoef& bt) {
  if (H)
    hipLaunchKernelGGL((gmf_kernel<KIND, 2, OT, PA>), dim3(blocks), dim3(GWAVE), 0, st, n, A, F, J, H, al, alv, bt);
  else if (J)
    hipLaunchKernelGGL((gmf_kernel<KIND, 1, OT, PA>), dim3(blocks), dim3(GWAVE), 0, st, n, A, F, J, H, al, alv, bt);
  else
    hipLaunchKernelGGL((gmf_kernel<KIND, 0, OT, PA>), dim3(blocks), dim3(GWAVE), 0, st, n, A, F, J, H, al, alv, bt);
}

template <class OT>
static void glaunch_any(hipStream_t st, int kind, int64_t n, const double* A, OT* F, OT* J, OT* H, double al,
                        const GmfCoef& bt) {
  const unsigned blocks = (unsigned)((n + GWAVE - 1) / GWAVE);
  switch (kind) {
    case EXP: glaunch_kind<EXP, OT, false>(st, blocks, n, A, F, J, H, al, nullptr, bt); break;
    case LOG: glaunch_kind<LOG, OT, false>(st, blocks, n, A, F, J, H, al, nullptr, bt); break;
    default: glaunch_kind<POW, OT, false>(st, blocks, n, A, F, J, H, al, nullptr, bt); break;
  }
}

template <class OT>
static void glaunch_pow(hipStream_t st, int64_t n, const double* alv, const double* A, OT* F, OT* J, OT* H,
                        const GmfCoef& recips) {
  const unsigned blocks = (unsigned)((n + GWAVE - 1) / GWAVE);
  glaunch_kind<POW, OT, true>(st, blocks, n, A, F, J, H, 1.0, alv, recips);
}

void gmf_launch(hipStream_t st, int kind, int64_t n, const double* A, double* F, double* J, double* H, double al,
                const GmfCoef& bt) {
  glaunch_any<double>(st, kind, n, A, F, J, H, al, bt);
}

void gmf_launch_mixed(hipStream_t st, int kind, int64_t n, const double* A, float* F, float* J, float* H, double al,
                      const GmfCoef& bt) {
  glaunch_any<float>(st, kind, n, A, F, J, H, al, bt);
}

void gmf_launch_pow(hipStream_t st, int64_t n, const double* alv, const double* A, double* F, double* J, double* H,
                    const GmfCoef& recips) {
  glaunch_pow<double>(st, n, alv, A, F, J, H, recips);
}

void gmf_launch_pow_mixed(hipStream_t st, int64_t n, const double* alv, const double* A, float* F, float* J, float* H,
                          const GmfCoef& recips) {
  glaunch_pow<float>(st, n, alv, A, F, J, H, recips);
}

}  // namespace mf
