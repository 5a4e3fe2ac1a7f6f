// gmf_kernels.hip -- general real 3x3 matrices on gfx950, one wavefront (64 matrices) per workgroup.
//   phase 1, one matrix per lane: eigenvalues, divided differences, real basis, contracted tensors (gmf_device.hpp);
//            r_m (27 values) and D3 (27) staged in LDS per matrix, D2 (9) and F (9) in a second region
//   phase 2, lane groups: 27 lanes per matrix, two matrices per pass, assemble J (81: 3 per lane) and H (729: 27 per
//            lane) into an LDS tile in the API layout, which the wave then writes as one contiguous block (2 x 5.8 KB
//            of consecutive output per pass; whole cache lines per store instruction)
// OT = double: the FP64 form; OT = float: the mixed form (FP64 per-lane math, FP32 staged tensors, J / H assembly
// and outputs).
#include "gmf_device.hpp"
#include "smf_launch.hpp"

namespace mf {

constexpr int GWAVE = 64;

// LDS layout (elements of OT):
//   main  [64][MS]      r (27), then D3 (27) and sc2 at 54 (order 2)
//   x     [XS]          [0, 1458): the J (2 x 81) / H (2 x 729) output tile of a pass;
//                       before the H passes also [XD2, +640): D2 (9) and sc1 per matrix, [XF, +576): F per matrix
template <int ORDER>
struct GmfLds {
  static constexpr int MS = ORDER >= 2 ? 55 : 27;
  static constexpr int XD2 = 162, XF = ORDER == 0 ? 0 : XD2 + 640;
  static constexpr int XS = ORDER == 0 ? 576 : (ORDER == 1 ? XF + 576 : 1458);
  static constexpr int MAIN = ORDER == 0 ? 0 : 64 * MS;
};

template <int KIND, int ORDER, class OT>
__global__ void __launch_bounds__(GWAVE) gmf_kernel(int64_t n, const double* __restrict__ A, OT* __restrict__ F,
                                                    OT* __restrict__ J, OT* __restrict__ H, double al, GmfCoef bt) {
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

  // ---- phase 1
  {
    GmfOut o;
    gmf_lane<KIND, ORDER>(a, al, bt.b, o);
#pragma unroll
    for (int k = 0; k < 9; k++) xr[L::XF + lane * 9 + k] = (OT)o.F[k];
    if (ORDER >= 1) {
      OT* my = mainr + lane * L::MS;
#pragma unroll
      for (int k = 0; k < 27; k++) my[k] = (OT)o.r[k / 9][k % 9];
#pragma unroll
      for (int k = 0; k < 9; k++) xr[L::XD2 + lane * 10 + k] = (OT)o.D2[k];
      xr[L::XD2 + lane * 10 + 9] = (OT)o.sc1;
    }
    if (ORDER >= 2) {
      OT* my = mainr + lane * L::MS;
#pragma unroll
      for (int k = 0; k < 27; k++) my[27 + k] = (OT)o.D3[k];
      my[54] = (OT)o.sc2;
    }
  }
  __syncthreads();
  for (int t = lane; t < 9 * nm; t += GWAVE) F[base * 9 + t] = xr[L::XF + t];
  if (ORDER < 1) return;

  // ---- phase 2: lane (g, L) = matrix g of the pass, position L = 9 u + 3 v + w of a 27-lane group
  constexpr int G = 27, MPP = GWAVE / G;
  const int g = lane / G, Lg = lane - g * G;
  for (int p0 = 0; p0 < nm; p0 += MPP) {          // J: J_abcd for a = 0..2 at 27 a + (9b + 3c + d)
    const int cnt = nm - p0 < MPP ? nm - p0 : MPP;
    if (g < cnt) {
      const int j = p0 + g;
      OT out[3];
      gmf_j_lane(mainr + j * L::MS, xr + L::XD2 + j * 10, xr[L::XD2 + j * 10 + 9], Lg / 9, (Lg / 3) % 3, Lg % 3, out);
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
    if (g < cnt) {
      const OT* s = mainr + (p0 + g) * L::MS;
      OT out[27];
      gmf_h_lane(s, s + 27, s[54], Lg / 9, (Lg / 3) % 3, Lg % 3, out);
#pragma unroll
      for (int k = 0; k < 27; k++) xr[g * 729 + 27 * k + Lg] = out[k];
    }
    __syncthreads();
    for (int t = lane; t < 729 * cnt; t += GWAVE) H[(base + p0) * 729 + t] = xr[t];
    __syncthreads();
  }
}

template <int KIND, class OT>
static void glaunch_kind(hipStream_t st, unsigned blocks, int64_t n, const double* A, OT* F, OT* J, OT* H, double al,
                         const GmfCoef& bt) {
  if (H) hipLaunchKernelGGL((gmf_kernel<KIND, 2, OT>), dim3(blocks), dim3(GWAVE), 0, st, n, A, F, J, H, al, bt);
  else if (J) hipLaunchKernelGGL((gmf_kernel<KIND, 1, OT>), dim3(blocks), dim3(GWAVE), 0, st, n, A, F, J, H, al, bt);
  else hipLaunchKernelGGL((gmf_kernel<KIND, 0, OT>), dim3(blocks), dim3(GWAVE), 0, st, n, A, F, J, H, al, bt);
}

template <class OT>
static void glaunch_any(hipStream_t st, int kind, int64_t n, const double* A, OT* F, OT* J, OT* H, double al,
                        const GmfCoef& bt) {
  const unsigned blocks = (unsigned)((n + GWAVE - 1) / GWAVE);
  switch (kind) {
    case EXP: glaunch_kind<EXP, OT>(st, blocks, n, A, F, J, H, al, bt); break;
    case LOG: glaunch_kind<LOG, OT>(st, blocks, n, A, F, J, H, al, bt); break;
    default: glaunch_kind<POW, OT>(st, blocks, n, A, F, J, H, al, bt); break;
  }
}

void gmf_launch(hipStream_t st, int kind, int64_t n, const double* A, double* F, double* J, double* H, double al,
                const GmfCoef& bt) {
  glaunch_any<double>(st, kind, n, A, F, J, H, al, bt);
}

void gmf_launch_mixed(hipStream_t st, int kind, int64_t n, const double* A, float* F, float* J, float* H, double al,
                      const GmfCoef& bt) {
  glaunch_any<float>(st, kind, n, A, F, J, H, al, bt);
}

}  // namespace mf
