// gmf_kernels.hip -- general real 3x3 matrices on gfx950.  Phase 1, one matrix per lane (64 per wavefront =
// workgroup): eigenvalues, divided differences, real basis, contracted tensors (gmf_device.hpp).  Phase 2, lane
// groups: the results of the 64 matrices are staged in LDS and 27 lanes per matrix assemble J (81) and H (729) from
// r_m (27 values) and D2 (9) / D3 (27), each lane owning 3 consecutive output positions (J) or 27 (H) so that the
// stores are coalesced (216 B runs) -- H alone is 5.8 KB per matrix and sets the HBM roofline of the kernel.
#include "gmf_device.hpp"
#include "smf_launch.hpp"

namespace mf {

constexpr int GWAVE = 64;

template <int ORDER>
struct GmfTile {                      // per-matrix staging layout (stride odd: conflict-free 64-bit LDS rows)
  static constexpr int R = 0, D3 = 27, D2 = ORDER >= 2 ? 54 : 27, F = D2 + 9, SC1 = F + 9, SC2 = SC1 + 1;
  static constexpr int STRIDE = ORDER == 0 ? 9 : (ORDER == 1 ? 47 : 75);
  static constexpr int FOFF = ORDER == 0 ? 0 : F;
};

template <int KIND, int ORDER>
__global__ void __launch_bounds__(GWAVE) gmf_kernel(int64_t n, const double* __restrict__ A, double* __restrict__ F,
                                                    double* __restrict__ J, double* __restrict__ H, double al,
                                                    GmfCoef bt) {
  using T = GmfTile<ORDER>;
  __shared__ double tile[GWAVE * T::STRIDE];
  const int lane = threadIdx.x;
  const int64_t base = (int64_t)blockIdx.x * GWAVE;
  const int nm = (int)(n - base < GWAVE ? n - base : GWAVE);

  for (int t = lane; t < 9 * nm; t += GWAVE) tile[t] = A[base * 9 + t];
  __syncthreads();
  double a[9];
#pragma unroll
  for (int k = 0; k < 9; k++) a[k] = lane < nm ? tile[9 * lane + k] : (k % 4 == 0 ? 1.0 : 0.0);
  __syncthreads();

  {
    GmfOut o;
    gmf_lane<KIND, ORDER>(a, al, bt.b, o);
    double* my = tile + lane * T::STRIDE;
#pragma unroll
    for (int k = 0; k < 9; k++) my[T::FOFF + k] = o.F[k];
    if (ORDER >= 1) {
#pragma unroll
      for (int k = 0; k < 27; k++) my[T::R + k] = o.r[k / 9][k % 9];
#pragma unroll
      for (int k = 0; k < 9; k++) my[T::D2 + k] = o.D2[k];
      my[T::SC1] = o.sc1;
    }
    if (ORDER >= 2) {
#pragma unroll
      for (int k = 0; k < 27; k++) my[T::D3 + k] = o.D3[k];
      my[T::SC2] = o.sc2;
    }
  }
  __syncthreads();

  // ---- F: 9 per matrix, contiguous over the wave
  for (int t = lane; t < 9 * nm; t += GWAVE) {
    const int j = t / 9, k = t - 9 * j;
    __builtin_nontemporal_store(tile[j * T::STRIDE + T::FOFF + k], F + base * 9 + t);
  }
  if (ORDER < 1) return;

  // ---- J: lane (b, c, d) = 9b + 3c + d of a 27-lane group writes J_abcd for a = 0..2 (positions 27a + lane)
  constexpr int G = 27, MPP = GWAVE / G;
  const int g = lane / G, L = lane - g * G;
  if (g < MPP) {
    for (int j = g; j < nm; j += MPP) {
      const double* s = tile + j * T::STRIDE;
      double out[3];
      gmf_j_lane(s + T::R, s + T::D2, s[T::SC1], L / 9, (L / 3) % 3, L % 3, out);
      double* o = J + (base + j) * 81 + L;
#pragma unroll
      for (int aa = 0; aa < 3; aa++) __builtin_nontemporal_store(out[aa], o + 27 * aa);
    }
  }
  if (ORDER < 2) return;

  // ---- H: lane (d, e, f) = 9d + 3e + f writes H_abcdef for (a, b, c) = 0..26 (positions 27 (9a + 3b + c) + lane)
  if (g < MPP) {
    for (int j = g; j < nm; j += MPP) {
      const double* s = tile + j * T::STRIDE;
      double out[27];
      gmf_h_lane(s + T::R, s + T::D3, s[T::SC2], L / 9, (L / 3) % 3, L % 3, out);
      double* o = H + (base + j) * 729 + L;
#pragma unroll
      for (int k = 0; k < 27; k++) __builtin_nontemporal_store(out[k], o + 27 * k);
    }
  }
}

template <int KIND>
static void glaunch_kind(hipStream_t st, unsigned blocks, int64_t n, const double* A, double* F, double* J, double* H,
                         double al, const GmfCoef& bt) {
  if (H) hipLaunchKernelGGL((gmf_kernel<KIND, 2>), dim3(blocks), dim3(GWAVE), 0, st, n, A, F, J, H, al, bt);
  else if (J) hipLaunchKernelGGL((gmf_kernel<KIND, 1>), dim3(blocks), dim3(GWAVE), 0, st, n, A, F, J, H, al, bt);
  else hipLaunchKernelGGL((gmf_kernel<KIND, 0>), dim3(blocks), dim3(GWAVE), 0, st, n, A, F, J, H, al, bt);
}

void gmf_launch(hipStream_t st, int kind, int64_t n, const double* A, double* F, double* J, double* H, double al,
                const GmfCoef& bt) {
  const unsigned blocks = (unsigned)((n + GWAVE - 1) / GWAVE);
  switch (kind) {
    case EXP: glaunch_kind<EXP>(st, blocks, n, A, F, J, H, al, bt); break;
    case LOG: glaunch_kind<LOG>(st, blocks, n, A, F, J, H, al, bt); break;
    default: glaunch_kind<POW>(st, blocks, n, A, F, J, H, al, bt); break;
  }
}

}  // namespace mf
