// gmf_kernels.hip -- the homogeneous general-matrix kernels on gfx950: one wavefront (64 consecutive matrices) per
// workgroup, the tile body of gmf_tile.hpp over RunRows.
#include "gmf_tile.hpp"
#include "smf_launch.hpp"

namespace mf {

// FP32 outputs at order 1: a register budget of four wavefronts per SIMD (128 VGPRs; the LDS of 8.9 KB per
// wavefront allows 18 per CU)
template <int KIND, int ORDER, class OT, bool PA>
__global__ void __launch_bounds__(WAVE) __attribute__((amdgpu_waves_per_eu((ORDER == 1 && sizeof(OT) == 4) ? 4 : 1)))
gmf_kernel(int64_t n, const double* __restrict__ A, OT* __restrict__ F,
                                                   OT* __restrict__ J, OT* __restrict__ H, double al,
                                                   const double* __restrict__ alv, GmfCoef bt) {
  __shared__ __attribute__((aligned(16))) unsigned char lds[GmfLds<ORDER, OT>::BYTES];
  const int64_t base = (int64_t)blockIdx.x * WAVE;
  const int nm = (int)(n - base < WAVE ? n - base : WAVE);
  gmf_tile<KIND, ORDER, OT, PA>(RunRows{base}, nm, A, F, J, H, al, alv, bt.b, lds, threadIdx.x);
}

template <int KIND, class OT, bool PA>
static void glaunch_kind(hipStream_t st, unsigned blocks, int64_t n, const double* A, OT* F, OT* J, OT* H, double al,
                         const double* alv, const GmfCoef& bt) {
  if (H)
    hipLaunchKernelGGL((gmf_kernel<KIND, 2, OT, PA>), dim3(blocks), dim3(WAVE), 0, st, n, A, F, J, H, al, alv, bt);
  else if (J)
    hipLaunchKernelGGL((gmf_kernel<KIND, 1, OT, PA>), dim3(blocks), dim3(WAVE), 0, st, n, A, F, J, H, al, alv, bt);
  else
    hipLaunchKernelGGL((gmf_kernel<KIND, 0, OT, PA>), dim3(blocks), dim3(WAVE), 0, st, n, A, F, J, H, al, alv, bt);
}

template <class OT>
static void glaunch_any(hipStream_t st, int kind, int64_t n, const double* A, OT* F, OT* J, OT* H, double al,
                        const GmfCoef& bt) {
  const unsigned blocks = (unsigned)((n + WAVE - 1) / WAVE);
  switch (kind) {
    case EXP: glaunch_kind<EXP, OT, false>(st, blocks, n, A, F, J, H, al, nullptr, bt); break;
    case LOG: glaunch_kind<LOG, OT, false>(st, blocks, n, A, F, J, H, al, nullptr, bt); break;
    default: glaunch_kind<POW, OT, false>(st, blocks, n, A, F, J, H, al, nullptr, bt); break;
  }
}

template <class OT>
static void glaunch_pow(hipStream_t st, int64_t n, const double* alv, const double* A, OT* F, OT* J, OT* H,
                        const GmfCoef& recips) {
  const unsigned blocks = (unsigned)((n + WAVE - 1) / WAVE);
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
