// smf_kernels.hip -- the homogeneous symmetric kernels on gfx950: one wavefront (64 consecutive matrices) per
// workgroup, the tile body of smf_tile.hpp over RunRows (every output region of a workgroup one contiguous block).
#include "smf_launch.hpp"
#include "smf_tile.hpp"

namespace mf {

template <int KIND, bool MAT, int ORDER, class OT, bool PAIRS, bool PA>
__global__ void __launch_bounds__(WAVE) smf_kernel(int64_t n, const double* __restrict__ x, OT* __restrict__ F,
                                                   OT* __restrict__ J, OT* __restrict__ H, double al,
                                                   const double* __restrict__ alv, SmfCoef c2) {
  using L = SmfLds<MAT, ORDER, OT>;
  __shared__ __attribute__((aligned(16))) unsigned char lds[L::BYTES];
  __shared__ unsigned short lutJ[ORDER >= 1 ? smf_nj(MAT) : 1], lutH[ORDER >= 2 ? smf_nh(MAT) : 1];
  const int64_t base = (int64_t)blockIdx.x * WAVE;
  const int nm = (int)(n - base < WAVE ? n - base : WAVE);
  smf_luts<MAT, ORDER>(lutJ, lutH, threadIdx.x);
  smf_tile<KIND, MAT, ORDER, OT, PAIRS, PA>(RunRows{base}, nm, x, F, J, H, al, alv, c2.c, lds, lutJ, lutH,
                                            threadIdx.x);
}

template <int KIND, bool MAT, class OT, bool PAIRS, bool PA>
static void launch_order(hipStream_t st, unsigned blocks, int64_t n, const double* x, OT* F, OT* J, OT* H, double al,
                         const double* alv, const SmfCoef& c2) {
  if (H)
    hipLaunchKernelGGL((smf_kernel<KIND, MAT, 2, OT, PAIRS, PA>), dim3(blocks), dim3(WAVE), 0, st, n, x, F, J, H, al,
                       alv, c2);
  else if (J)
    hipLaunchKernelGGL((smf_kernel<KIND, MAT, 1, OT, PAIRS, PA>), dim3(blocks), dim3(WAVE), 0, st, n, x, F, J, H, al,
                       alv, c2);
  else
    hipLaunchKernelGGL((smf_kernel<KIND, MAT, 0, OT, false, PA>), dim3(blocks), dim3(WAVE), 0, st, n, x, F, J, H, al,
                       alv, c2);
}

// paired stores need J and H aligned to two elements (the per-wave regions then stay aligned: 64 and 10 matrices of
// 81 / 729 (36 / 216) elements are even counts)
template <int KIND, bool MAT, class OT, bool PA>
static void launch_kind(hipStream_t st, unsigned blocks, int64_t n, const double* x, OT* F, OT* J, OT* H, double al,
                        const double* alv, const SmfCoef& c2) {
  const uintptr_t al2 = 2 * sizeof(OT) - 1;
  if ((!J || ((uintptr_t)J & al2) == 0) && (!H || ((uintptr_t)H & al2) == 0))
    launch_order<KIND, MAT, OT, true, PA>(st, blocks, n, x, F, J, H, al, alv, c2);
  else
    launch_order<KIND, MAT, OT, false, PA>(st, blocks, n, x, F, J, H, al, alv, c2);
}

template <class OT>
static void launch_any(hipStream_t st, int kind, bool mat, int64_t n, const double* x, OT* F, OT* J, OT* H, double al,
                       const SmfCoef& c2) {
  const unsigned blocks = (unsigned)((n + WAVE - 1) / WAVE);
#define MF_CASE(K)                                                                        \
  if (mat) launch_kind<K, true, OT, false>(st, blocks, n, x, F, J, H, al, nullptr, c2);   \
  else launch_kind<K, false, OT, false>(st, blocks, n, x, F, J, H, al, nullptr, c2);
  switch (kind) {
    case EXP: MF_CASE(EXP) break;
    case LOG: MF_CASE(LOG) break;
    case SQRT: MF_CASE(SQRT) break;
    default: MF_CASE(POW) break;
  }
#undef MF_CASE
}

template <class OT>
static void launch_pow(hipStream_t st, bool mat, int64_t n, const double* alv, const double* x, OT* F, OT* J, OT* H) {
  const unsigned blocks = (unsigned)((n + WAVE - 1) / WAVE);
  const SmfCoef unused{};
  if (mat) launch_kind<POW, true, OT, true>(st, blocks, n, x, F, J, H, 1.0, alv, unused);
  else launch_kind<POW, false, OT, true>(st, blocks, n, x, F, J, H, 1.0, alv, unused);
}

void smf_launch(hipStream_t st, int kind, bool mat, int64_t n, const double* x, double* F, double* J, double* H,
                double al, const SmfCoef& c2) {
  launch_any<double>(st, kind, mat, n, x, F, J, H, al, c2);
}

void smf_launch_mixed(hipStream_t st, int kind, bool mat, int64_t n, const double* x, float* F, float* J, float* H,
                      double al, const SmfCoef& c2) {
  launch_any<float>(st, kind, mat, n, x, F, J, H, al, c2);
}

void smf_launch_pow(hipStream_t st, bool mat, int64_t n, const double* alv, const double* x, double* F, double* J,
                    double* H) {
  launch_pow<double>(st, mat, n, alv, x, F, J, H);
}

void smf_launch_pow_mixed(hipStream_t st, bool mat, int64_t n, const double* alv, const double* x, float* F, float* J,
                          float* H) {
  launch_pow<float>(st, mat, n, alv, x, F, J, H);
}

}  // namespace mf
