// gmf_kernels.hip -- the homogeneous general-matrix kernels on gfx950: workgroups of gmf_nw wavefronts, each
// wavefront one tile of 64 consecutive matrices (the tile body of gmf_tile.hpp over RunRows); the wavefronts of a
// workgroup of several share out their group series by length (GmfSeriesShared) and otherwise run independently.
#include "gmf_tile.hpp"
#include "smf_launch.hpp"

namespace mf {

// wavefronts (tiles) per workgroup: four that share out their group series (GmfSeriesShared) for the FP32 order-1
// kernels of the functions with long series (log, powers; exp's series are short), one elsewhere.  Measured on
// MI355X: sharing takes 13-16% off those kernels, which run four workgroups per CU, so that the other workgroups fill
// the SIMDs while a workgroup's waves wait for its longest series; the FP64 and order-2 kernels (two workgroups per
// CU by registers and LDS) lost 1-5% with it
template <int KIND, int ORDER, class OT>
constexpr int gmf_nw() { return KIND != EXP && ORDER == 1 && sizeof(OT) == 4 ? 4 : 1; }

// LDS of a workgroup: the tiles' regions (16-byte aligned), which the shared series also uses during phase 1
template <int KIND, int ORDER, class OT>
constexpr int gmf_wg_lds() {
  constexpr int nw = gmf_nw<KIND, ORDER, OT>();
  constexpr int tb = (GmfLds<ORDER, OT>::BYTES + 15) / 16 * 16, pool = nw > 1 ? GmfPool<ORDER, nw>::BYTES : 0;
  return nw * tb > pool ? nw * tb : pool;
}

// Register budgets: FP32 outputs at order 1 four wavefronts per SIMD (128 VGPRs; the LDS of 8.9 KB per tile, or
// 36.9 KB per workgroup of four, allows sixteen per CU); order 2 two (256 registers in all: left to itself the
// compiler may add a few accumulation registers to its 256 VGPRs, which leaves one wavefront per SIMD -- measured 45%
// slower)
template <int KIND, int ORDER, class OT, bool PA>
__global__ void __launch_bounds__((WAVE * gmf_nw<KIND, ORDER, OT>()))
    __attribute__((amdgpu_waves_per_eu((ORDER == 1 && sizeof(OT) == 4) ? 4 : (ORDER == 2 ? 2 : 1))))
    gmf_kernel(int64_t n, const double* __restrict__ A, OT* __restrict__ F, OT* __restrict__ J, OT* __restrict__ H,
               double al, const double* __restrict__ alv, GmfCoef bt) {
  constexpr int TB = (GmfLds<ORDER, OT>::BYTES + 15) / 16 * 16, NW = gmf_nw<KIND, ORDER, OT>();
  __shared__ __attribute__((aligned(16))) unsigned char lds[gmf_wg_lds<KIND, ORDER, OT>()];
  const int w = threadIdx.x / WAVE, lane = threadIdx.x - w * WAVE;
  const int64_t base = ((int64_t)blockIdx.x * NW + w) * WAVE, left = n - base;
  const int nm = left <= 0 ? 0 : (left < WAVE ? (int)left : WAVE);   // (a wavefront past the end: no matrices)
  gmf_tile<KIND, ORDER, OT, PA, RunRows, NW>(RunRows{base}, nm, A, F, J, H, al, alv, bt.b, lds + w * TB, lane,
                                                 lds, threadIdx.x);
}

template <int KIND, int ORDER, class OT>
static unsigned grid_for(int64_t n) {
  constexpr int64_t per = (int64_t)WAVE * gmf_nw<KIND, ORDER, OT>();
  return (unsigned)((n + per - 1) / per);
}

template <int KIND, class OT, bool PA>
static void glaunch_kind(hipStream_t st, int64_t n, const double* A, OT* F, OT* J, OT* H, double al,
                         const double* alv, const GmfCoef& bt) {
  if (H)
    hipLaunchKernelGGL((gmf_kernel<KIND, 2, OT, PA>), dim3(grid_for<KIND, 2, OT>(n)),
                       dim3(WAVE * gmf_nw<KIND, 2, OT>()), 0, st, n, A, F, J, H, al, alv, bt);
  else if (J)
    hipLaunchKernelGGL((gmf_kernel<KIND, 1, OT, PA>), dim3(grid_for<KIND, 1, OT>(n)),
                       dim3(WAVE * gmf_nw<KIND, 1, OT>()), 0, st, n, A, F, J, H, al, alv, bt);
  else
    hipLaunchKernelGGL((gmf_kernel<KIND, 0, OT, PA>), dim3(grid_for<KIND, 0, OT>(n)),
                       dim3(WAVE * gmf_nw<KIND, 0, OT>()), 0, st, n, A, F, J, H, al, alv, bt);
}

template <class OT>
static void glaunch_any(hipStream_t st, int kind, int64_t n, const double* A, OT* F, OT* J, OT* H, double al,
                        const GmfCoef& bt) {
  switch (kind) {
    case EXP: glaunch_kind<EXP, OT, false>(st, n, A, F, J, H, al, nullptr, bt); break;
    case LOG: glaunch_kind<LOG, OT, false>(st, n, A, F, J, H, al, nullptr, bt); break;
    default: glaunch_kind<POW, OT, false>(st, n, A, F, J, H, al, nullptr, bt); break;
  }
}

template <class OT>
static void glaunch_pow(hipStream_t st, int64_t n, const double* alv, const double* A, OT* F, OT* J, OT* H,
                        const GmfCoef& recips) {
  glaunch_kind<POW, OT, true>(st, n, A, F, J, H, 1.0, alv, recips);
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
