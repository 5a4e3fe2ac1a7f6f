// gmf_kernels.hip -- the homogeneous general-matrix kernels on gfx950: workgroups of NW wavefronts, each
// wavefront one tile of 64 consecutive matrices (the tile body of gmf_tile.hpp over RunRows); the wavefronts of a
// workgroup of several share out their group series by length (GmfSeriesShared) and otherwise run independently.
#include <atomic>
#include <mutex>

#include "gmf_tile.hpp"
#include "smf_launch.hpp"

namespace mf {

// wavefronts (tiles) per workgroup, NW = 1, 2 or 4, chosen per function slot (exp, log, pow, sqrt, per-matrix
// exponents), order and precision at run time (gmf_hip_set_tiles): the wavefronts of a workgroup of several share out
// their group series by length (GmfSeriesShared).  Default: four for the FP32 order-1 kernels and the order-0 kernels
// of the functions with long series (log, pow, sqrt with one exponent; exp's series are short), one elsewhere.
// Measured on MI355X over 2^20 matrices of the general mix: sharing takes 10-12% off the FP32 order-1 kernels, which
// run four workgroups per CU, so that the other workgroups fill the SIMDs while a workgroup's waves wait for its
// longest series, and 3-4% off order 0; the FP64 order-1 and the order-2 kernels (two or three wavefronts per SIMD by
// registers) lose 4-24% with two or four
int gmf_nw_default(int slot, int order, bool mixed) {
  const bool longs = slot == LOG || slot == POW || slot == SQRT;
  return longs && (order == 0 || (order == 1 && mixed)) ? 4 : (slot == 4 && order == 1 && mixed ? 4 : 1);
}

std::atomic<int> g_nw[5][3][2];
std::once_flag g_nw_once;

void nw_init() {
  std::call_once(g_nw_once, [] {
    for (int k = 0; k < 5; k++)
      for (int o = 0; o < 3; o++)
        for (int x = 0; x < 2; x++) g_nw[k][o][x].store(gmf_nw_default(k, o, x != 0));
  });
}

int gmf_tiles_get(int slot, int order, bool mixed) {
  nw_init();
  return g_nw[slot][order][mixed ? 1 : 0].load();
}
void gmf_tiles_set(int slot, int order, bool mixed, int nw) {
  nw_init();
  g_nw[slot][order][mixed ? 1 : 0].store(nw);
}

// LDS of a workgroup: the tiles' regions (16-byte aligned), which the shared series also uses during phase 1
template <int KIND, int ORDER, class OT, int NW>
constexpr int gmf_wg_lds() {
  constexpr int tb = (GmfLds<ORDER, OT>::BYTES + 15) / 16 * 16, pool = NW > 1 ? GmfPool<ORDER, NW>::BYTES : 0;
  return NW * tb > pool ? NW * tb : pool;
}

// Register budgets (per order and precision, whatever NW): FP32 outputs at order 1 four wavefronts per SIMD (128
// VGPRs; the LDS of 8.9 KB per tile, or 36.9 KB per workgroup of four, allows sixteen per CU; the compiler spills
// 35-88 VGPRs there, measured faster than three or two wavefronts per SIMD without spills: 0.70 against 0.80 and
// 1.00 ms for log over 2^20 general matrices); order 2 two (256 registers in all: left to itself the compiler may add
// a few accumulation registers to its 256 VGPRs, which leaves one wavefront per SIMD -- measured 45% slower)
template <int KIND, int ORDER, class OT, bool PA, int NW>
__global__ void __launch_bounds__((WAVE * NW))
    __attribute__((amdgpu_waves_per_eu((ORDER == 1 && sizeof(OT) == 4) ? 4 : (ORDER == 2 ? 2 : 1))))
    gmf_kernel(int64_t n, const double* __restrict__ A, OT* __restrict__ F, OT* __restrict__ J, OT* __restrict__ H,
               double al, const double* __restrict__ alv, GmfCoef bt) {
  constexpr int TB = (GmfLds<ORDER, OT>::BYTES + 15) / 16 * 16;
  __shared__ __attribute__((aligned(16))) unsigned char lds[gmf_wg_lds<KIND, ORDER, OT, NW>()];
  const int w = threadIdx.x / WAVE, lane = threadIdx.x - w * WAVE;
  const int64_t base = ((int64_t)blockIdx.x * NW + w) * WAVE, left = n - base;
  const int nm = left <= 0 ? 0 : (left < WAVE ? (int)left : WAVE);   // (a wavefront past the end: no matrices)
  gmf_tile<KIND, ORDER, OT, PA, RunRows, NW>(RunRows{base}, nm, A, F, J, H, al, alv, bt.b, lds + w * TB, lane, lds,
                                             threadIdx.x);
}

template <int KIND, int ORDER, class OT, bool PA, int NW>
static void glaunch_nw(hipStream_t st, int64_t n, const double* A, OT* F, OT* J, OT* H, double al, const double* alv,
                       const GmfCoef& bt) {
  const int64_t per = (int64_t)WAVE * NW;
  hipLaunchKernelGGL((gmf_kernel<KIND, ORDER, OT, PA, NW>), dim3((unsigned)((n + per - 1) / per)), dim3(WAVE * NW), 0,
                     st, n, A, F, J, H, al, alv, bt);
}

template <int KIND, int ORDER, class OT, bool PA>
static void glaunch_order(hipStream_t st, int nw, int64_t n, const double* A, OT* F, OT* J, OT* H, double al,
                          const double* alv, const GmfCoef& bt) {
  if (nw == 4) glaunch_nw<KIND, ORDER, OT, PA, 4>(st, n, A, F, J, H, al, alv, bt);
  else if (nw == 2) glaunch_nw<KIND, ORDER, OT, PA, 2>(st, n, A, F, J, H, al, alv, bt);
  else glaunch_nw<KIND, ORDER, OT, PA, 1>(st, n, A, F, J, H, al, alv, bt);
}

template <int KIND, class OT, bool PA>
static void glaunch_kind(hipStream_t st, int slot, int64_t n, const double* A, OT* F, OT* J, OT* H, double al,
                         const double* alv, const GmfCoef& bt) {
  const int order = H ? 2 : (J ? 1 : 0), nw = gmf_tiles_get(slot, order, sizeof(OT) == 4);
  if (H) glaunch_order<KIND, 2, OT, PA>(st, nw, n, A, F, J, H, al, alv, bt);
  else if (J) glaunch_order<KIND, 1, OT, PA>(st, nw, n, A, F, J, H, al, alv, bt);
  else glaunch_order<KIND, 0, OT, PA>(st, nw, n, A, F, J, H, al, alv, bt);
}

template <class OT>
static void glaunch_any(hipStream_t st, int slot, int kind, int64_t n, const double* A, OT* F, OT* J, OT* H,
                        double al, const GmfCoef& bt) {
  switch (kind) {
    case EXP: glaunch_kind<EXP, OT, false>(st, slot, n, A, F, J, H, al, nullptr, bt); break;
    case LOG: glaunch_kind<LOG, OT, false>(st, slot, n, A, F, J, H, al, nullptr, bt); break;
    default: glaunch_kind<POW, OT, false>(st, slot, n, A, F, J, H, al, nullptr, bt); break;
  }
}

void gmf_launch(hipStream_t st, int slot, int kind, int64_t n, const double* A, double* F, double* J, double* H,
                double al, const GmfCoef& bt) {
  glaunch_any<double>(st, slot, kind, n, A, F, J, H, al, bt);
}

void gmf_launch_mixed(hipStream_t st, int slot, int kind, int64_t n, const double* A, float* F, float* J, float* H,
                      double al, const GmfCoef& bt) {
  glaunch_any<float>(st, slot, kind, n, A, F, J, H, al, bt);
}

void gmf_launch_pow(hipStream_t st, int64_t n, const double* alv, const double* A, double* F, double* J, double* H,
                    const GmfCoef& recips) {
  glaunch_kind<POW, double, true>(st, 4, n, A, F, J, H, 1.0, alv, recips);
}

void gmf_launch_pow_mixed(hipStream_t st, int64_t n, const double* alv, const double* A, float* F, float* J, float* H,
                          const GmfCoef& recips) {
  glaunch_kind<POW, float, true>(st, 4, n, A, F, J, H, 1.0, alv, recips);
}

}  // namespace mf
