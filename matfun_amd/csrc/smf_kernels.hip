// smf_kernels.hip -- the homogeneous symmetric kernels on gfx950: one wavefront (64 consecutive matrices) per
// workgroup, the tile body of smf_tile.hpp over RunRows (every output region of a workgroup one contiguous block).
#include <atomic>

#include "smf_launch.hpp"
#include "smf_tile.hpp"

namespace mf {

// H store tile of the order-2 kernels per (function, layout, precision): [kind 0..3, or 4 = pow with per-matrix
// exponents][matrix layout][mixed] -> TILE_SEP10 (10 matrices per phase-B pass, V / D2 in their own LDS block) or
// 10, 8, 5 (matrices per pass, V / D2 in the phase-A region).  Less LDS per wavefront means more wavefronts per CU;
// no result depends on the choice.  Defaults measured on MI355X; smf_hip_set_tile (the tuner) changes them.
static std::atomic<int> g_tile[5][2][2] = {
    {{{TILE_SEP10}, {8}}, {{TILE_SEP10}, {8}}}, {{{TILE_SEP10}, {8}}, {{TILE_SEP10}, {8}}},
    {{{TILE_SEP10}, {8}}, {{TILE_SEP10}, {8}}}, {{{TILE_SEP10}, {8}}, {{TILE_SEP10}, {8}}},
    {{{TILE_SEP10}, {8}}, {{TILE_SEP10}, {8}}}};

int smf_tile_get(int slot, bool mat, bool mixed) { return g_tile[slot][mat][mixed].load(std::memory_order_relaxed); }
void smf_tile_set(int slot, bool mat, bool mixed, int hpp) { g_tile[slot][mat][mixed].store(hpp, std::memory_order_relaxed); }

template <int KIND, bool MAT, int ORDER, class OT, int VEC, bool PA, int HPP = HPP_DEFAULT, bool SEP = true,
          bool PACK = false>
__global__ void __launch_bounds__(WAVE) smf_kernel(int64_t n, const double* __restrict__ x, OT* __restrict__ F,
                                                   OT* __restrict__ J, OT* __restrict__ H, double al,
                                                   const double* __restrict__ alv, SmfCoef c2) {
  using L = SmfLds<MAT, ORDER, OT, HPP, SEP>;
  constexpr bool LUT = !PACK;
  __shared__ __attribute__((aligned(16))) unsigned char lds[L::BYTES];
  __shared__ short lutJ[LUT && ORDER >= 1 ? smf_nj(MAT) + LUT_PAD : 1];
  __shared__ short lutH[LUT && ORDER >= 2 ? smf_nh(MAT) + LUT_PAD : 1];
  const int64_t base = (int64_t)blockIdx.x * WAVE;
  const int nm = (int)(n - base < WAVE ? n - base : WAVE);
  if (LUT) smf_luts<MAT, ORDER, OT>(lutJ, lutH, threadIdx.x);
  smf_tile<KIND, MAT, ORDER, OT, VEC, PA, RunRows, HPP, SEP, PACK>(RunRows{base}, nm, x, F, J, H, al, alv, c2.c, lds,
                                                                   lutJ, lutH, threadIdx.x);
}

#define MF_LAUNCH(ORD, V, HP, SP)                                                                                    \
  hipLaunchKernelGGL((smf_kernel<KIND, MAT, ORD, OT, V, PA, HP, SP>), dim3(blocks), dim3(WAVE), 0, st, n, x, F, J, \
                     H, al, alv, c2)

// order 2: the tile (TILE_SEP10, or 10 / 8 / 5 matrices per pass sharing the phase-A region); VEC elements per store
template <int KIND, bool MAT, class OT, int VEC, bool PA>
static void launch_h(hipStream_t st, unsigned blocks, int64_t n, const double* x, OT* F, OT* J, OT* H, double al,
                     const double* alv, const SmfCoef& c2, int tile) {
  if (VEC == 4 && MAT) tile = 8;                 // (FP32 quadruples in the matrix layout: the caller chose tile 8)
  if (tile == 10) MF_LAUNCH(2, VEC, 10, false);
  else if (tile == 8) MF_LAUNCH(2, VEC, 8, false);
  else if (tile == 5) MF_LAUNCH(2, VEC, 5, false);
  else MF_LAUNCH(2, VEC, 10, true);
}

// Vector stores need J and H aligned to the vector, and every block a wave writes to start aligned: the J block of 64
// matrices always does; an H pass of HPP matrices x NH elements keeps 16-byte alignment for FP32 quadruples only if
// HPP NH is a multiple of 4 (matrix layout: tile 8; Voigt: any tile), FP64 / FP32 pairs with any tile.
template <int KIND, bool MAT, class OT, bool PA>
static void launch_kind(hipStream_t st, unsigned blocks, int64_t n, const double* x, OT* F, OT* J, OT* H, double al,
                        const double* alv, const SmfCoef& c2) {
  auto aligned = [&](uintptr_t a) { return (!J || ((uintptr_t)J & (a - 1)) == 0) && (!H || ((uintptr_t)H & (a - 1)) == 0); };
  const bool a2 = aligned(2 * sizeof(OT)), a16 = aligned(16);
  const int tile = a2 ? smf_tile_get(PA ? 4 : KIND, MAT, sizeof(OT) == 4) : TILE_SEP10;
  constexpr bool F32 = sizeof(OT) == 4;
  if (!J) {
    MF_LAUNCH(0, 1, HPP_DEFAULT, true);
  } else if (!H) {
    if (F32 && a16) MF_LAUNCH(1, (F32 ? 4 : 2), HPP_DEFAULT, true);
    else if (a2) MF_LAUNCH(1, 2, HPP_DEFAULT, true);
    else MF_LAUNCH(1, 1, HPP_DEFAULT, true);
  } else if (F32 && a16 && (!MAT || tile == 8)) {
    launch_h<KIND, MAT, OT, (F32 ? 4 : 2), PA>(st, blocks, n, x, F, J, H, al, alv, c2, tile);
  } else if (a2) {
    launch_h<KIND, MAT, OT, 2, PA>(st, blocks, n, x, F, J, H, al, alv, c2, tile);
  } else {
    MF_LAUNCH(2, 1, HPP_DEFAULT, true);
  }
}
#undef MF_LAUNCH

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

// packed outputs (Voigt input; F 6, J 21, H 126 per matrix): the default tile (10 matrices per H pass, whose
// 1260-element pass and the 1344-element J block of a wave keep 16-byte alignment for FP32 quadruples too)
template <int KIND, class OT>
static void launch_packed_kind(hipStream_t st, unsigned blocks, int64_t n, const double* x, OT* F, OT* J, OT* H,
                               double al, const SmfCoef& c2) {
  auto aligned = [&](uintptr_t a) { return (!J || ((uintptr_t)J & (a - 1)) == 0) && (!H || ((uintptr_t)H & (a - 1)) == 0); };
  constexpr int V2 = sizeof(OT) == 4 ? 4 : 2;
  const double* alv = nullptr;
#define MF_PK(ORD, V)                                                                                               \
  hipLaunchKernelGGL((smf_kernel<KIND, false, ORD, OT, V, false, HPP_DEFAULT, true, true>), dim3(blocks), dim3(WAVE), \
                     0, st, n, x, F, J, H, al, alv, c2)
  if (!J) MF_PK(0, 1);
  else if (!H) {
    if (aligned(16)) MF_PK(1, V2); else if (aligned(2 * sizeof(OT))) MF_PK(1, 2); else MF_PK(1, 1);
  } else {
    if (aligned(16)) MF_PK(2, V2); else if (aligned(2 * sizeof(OT))) MF_PK(2, 2); else MF_PK(2, 1);
  }
#undef MF_PK
}

template <class OT>
static void launch_packed(hipStream_t st, int kind, int64_t n, const double* x, OT* F, OT* J, OT* H, double al,
                          const SmfCoef& c2) {
  const unsigned blocks = (unsigned)((n + WAVE - 1) / WAVE);
  switch (kind) {
    case EXP: launch_packed_kind<EXP, OT>(st, blocks, n, x, F, J, H, al, c2); break;
    case LOG: launch_packed_kind<LOG, OT>(st, blocks, n, x, F, J, H, al, c2); break;
    case SQRT: launch_packed_kind<SQRT, OT>(st, blocks, n, x, F, J, H, al, c2); break;
    default: launch_packed_kind<POW, OT>(st, blocks, n, x, F, J, H, al, c2); break;
  }
}

void smf_launch_packed(hipStream_t st, int kind, int64_t n, const double* x, double* F, double* J, double* H,
                       double al, const SmfCoef& c2) {
  launch_packed<double>(st, kind, n, x, F, J, H, al, c2);
}

void smf_launch_packed_mixed(hipStream_t st, int kind, int64_t n, const double* x, float* F, float* J, float* H,
                             double al, const SmfCoef& c2) {
  launch_packed<float>(st, kind, n, x, F, J, H, al, c2);
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
