// batch_kernels.hpp -- the kernels of heterogeneous batches (mf_hip_batch_eval): one persistent kernel over
// class-homogeneous tiles of any class, and one kernel per class.
//
// A class is (family, function, order): 24 of them.  Each tile holds up to 64 matrices of one class (one per lane),
// listed by their rows (input / F, J, H) in device tables; the tile bodies are those of the homogeneous kernels
// (smf_tile.hpp, gmf_tile.hpp) over ListRows, so a matrix gets bitwise the result of the homogeneous call.
// Classes: smf exp / log / sqrt with their fixed series tables; smf pow with per-matrix exponents (PA); gmf exp /
// log; gmf sqrt as pow 1/2 with its table (as gmf_hip_eval); gmf pow with per-matrix exponents.
// LOOP: each workgroup takes tiles t0 + b, t0 + b + gridDim.x, ... (persistent); else one tile per workgroup.
// The kernel files that include this header (batch_fused.hip, batch_smf.hip, batch_gmf.hip) are compiled with
// MachineLICM off, and the tile bodies get the lane and the table pointers through opaque() each tile: with them
// the persistent loop keeps no per-tile state live across tiles, so its register use stays that of one tile.
#pragma once
#include "gmf_tile.hpp"
#include "philox.hpp"
#include "smf_tile.hpp"

namespace mf {

constexpr int NCLS = 24;                        // family (2) x function slot (4) x order (3)
MF_DEV constexpr int cls_of(int fam, int slot, int order) { return fam * 12 + slot * 3 + order; }
MF_DEV constexpr int cls_fam(int c) { return c / 12; }
MF_DEV constexpr int cls_slot(int c) { return (c % 12) / 3; }
MF_DEV constexpr int cls_order(int c) { return c % 3; }
constexpr int SLOT_POW = 3;                     // function slots: 0 exp, 1 log, 2 sqrt, 3 pow (per-matrix exponents)

struct BatchTile {
  int32_t cls;
  int32_t nm;
  int64_t off;          // its rows: tables [off, off + nm)
};

struct BatchTables {     // the same for every plan
  double smf_c2[3][SMF_NTERMS];   // exp, log, sqrt
  GmfCoef gmf[3];                 // exp, log, pow 1/2
  GmfCoef recips;                 // 1 / j (per-matrix exponents)
};

// the Newton perturbation applied by the evaluation itself: each matrix's inputs multiplied by perturb_factors
// (philox.hpp) of its row before it is evaluated, and written back over A (mf_hip_perturb's result, bitwise)
struct BatchPerturb {
  double sigma;
  uint64_t seed;
  uint32_t iteration;
};

struct BatchArgs {
  const BatchTile* tiles;
  int64_t t0, t1;                 // tile range of this launch
  const int64_t* rm;
  const int64_t* rj;
  const int64_t* rh;
  const double* A;
  const double* alv;              // per-matrix exponents (the caller's, or the plan's group exponents)
  int perturb;                    // 1: the inputs perturbed first (pt), written back through Aw (= A)
  double* Aw;
  BatchPerturb pt;
  // a launch over one class's tiles (full but the last): they list the class's cn matrices from table offset coff
  int64_t coff, cn;
};

// tile t of a launch over one class's tiles, from its place (no descriptor load)
template <int CLS>
MF_DEV BatchTile class_tile(const BatchArgs& a, int64_t t) {
  const int64_t k = (t - a.t0) * WAVE, left = a.cn - k;
  return BatchTile{CLS, (int32_t)(left < WAVE ? left : WAVE), a.coff + k};
}

// the tile bodies' input transform (pre_inputs in tile_rows.hpp) for BatchArgs::perturb
struct PerturbPre {
  double* x;
  BatchPerturb pt;
  bool on, sym;
  MF_DEV bool active() const { return on; }
  MF_DEV void apply(double* a, int64_t m) const {
    double f[9];
    perturb_factors(pt.seed, pt.iteration, (uint64_t)m, sym, pt.sigma, f);
#pragma unroll
    for (int k = 0; k < 9; k++) a[k] *= f[k];
  }
};

// ---- host side: each kernel file holds its own copy of the tables in constant memory (code objects do not share
// symbols), written by its upload function; launches
int batch_upload_fused(const BatchTables& t);       // 0, or -2 (device error)
int batch_upload_smf(const BatchTables& t);
int batch_upload_gmf(const BatchTables& t);
void batch_launch_fused(hipStream_t st, unsigned grid, bool loop, const BatchArgs& a, double* F, double* J, double* H);
void batch_launch_fused(hipStream_t st, unsigned grid, bool loop, const BatchArgs& a, float* F, float* J, float* H);
void batch_launch_smf(int cls, hipStream_t st, unsigned grid, bool loop, const BatchArgs& a, double* F, double* J,
                      double* H);
void batch_launch_smf(int cls, hipStream_t st, unsigned grid, bool loop, const BatchArgs& a, float* F, float* J,
                      float* H);
void batch_launch_gmf(int cls, hipStream_t st, unsigned grid, bool loop, const BatchArgs& a, double* F, double* J,
                      double* H);
void batch_launch_gmf(int cls, hipStream_t st, unsigned grid, bool loop, const BatchArgs& a, float* F, float* J,
                      float* H);
// a gmf class of order >= 1 (not exp) finished from staged data (gmf_staged_list over the class's matrices in tile
// order: matrix k of the class at k = tile offset - off0 + place in the tile), one workgroup per tile
bool batch_gmf_stageable(int cls);
void batch_launch_gmf_staged(int cls, hipStream_t st, unsigned grid, const BatchArgs& a, int64_t off0,
                             const GmfStaged& sg, double* F, double* J, double* H);
void batch_launch_gmf_staged(int cls, hipStream_t st, unsigned grid, const BatchArgs& a, int64_t off0,
                             const GmfStaged& sg, float* F, float* J, float* H);

#ifdef MF_BATCH_KERNELS
// constant memory: uniform reads of it are scalar loads, so the tables take no vector registers
static __constant__ BatchTables g_tab;

// one tile of class CLS (lane and the table pointers opaque per tile)
template <int CLS, class OT, class R = ListRows>
MF_DEV void run_tile(const BatchArgs& a, const BatchTile& tl, const R& rows, OT* F, OT* J, OT* H,
                     unsigned char* lds, const short* lutJ, const short* lutH) {
  constexpr int FAM = cls_fam(CLS), SLOT = cls_slot(CLS), ORDER = cls_order(CLS);
  const int lane = opaque((int)threadIdx.x);
  const PerturbPre pre{a.Aw, a.pt, a.perturb != 0, FAM == 0};
  if constexpr (FAM == 0) {
    constexpr int KIND = SLOT == 0 ? EXP : (SLOT == 1 ? LOG : (SLOT == 2 ? SQRT : POW));
    const double* c2 = opaque(g_tab.smf_c2[SLOT < 3 ? SLOT : 0]);
    smf_tile<KIND, true, ORDER, OT, 1, SLOT == SLOT_POW, R, HPP_DEFAULT, false, false, PerturbPre>(
        rows, tl.nm, a.A, F, J, H, SLOT == 2 ? 0.5 : 0.0, a.alv, c2, lds, lutJ, lutH, lane, pre);
  } else {
    constexpr int KIND = SLOT == 0 ? EXP : (SLOT == 1 ? LOG : POW);
    const double* bt = opaque(SLOT == SLOT_POW ? g_tab.recips.b : g_tab.gmf[SLOT].b);
    gmf_tile<KIND, ORDER, OT, SLOT == SLOT_POW, R, 1, false, PerturbPre>(
        rows, tl.nm, a.A, F, J, H, SLOT == 2 ? 0.5 : 0.0, a.alv, bt, lds, lane, nullptr, 0,
        GmfStaged{nullptr, nullptr, nullptr, nullptr, 0, -1}, pre);
  }
}

MF_DEV void stage_rows(const BatchArgs& a, const BatchTile& tl, int64_t* rows, int lane) {
  if (lane < tl.nm) {
    rows[lane] = a.rm[tl.off + lane];
    rows[WAVE + lane] = a.rj[tl.off + lane];
    rows[2 * WAVE + lane] = a.rh[tl.off + lane];
  }
  __syncthreads();
}

// ListRows32 of a tile: its input rows into LDS (32-bit: a plan holds fewer than 2^31 matrices), its J and H rows
// in registers, the run masks of the three regions by ballot
// (ORDER: the class's order; the J / H rows of a class that writes no J / H are not read)
template <int ORDER>
MF_DEV ListRows32 stage_rows32(const BatchArgs& a, const BatchTile& tl, int32_t* rows, int lane) {
  const bool in = lane < tl.nm;
  int64_t m = 0, jr = 0, hr = 0;
  if (in) {
    m = a.rm[tl.off + lane];
    if (ORDER >= 1) jr = a.rj[tl.off + lane];
    if (ORDER >= 2) hr = a.rh[tl.off + lane];
    rows[lane] = (int32_t)m;
  }
  const int64_t pm = __shfl_up(m, 1), pj = __shfl_up(jr, 1), ph = __shfl_up(hr, 1);
  const bool on = in && lane > 0;
  const ListRows32 R{rows, jr, hr, 0,
                     {__ballot(on && m == pm + 1), __ballot(on && jr == pj + 1), __ballot(on && hr == ph + 1)}};
  __syncthreads();
  return R;
}

template <int C, class OT>
MF_DEV void dispatch_from(int cls, const BatchArgs& a, const BatchTile& tl, const ListRows& rows, OT* F, OT* J, OT* H,
                          unsigned char* lds, const short* lutJ, const short* lutH) {
  if constexpr (C < NCLS) {
    if (cls == C) run_tile<C, OT>(a, tl, rows, F, J, H, lds, lutJ, lutH);
    else dispatch_from<C + 1, OT>(cls, a, tl, rows, F, J, H, lds, lutJ, lutH);
  }
}

template <class OT>
constexpr int batch_lds_bytes() {
  constexpr int a = SmfLds<true, 2, OT>::BYTES, b = GmfLds<2, OT>::BYTES;
  return a > b ? a : b;
}

// all classes (registers and LDS of the largest, gmf order 2: two waves per SIMD)
template <class OT, bool LOOP>
__global__ void __launch_bounds__(WAVE) __attribute__((amdgpu_waves_per_eu(2)))
batch_kernel(BatchArgs a, OT* __restrict__ F, OT* __restrict__ J, OT* __restrict__ H) {
  __shared__ __attribute__((aligned(16))) unsigned char lds[batch_lds_bytes<OT>()];
  __shared__ short lutJ[81 + LUT_PAD], lutH[729 + LUT_PAD];
  __shared__ int64_t rows[3 * WAVE];
  const int lane = threadIdx.x;
  smf_luts<true, 2, OT>(lutJ, lutH, lane);
  const ListRows R{rows, rows + WAVE, rows + 2 * WAVE};
  for (int64_t t = a.t0 + blockIdx.x; t < a.t1; t += gridDim.x) {
    const BatchTile tl = a.tiles[t];
    stage_rows(a, tl, rows, lane);
    dispatch_from<0, OT>(tl.cls, a, tl, R, F, J, H, lds, lutJ, lutH);
    if (!LOOP) break;
  }
}

// one class: the tiles t0 .. t1 - 1 are all of class CLS
template <int CLS, class OT, bool LOOP>
__global__ void __launch_bounds__(WAVE) __attribute__((amdgpu_waves_per_eu(2)))
batch_class_kernel(BatchArgs a, OT* __restrict__ F, OT* __restrict__ J, OT* __restrict__ H) {
  constexpr int FAM = cls_fam(CLS), ORDER = cls_order(CLS);
  constexpr int B = FAM == 0 ? SmfLds<true, ORDER, OT>::BYTES : GmfLds<ORDER, OT>::BYTES;
  __shared__ __attribute__((aligned(16))) unsigned char lds[B];
  __shared__ short lutJ[FAM == 0 && ORDER >= 1 ? 81 + LUT_PAD : 1], lutH[FAM == 0 && ORDER >= 2 ? 729 + LUT_PAD : 1];
  // the rows staged as ListRows32: 256 B of LDS (input rows), J and H rows in registers
  __shared__ int32_t rows32[WAVE];
  const int lane = threadIdx.x;
  if (FAM == 0) smf_luts<true, ORDER, OT>(lutJ, lutH, lane);
  for (int64_t t = a.t0 + blockIdx.x; t < a.t1; t += gridDim.x) {
    const BatchTile tl = class_tile<CLS>(a, t);
    run_tile<CLS, OT>(a, tl, stage_rows32<ORDER>(a, tl, rows32, lane), F, J, H, lds, lutJ, lutH);
    if (!LOOP) break;
  }
}

// one tile of a gmf class from staged data (the finishing pass of the staged evaluation over the tile's listed rows)
template <int CLS, class OT>
__global__ void __launch_bounds__(WAVE) __attribute__((amdgpu_waves_per_eu(2)))
batch_staged_kernel(BatchArgs a, int64_t off0, GmfStaged sg, OT* __restrict__ F, OT* __restrict__ J,
                    OT* __restrict__ H) {
  constexpr int SLOT = cls_slot(CLS), ORDER = cls_order(CLS);
  constexpr int KIND = SLOT == 0 ? EXP : (SLOT == 1 ? LOG : POW);
  static_assert(cls_fam(CLS) == 1 && ORDER >= 1, "gmf classes of order >= 1");
  __shared__ __attribute__((aligned(16))) unsigned char lds[GmfLds<ORDER, OT>::BYTES];
  __shared__ int32_t rows32[WAVE];
  const int lane = threadIdx.x;
  const int64_t t = a.t0 + blockIdx.x;
  if (t >= a.t1) return;
  const BatchTile tl = class_tile<CLS>(a, t);
  const double* bt = SLOT == SLOT_POW ? g_tab.recips.b : g_tab.gmf[SLOT].b;
  const GmfStaged s{sg.nre, sg.nim, sg.sre, sg.sim, 0, tl.off - off0};
  const ListRows32 rows = stage_rows32<ORDER>(a, tl, rows32, lane);
  gmf_tile<KIND, ORDER, OT, SLOT == SLOT_POW, ListRows32, 1, true>(rows, tl.nm, a.A, F, J, H, SLOT == 2 ? 0.5 : 0.0,
                                                                   a.alv, bt, lds, lane, nullptr, 0, s);
}

template <int C, int C1, class OT>
void launch_class_from(int cls, hipStream_t st, unsigned grid, bool loop, const BatchArgs& a, OT* F, OT* J, OT* H) {
  if constexpr (C < C1) {
    if (cls != C) {
      launch_class_from<C + 1, C1, OT>(cls, st, grid, loop, a, F, J, H);
    } else if (loop) {
      hipLaunchKernelGGL((batch_class_kernel<C, OT, true>), dim3(grid), dim3(WAVE), 0, st, a, F, J, H);
    } else {
      hipLaunchKernelGGL((batch_class_kernel<C, OT, false>), dim3(grid), dim3(WAVE), 0, st, a, F, J, H);
    }
  }
}

inline int upload_tables(const BatchTables& t) {
  return hipMemcpyToSymbol(HIP_SYMBOL(g_tab), &t, sizeof(t), 0, hipMemcpyHostToDevice) == hipSuccess ? 0 : -2;
}
#endif

}  // namespace mf
