// gmf_sorted.hip -- the general-matrix evaluation with the matrices ordered by series length.
//
// A wavefront runs its group-series loop to the longest series among its 64 lanes, and the lengths spread widely
// (general mix, log order 2: mean 51 terms, median 43, while a wavefront's longest averages 165), so most lanes idle
// through most of the loop.  Here the batch is evaluated in three steps on the stream:
//   1. gmf_klen_kernel   one lane per matrix runs the per-matrix math up to the series length K (gmf_lane with
//                        GmfSeriesLength) and stores the key 200 - K (one byte; longest first);
//   2. a counting sort of the keys (key_hist_kernel: a histogram per chunk of the batch; key_scan_kernel: every
//                        chunk's offset per key; key_scatter_kernel: each matrix's position) into a permutation;
//   3. gmf_perm_kernel   the tile body of the homogeneous kernel (gmf_tile) over tiles of 64 consecutive entries of
//                        the permutation (PermRows): every matrix is computed by the same code as in gmf_hip_eval,
//                        so its results are those of gmf_hip_eval bitwise; only the grouping of matrices into
//                        wavefronts changes, and with it how long each wavefront's series loop runs.
// Workspace (caller-owned device memory, gmf_sort_workspace): keys (1 B per matrix), the permutation (4 B per matrix),
// two tables of KSORT_MAXCH x KBINS counts.  A batch larger than the workspace holds runs in consecutive parts.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <atomic>
#include <mutex>

#include "gmf_sorted.hpp"
#include "gmf_tile.hpp"
#include "philox.hpp"
#include "smf_launch.hpp"

namespace mf {

namespace {

constexpr int KLEN_THREADS = 256, SORT_THREADS = 1024;

size_t align256(size_t b) { return (b + 255) / 256 * 256; }

// the keys 200 - K of matrices [0, n) (one lane per matrix)
template <int KIND, int ORDER, bool PA>
__global__ void __launch_bounds__(KLEN_THREADS)
    gmf_klen_kernel(int64_t n, const double* __restrict__ A, double al, const double* __restrict__ alv, GmfCoef bt,
                    uint8_t* __restrict__ key) {
  const int64_t i = (int64_t)blockIdx.x * KLEN_THREADS + threadIdx.x;
  if (i >= n) return;
  double a[9];
#pragma unroll
  for (int k = 0; k < 9; k++) a[k] = A[9 * i + k];
  if (PA) {                                        // (as gmf_tile: a non-finite exponent computes with 1)
    al = alv[i];
    al = fabs(al) <= 1.7976931348623157e308 ? al : 1.0;
  }
  GmfOut o;
  gmf_lane<KIND, ORDER, PA>(a, al, bt.b, o, g_gmf_recip.r, GmfSeriesLength{});
  key[i] = (uint8_t)(GMF_KMAX - o.K);
}

// the series inputs and the key of matrices k = 0 .. n-1 (one lane per matrix; the first pass of the staged
// evaluation): matrix k is row rows[k] of A (and of alv) when rows is given, else row k; outputs at k.  Aw (= A,
// hence A without __restrict__): the inputs first multiplied by the matrix's Newton perturbation factors and written
// back
template <int KIND, int ORDER, bool PA>
__global__ void __launch_bounds__(KLEN_THREADS)
    gmf_prep_kernel(int64_t n, const int64_t* __restrict__ rows, const double* A, double al,
                    const double* __restrict__ alv, GmfCoef bt, uint8_t* __restrict__ key, GmfSeriesIn* __restrict__ sin,
                    GmfNodesRe* __restrict__ nre, double* __restrict__ nim, double* Aw, double sigma, uint64_t seed,
                    uint32_t iteration) {
  const int64_t k = (int64_t)blockIdx.x * KLEN_THREADS + threadIdx.x;
  if (k >= n) return;
  const int64_t i = rows ? rows[k] : k;
  double a[9];
#pragma unroll
  for (int k = 0; k < 9; k++) a[k] = A[9 * i + k];
  if (Aw) {
    double f[9];
    perturb_factors(seed, iteration, (uint64_t)i, false, sigma, f);
#pragma unroll
    for (int k = 0; k < 9; k++) {
      a[k] *= f[k];
      Aw[9 * i + k] = a[k];
    }
  }
  if (PA) {
    al = alv[i];
    al = fabs(al) <= 1.7976931348623157e308 ? al : 1.0;
  }
  GmfOut o;
  gmf_lane<KIND, ORDER, PA>(a, al, bt.b, o, g_gmf_recip.r, GmfSeriesLength{sin + k, nre + k, nim + 6 * k});
  key[k] = (uint8_t)(o.cls == 0 ? GMF_KMAX - o.K : GMF_KMAX + 1 + (o.cls - 1) * KQ + (GMF_KMAX - o.K) / 8);
}

// the group series of the matrices in sorted order: lane t of the batch takes matrix perm[t] (its inputs from
// sin, its ND sums at the matrix's own row: real parts to sre, imaginary parts to sim for series class 2 only).  Each
// matrix runs its class's loop (the real loop without or with the pole factors, or the complex loop), so its sums do
// not depend on which matrices share its wavefront (the order within a key is that of the sort's atomics); the keys
// order the series by class, so a wavefront runs more than one loop only where the classes meet
template <int KIND, int ORDER, bool PA>
__global__ void __launch_bounds__(KLEN_THREADS)
    gmf_series_kernel(int64_t n, const int32_t* __restrict__ perm, const GmfSeriesIn* __restrict__ sin, GmfCoef bt,
                      double* __restrict__ sre, double* __restrict__ sim) {
  constexpr int ND = ORDER == 0 ? 3 : (ORDER == 1 ? 9 : 19);
  const int64_t t = (int64_t)blockIdx.x * KLEN_THREADS + threadIdx.x;
  if (t >= n) return;
  const int64_t i = perm[t];
  const GmfSeriesIn s = sin[i];
  const int cls = gmf_series_class(s);
  cpx acc[ND];
  if (cls == 0) gmf_series_real<KIND, ORDER, PA, false>(s, bt.b, acc);
  else if (cls == 1) gmf_series_real<KIND, ORDER, PA, true>(s, bt.b, acc);
  else gmf_series<KIND, ORDER, PA>(s, bt.b, acc);
#pragma unroll
  for (int c = 0; c < ND; c++) sre[i * ND + c] = acc[c].re;
  if (cls == 2) {
#pragma unroll
    for (int c = 0; c < ND; c++) sim[i * ND + c] = acc[c].im;
  }
}

// the tile body of the homogeneous kernel over consecutive matrices, each finished from its staged nodes and sums
template <int KIND, int ORDER, class OT, bool PA>
__global__ void __launch_bounds__(WAVE) __attribute__((amdgpu_waves_per_eu(ORDER == 2 ? 2 : 1)))
    gmf_finish_kernel(int64_t n, const double* __restrict__ A, OT* __restrict__ F, OT* __restrict__ J,
                      OT* __restrict__ H, double al, const double* __restrict__ alv, GmfCoef bt, GmfStaged staged) {
  __shared__ __attribute__((aligned(16))) unsigned char lds[GmfLds<ORDER, OT>::BYTES];
  const int lane = threadIdx.x;
  const int64_t base = (int64_t)blockIdx.x * WAVE, left = n - base;
  const int nm = left < WAVE ? (int)left : WAVE;
  gmf_tile<KIND, ORDER, OT, PA, RunRows, 1, true>(RunRows{base}, nm, A, F, J, H, al, alv, bt.b, lds, lane, nullptr, 0,
                                                  staged);
}

// hist[c][b]: how many of the keys in chunk c (matrices [c chunk, (c + 1) chunk)) equal b
__global__ void __launch_bounds__(SORT_THREADS)
    key_hist_kernel(int64_t n, int64_t chunk, const uint8_t* __restrict__ key, uint32_t* __restrict__ hist) {
  __shared__ uint32_t h[KBINS];
  for (int b = threadIdx.x; b < KBINS; b += SORT_THREADS) h[b] = 0;
  __syncthreads();
  const int64_t lo = (int64_t)blockIdx.x * chunk, hi = lo + chunk < n ? lo + chunk : n;
  for (int64_t i = lo + threadIdx.x; i < hi; i += SORT_THREADS) atomicAdd(&h[key[i]], 1u);
  __syncthreads();
  for (int b = threadIdx.x; b < KBINS; b += SORT_THREADS) hist[(int64_t)blockIdx.x * KBINS + b] = h[b];
}

// off[c][b] = (matrices with a key below b) + (matrices with key b in the chunks before c): where chunk c's matrices
// of key b start in the permutation.  One workgroup; the nch x KBINS table is scanned in LDS.
__global__ void __launch_bounds__(SORT_THREADS)
    key_scan_kernel(int nch, const uint32_t* __restrict__ hist, uint32_t* __restrict__ off) {
  __shared__ uint32_t t[KSORT_MAXCH * KBINS];
  __shared__ uint32_t base[KBINS];
  const int cells = nch * KBINS;
  for (int k = threadIdx.x; k < cells; k += SORT_THREADS) t[k] = hist[k];
  __syncthreads();
  if (threadIdx.x < KBINS) {                       // per key: running sum over the chunks (exclusive)
    const int b = threadIdx.x;
    uint32_t s = 0;
    for (int c = 0; c < nch; c++) {
      const uint32_t v = t[c * KBINS + b];
      t[c * KBINS + b] = s;
      s += v;
    }
    base[b] = s;                                   // (the key's total)
  }
  __syncthreads();
  if (threadIdx.x == 0) {                          // exclusive scan of the totals over the keys
    uint32_t s = 0;
    for (int b = 0; b < KBINS; b++) {
      const uint32_t v = base[b];
      base[b] = s;
      s += v;
    }
  }
  __syncthreads();
  for (int k = threadIdx.x; k < cells; k += SORT_THREADS) off[k] = t[k] + base[k % KBINS];
}

// perm[position] = matrix, positions of chunk c's matrices of key b from off[c][b] on (their order within a chunk
// and key is that of the LDS atomics: any order gives the same results)
__global__ void __launch_bounds__(SORT_THREADS)
    key_scatter_kernel(int64_t n, int64_t chunk, const uint8_t* __restrict__ key, const uint32_t* __restrict__ off,
                       int32_t* __restrict__ perm) {
  __shared__ uint32_t cur[KBINS];
  for (int b = threadIdx.x; b < KBINS; b += SORT_THREADS) cur[b] = off[(int64_t)blockIdx.x * KBINS + b];
  __syncthreads();
  const int64_t lo = (int64_t)blockIdx.x * chunk, hi = lo + chunk < n ? lo + chunk : n;
  for (int64_t i = lo + threadIdx.x; i < hi; i += SORT_THREADS) {
    const uint32_t pos = atomicAdd(&cur[key[i]], 1u);
    perm[pos] = (int32_t)i;
  }
}

// the homogeneous kernel's tile body over perm[64 t .. 64 t + 63]; register budgets as gmf_kernel's one-tile
// workgroups (order 2: two wavefronts per SIMD).  Workgroup b takes tile t = b step mod T
// (T tiles, step coprime to T, about T / SORT_MIX): the workgroups resident at one time hold tiles from every part
// of the length order -- long series (compute) next to short ones (stores) -- instead of the longest ones together
// and then the shortest (measured on MI355X against the plain order: 1-6% faster at order 2; SORT_MIX 16 or 64
// gained less)
constexpr int64_t SORT_MIX = 4;
template <int KIND, int ORDER, class OT, bool PA>
__global__ void __launch_bounds__(WAVE) __attribute__((amdgpu_waves_per_eu(ORDER == 2 ? 2 : 1)))
    gmf_perm_kernel(int64_t n, int64_t step, const int32_t* __restrict__ perm, const double* __restrict__ A,
                    OT* __restrict__ F, OT* __restrict__ J, OT* __restrict__ H, double al,
                    const double* __restrict__ alv, GmfCoef bt) {
  __shared__ __attribute__((aligned(16))) unsigned char lds[GmfLds<ORDER, OT>::BYTES];
  __shared__ int32_t rows[WAVE];
  const int lane = threadIdx.x;
  const int64_t T = (n + WAVE - 1) / WAVE, t = (int64_t)blockIdx.x * step % T;
  const int64_t base = t * WAVE, left = n - base;
  const int nm = left < WAVE ? (int)left : WAVE;
  const bool in = lane < nm;
  const int32_t r = in ? perm[base + lane] : 0;
  rows[lane] = r;
  const int32_t rp = __shfl_up(r, 1);
  const uint64_t cont = __ballot(in && lane > 0 && r == rp + 1);
  __syncthreads();
  gmf_tile<KIND, ORDER, OT, PA, PermRows>(PermRows{rows, {cont, cont, cont}}, nm, A, F, J, H, al, alv, bt.b, lds,
                                          lane);
}

template <int KIND, bool PA>
void klen_launch(hipStream_t st, int order, int64_t n, const double* A, double al, const double* alv,
                 const GmfCoef& bt, uint8_t* key) {
  const unsigned g = (unsigned)((n + KLEN_THREADS - 1) / KLEN_THREADS);
  if (order == 2)
    hipLaunchKernelGGL((gmf_klen_kernel<KIND, 2, PA>), dim3(g), dim3(KLEN_THREADS), 0, st, n, A, al, alv, bt, key);
  else if (order == 1)
    hipLaunchKernelGGL((gmf_klen_kernel<KIND, 1, PA>), dim3(g), dim3(KLEN_THREADS), 0, st, n, A, al, alv, bt, key);
  else
    hipLaunchKernelGGL((gmf_klen_kernel<KIND, 0, PA>), dim3(g), dim3(KLEN_THREADS), 0, st, n, A, al, alv, bt, key);
}

int64_t gcd64(int64_t a, int64_t b) { return b ? gcd64(b, a % b) : a; }

template <int KIND, class OT, bool PA>
void perm_launch(hipStream_t st, int64_t n, const int32_t* perm, const double* A, OT* F, OT* J, OT* H, double al,
                 const double* alv, const GmfCoef& bt) {
  const int64_t T = (n + WAVE - 1) / WAVE;
  int64_t step = T / SORT_MIX > 1 ? T / SORT_MIX : 1;
  while (gcd64(step, T) != 1) step++;                     // (a bijection of the tiles)
  const unsigned g = (unsigned)T;
  if (H)
    hipLaunchKernelGGL((gmf_perm_kernel<KIND, 2, OT, PA>), dim3(g), dim3(WAVE), 0, st, n, step, perm, A, F, J, H, al,
                       alv, bt);
  else if (J)
    hipLaunchKernelGGL((gmf_perm_kernel<KIND, 1, OT, PA>), dim3(g), dim3(WAVE), 0, st, n, step, perm, A, F, J, H, al,
                       alv, bt);
  else
    hipLaunchKernelGGL((gmf_perm_kernel<KIND, 0, OT, PA>), dim3(g), dim3(WAVE), 0, st, n, step, perm, A, F, J, H, al,
                       alv, bt);
}

// the workspace of a part of m matrices
struct SortSpace {
  uint8_t* key;
  int32_t* perm;
  uint32_t *hist, *off;
};
SortSpace carve(void* work, int64_t m) {
  unsigned char* p = static_cast<unsigned char*>(work);
  SortSpace s;
  s.key = p;
  p += align256((size_t)m);
  s.perm = reinterpret_cast<int32_t*>(p);
  p += align256(4 * (size_t)m);
  s.hist = reinterpret_cast<uint32_t*>(p);
  p += align256(4 * (size_t)KSORT_MAXCH * KBINS);
  s.off = reinterpret_cast<uint32_t*>(p);
  return s;
}

// one part of m <= capacity matrices: keys, sort, evaluation
template <int KIND, class OT, bool PA>
void sorted_part(hipStream_t st, int order, int64_t m, const double* A, OT* F, OT* J, OT* H, double al,
                 const double* alv, const GmfCoef& bt, void* work) {
  const SortSpace s = carve(work, m);
  // chunks of the histogram: at most KSORT_MAXCH, each a multiple of the workgroup size
  int64_t chunk = (m + KSORT_MAXCH - 1) / KSORT_MAXCH;
  chunk = (chunk + SORT_THREADS - 1) / SORT_THREADS * SORT_THREADS;
  const int nch = (int)((m + chunk - 1) / chunk);
  klen_launch<KIND, PA>(st, order, m, A, al, alv, bt, s.key);
  hipLaunchKernelGGL(key_hist_kernel, dim3(nch), dim3(SORT_THREADS), 0, st, m, chunk, s.key, s.hist);
  hipLaunchKernelGGL(key_scan_kernel, dim3(1), dim3(SORT_THREADS), 0, st, nch, s.hist, s.off);
  hipLaunchKernelGGL(key_scatter_kernel, dim3(nch), dim3(SORT_THREADS), 0, st, m, chunk, s.key, s.off, s.perm);
  perm_launch<KIND, OT, PA>(st, m, s.perm, A, F, J, H, al, alv, bt);
}

// the staged evaluation of m matrices at order `order`: series inputs and keys (natural order), the sort, the series
// in sorted order, then the tile body in natural order with the sums (contiguous stores).  Workspace: staged_bytes.
size_t staged_bytes(int64_t m, int order) {
  const int nd = order == 0 ? 3 : (order == 1 ? 9 : 19);
  return align256((size_t)m) + align256(4 * (size_t)m) + 2 * align256(4 * (size_t)KSORT_MAXCH * KBINS) +
         align256(sizeof(GmfSeriesIn) * (size_t)m) + align256(sizeof(GmfNodesRe) * (size_t)m) +
         align256(6 * sizeof(double) * (size_t)m) + 2 * align256(sizeof(double) * nd * (size_t)m);
}

// the passes before the finishing pass (nodes and series inputs, the sort, the series) of m matrices, matrix k being
// row rows[k] of A (or row k): the staged data, indexed by k, in the workspace (staged_bytes(m, ORDER))
template <int KIND, int ORDER, bool PA>
GmfStaged staged_front(hipStream_t st, int64_t m, const int64_t* rows, const double* A, double al, const double* alv,
                       const GmfCoef& bt, void* work, double* Aw = nullptr, double sigma = 0.0, uint64_t seed = 0,
                       uint32_t iteration = 0) {
  const SortSpace s = carve(work, m);
  unsigned char* p = reinterpret_cast<unsigned char*>(s.off) + align256(4 * (size_t)KSORT_MAXCH * KBINS);
  GmfSeriesIn* sin = reinterpret_cast<GmfSeriesIn*>(p);
  p += align256(sizeof(GmfSeriesIn) * (size_t)m);
  constexpr int ND = ORDER == 0 ? 3 : (ORDER == 1 ? 9 : 19);
  GmfNodesRe* nre = reinterpret_cast<GmfNodesRe*>(p);
  p += align256(sizeof(GmfNodesRe) * (size_t)m);
  double* nim = reinterpret_cast<double*>(p);
  p += align256(6 * sizeof(double) * (size_t)m);
  double* sre = reinterpret_cast<double*>(p);
  p += align256(sizeof(double) * ND * (size_t)m);
  double* sim = reinterpret_cast<double*>(p);
  int64_t chunk = (m + KSORT_MAXCH - 1) / KSORT_MAXCH;
  chunk = (chunk + SORT_THREADS - 1) / SORT_THREADS * SORT_THREADS;
  const int nch = (int)((m + chunk - 1) / chunk);
  const unsigned gl = (unsigned)((m + KLEN_THREADS - 1) / KLEN_THREADS);
  hipLaunchKernelGGL((gmf_prep_kernel<KIND, ORDER, PA>), dim3(gl), dim3(KLEN_THREADS), 0, st, m, rows, A, al, alv, bt,
                     s.key, sin, nre, nim, Aw, sigma, seed, iteration);
  hipLaunchKernelGGL(key_hist_kernel, dim3(nch), dim3(SORT_THREADS), 0, st, m, chunk, s.key, s.hist);
  hipLaunchKernelGGL(key_scan_kernel, dim3(1), dim3(SORT_THREADS), 0, st, nch, s.hist, s.off);
  hipLaunchKernelGGL(key_scatter_kernel, dim3(nch), dim3(SORT_THREADS), 0, st, m, chunk, s.key, s.off, s.perm);
  hipLaunchKernelGGL((gmf_series_kernel<KIND, ORDER, PA>), dim3(gl), dim3(KLEN_THREADS), 0, st, m, s.perm, sin, bt,
                     sre, sim);
  return GmfStaged{nre, nim, sre, sim, 0, -1};
}

template <int KIND, int ORDER, class OT, bool PA>
void staged_order(hipStream_t st, int64_t m, const double* A, OT* F, OT* J, OT* H, double al, const double* alv,
                  const GmfCoef& bt, void* work) {
  const GmfStaged sg = staged_front<KIND, ORDER, PA>(st, m, nullptr, A, al, alv, bt, work);
  hipLaunchKernelGGL((gmf_finish_kernel<KIND, ORDER, OT, PA>), dim3((unsigned)((m + WAVE - 1) / WAVE)), dim3(WAVE), 0,
                     st, m, A, F, J, H, al, alv, bt, sg);
}

template <int KIND, class OT, bool PA>
void staged_part(hipStream_t st, int order, int64_t m, const double* A, OT* F, OT* J, OT* H, double al,
                 const double* alv, const GmfCoef& bt, void* work) {
  if (order == 2) staged_order<KIND, 2, OT, PA>(st, m, A, F, J, H, al, alv, bt, work);
  else if (order == 1) staged_order<KIND, 1, OT, PA>(st, m, A, F, J, H, al, alv, bt, work);
  else staged_order<KIND, 0, OT, PA>(st, m, A, F, J, H, al, alv, bt, work);
}

size_t perm_bytes(int64_t m) {
  return align256((size_t)m) + align256(4 * (size_t)m) + 2 * align256(4 * (size_t)KSORT_MAXCH * KBINS);
}

// the most matrices a workspace of `bytes` takes in one part with the given layout (per-matrix bytes `per`, plus
// the fixed tables), for the byte count function `need`
template <class Need>
int64_t capacity_of(size_t bytes, size_t per, Need need) {
  const size_t fixed = 2 * align256(4 * (size_t)KSORT_MAXCH * KBINS);
  if (bytes < fixed) return 0;
  int64_t m = (int64_t)((bytes - fixed) / per);
  while (m > 0 && need(m) > bytes) m--;                     // (a few steps: the alignment padding)
  return m;
}

// the strategy table: [function slot: exp, log, pow, sqrt, per-matrix exponents][order][precision]
std::atomic<int> g_strategy[5][3][2];
std::once_flag g_strategy_once;

// defaults measured on MI355X (general mix, 2^20 matrices): short series (exp) and order 0 gain nothing from the
// length order (its passes cost more than the shorter loops save); for log / powers the staged evaluation is faster
// at order 1 (20-29%) and at order 2 (mixed 24-25%, FP64 3-10%: its series in real arithmetic for the wavefronts
// whose offsets are all real)
int default_strategy(int slot, int order, int mixed) {
  (void)mixed;
  if (slot == 0 || order == 0) return GMF_SORT_PLAIN;
  return GMF_SORT_STAGED;
}

void init_strategies() {
  std::call_once(g_strategy_once, [] {
    for (int k = 0; k < 5; k++)
      for (int o = 0; o < 3; o++)
        for (int x = 0; x < 2; x++) g_strategy[k][o][x].store(default_strategy(k, o, x));
  });
}

constexpr int64_t STAGED_MIN_PART = (int64_t)1 << 15;   // (smaller parts would leave the GPU half idle per pass)

void plain_launch(hipStream_t st, int slot, int kind, int64_t n, const double* A, double* F, double* J, double* H,
                  double al, const double* alv, const GmfCoef& bt) {
  if (alv) gmf_launch_pow(st, n, alv, A, F, J, H, bt);
  else gmf_launch(st, slot, kind, n, A, F, J, H, al, bt);
}
void plain_launch(hipStream_t st, int slot, int kind, int64_t n, const double* A, float* F, float* J, float* H,
                  double al, const double* alv, const GmfCoef& bt) {
  if (alv) gmf_launch_pow_mixed(st, n, alv, A, F, J, H, bt);
  else gmf_launch_mixed(st, slot, kind, n, A, F, J, H, al, bt);
}

}  // namespace

size_t gmf_staged_list_bytes(int64_t m, int order) { return m > 0 ? staged_bytes(m, order) : 0; }

GmfStaged gmf_staged_list(hipStream_t st, int kind, int order, bool pa, int64_t m, const int64_t* rows,
                          const double* A, double al, const double* alv, const GmfCoef& bt, void* work, double* Aw,
                          double sigma, uint64_t seed, uint32_t iteration) {
#define MF_FRONT(K, P)                                                                                      \
  if (order == 2) return staged_front<K, 2, P>(st, m, rows, A, al, alv, bt, work, Aw, sigma, seed, iteration); \
  if (order == 1) return staged_front<K, 1, P>(st, m, rows, A, al, alv, bt, work, Aw, sigma, seed, iteration); \
  return staged_front<K, 0, P>(st, m, rows, A, al, alv, bt, work, Aw, sigma, seed, iteration);
  if (pa) { MF_FRONT(POW, true) }
  if (kind == EXP) { MF_FRONT(EXP, false) }
  if (kind == LOG) { MF_FRONT(LOG, false) }
  MF_FRONT(POW, false)
#undef MF_FRONT
}

int gmf_sort_strategy_get(int slot, int order, bool mixed) {
  init_strategies();
  return g_strategy[slot][order][mixed ? 1 : 0].load();
}
void gmf_sort_strategy_set(int slot, int order, bool mixed, int strategy) {
  init_strategies();
  g_strategy[slot][order][mixed ? 1 : 0].store(strategy);
}

size_t gmf_sort_workspace(int64_t n) {
  if (n < 0) return 0;
  const size_t a = staged_bytes(n, 2), b = perm_bytes(n);
  return a > b ? a : b;
}

int64_t gmf_sort_capacity(size_t bytes) { return capacity_of(bytes, 5, perm_bytes); }

template <class OT>
int gmf_sorted_launch(hipStream_t st, int slot, int kind, int order, int64_t n, const double* A, OT* F, OT* J, OT* H,
                      double al, const double* alv, const GmfCoef& bt, void* work, size_t bytes) {
  int strat = gmf_sort_strategy_get(slot, order, sizeof(OT) == 4);
  int64_t cap = 0;
  if (strat == GMF_SORT_STAGED) {
    const int nd = order == 0 ? 3 : (order == 1 ? 9 : 19);
    const size_t per = 5 + sizeof(GmfSeriesIn) + sizeof(GmfNodesRe) + (6 + 2 * nd) * sizeof(double);
    cap = capacity_of(bytes, per, [order](int64_t m) { return staged_bytes(m, order); });
    if (cap < (n < STAGED_MIN_PART ? n : STAGED_MIN_PART)) strat = GMF_SORT_PERM;   // (too little room: permuted)
  }
  if (strat == GMF_SORT_PERM) cap = gmf_sort_capacity(bytes);
  if (strat == GMF_SORT_PLAIN) cap = MAX_PER_LAUNCH;
  cap = cap < MAX_PER_LAUNCH ? cap : MAX_PER_LAUNCH;            // (32-bit permutation entries, grids below 2^32)
  if (cap <= 0) return -1;
  for (int64_t lo = 0; lo < n; lo += cap) {
    const int64_t m = n - lo < cap ? n - lo : cap;
    const double* Ap = A + 9 * lo;
    OT *Fp = F + 9 * lo, *Jp = J ? J + 81 * lo : nullptr, *Hp = H ? H + 729 * lo : nullptr;
    const double* ap = alv ? alv + lo : nullptr;
    if (strat == GMF_SORT_PLAIN) {
      plain_launch(st, slot, kind, m, Ap, Fp, Jp, Hp, al, ap, bt);
    } else if (strat == GMF_SORT_STAGED) {
      if (alv) staged_part<POW, OT, true>(st, order, m, Ap, Fp, Jp, Hp, 1.0, ap, bt, work);
      else if (kind == EXP) staged_part<EXP, OT, false>(st, order, m, Ap, Fp, Jp, Hp, al, nullptr, bt, work);
      else if (kind == LOG) staged_part<LOG, OT, false>(st, order, m, Ap, Fp, Jp, Hp, al, nullptr, bt, work);
      else staged_part<POW, OT, false>(st, order, m, Ap, Fp, Jp, Hp, al, nullptr, bt, work);
    } else {
      if (alv) sorted_part<POW, OT, true>(st, order, m, Ap, Fp, Jp, Hp, 1.0, ap, bt, work);
      else if (kind == EXP) sorted_part<EXP, OT, false>(st, order, m, Ap, Fp, Jp, Hp, al, nullptr, bt, work);
      else if (kind == LOG) sorted_part<LOG, OT, false>(st, order, m, Ap, Fp, Jp, Hp, al, nullptr, bt, work);
      else sorted_part<POW, OT, false>(st, order, m, Ap, Fp, Jp, Hp, al, nullptr, bt, work);
    }
    if (hipGetLastError() != hipSuccess) return -2;
  }
  return 0;
}
template int gmf_sorted_launch<double>(hipStream_t, int, int, int, int64_t, const double*, double*, double*, double*,
                                       double, const double*, const GmfCoef&, void*, size_t);
template int gmf_sorted_launch<float>(hipStream_t, int, int, int, int64_t, const double*, float*, float*, float*,
                                      double, const double*, const GmfCoef&, void*, size_t);

}  // namespace mf

// ---- gmf_hip_eval_sorted_noalloc: the length-sorted evaluation with its workspace in a static device array, for
// callers that cannot allocate (graph capture, serving loops): at most STATIC_MAX_N matrices per call.  Each device
// holds its own copy of `scratch` (21.2 MB) from the first call on it until the process ends.  The kernels write every
// byte of the workspace they read (keys, histograms, offsets, permutation), so nothing is cleared between calls.
// Calls that overlap on one device (concurrent streams, threads) share the array and race; calls on one stream
// follow each other.
namespace mf {
namespace {
constexpr int64_t STATIC_MAX_N = (int64_t)1 << 22;
constexpr size_t STATIC_BYTES = (size_t)STATIC_MAX_N + 4 * (size_t)STATIC_MAX_N +
                                2 * (((size_t)4 * KSORT_MAXCH * KBINS + 255) / 256 * 256);
static __device__ __align__(256) unsigned char scratch[STATIC_BYTES];

constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mutex;
void* g_scratch_ptr[MAX_DEVICES] = {};

// the device address of the current device's copy of `scratch` (looked up once per device), or nullptr
void* scratch_of_current_device() {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= MAX_DEVICES) return nullptr;
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch_ptr[dev]) {
    void* ptr = nullptr;
    if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) != hipSuccess || !ptr) return nullptr;
    g_scratch_ptr[dev] = ptr;
  }
  return g_scratch_ptr[dev];
}
}  // namespace

}  // namespace mf

extern "C" int gmf_hip_eval_sorted_noalloc(int kind, double alpha, int64_t n, const double* A, double* F, double* J,
                                           double* H, void* stream) {
  using namespace mf;
  if (kind < EXP || kind > SQRT || n < 0 || (H && !J) || (kind == POW && !(alpha > 0))) return -1;
  if (n > STATIC_MAX_N) return -1;                        // (beyond the static workspace: split the batch)
  const int slot = kind;
  if (kind == SQRT) kind = POW, alpha = 0.5;
  if (kind != POW) alpha = 0.0;
  if (n == 0) return 0;
  void* work = scratch_of_current_device();
  if (!work) return -2;
  GmfCoef bt;
  gmf_coefs(kind, alpha, bt);
  const int order = H ? 2 : (J ? 1 : 0);
  return gmf_sorted_launch<double>((hipStream_t)stream, slot, kind, order, n, A, F, J, H, alpha, nullptr, bt, work,
                                   STATIC_BYTES);
}
