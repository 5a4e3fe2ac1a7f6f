// gmf_sorted.hpp -- host side of the length-sorted general-matrix evaluation (gmf_sorted.hip)
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace mf {

struct GmfCoef;

// sort keys: the permuted evaluation 200 - K (series lengths K = 0 .. 200, longest first); the staged evaluation
// orders by series class first (gmf_series_class: real offsets without a pole, 200 - K; real with a pole and complex,
// 201 + (200 - K) / 8 and 227 + (200 - K) / 8), so that a wavefront of its series pass holds one class and runs the
// class's own loop
constexpr int KQ = (200 / 8) + 1;
constexpr int KBINS = 201 + 2 * KQ;
constexpr int KSORT_MAXCH = 128;     // histogram chunks of one part at most

// evaluation strategies of the workspace entry points (gmf_hip_eval_sorted*), per function slot (0 exp, 1 log, 2 pow,
// 3 sqrt, 4 per-matrix exponents), order and precision:
//   PLAIN  the homogeneous kernel in input order (gmf_hip_eval; no workspace used)
//   PERM   keys, sort, the tile body over the length-sorted permutation (bitwise gmf_hip_eval's results)
//   STAGED nodes and series inputs per matrix, the series in length order, the rest in input order (contiguous
//          stores; equal to gmf_hip_eval's results up to rounding)
constexpr int GMF_SORT_PLAIN = 0, GMF_SORT_PERM = 1, GMF_SORT_STAGED = 2;
int gmf_sort_strategy_get(int slot, int order, bool mixed);
void gmf_sort_strategy_set(int slot, int order, bool mixed, int strategy);

// workspace bytes for n matrices in one part (every strategy and order); the number of matrices a workspace of
// `bytes` takes in one part of the permuted evaluation (the least room a workspace must offer)
size_t gmf_sort_workspace(int64_t n);
int64_t gmf_sort_capacity(size_t bytes);

// the passes of the staged evaluation before its finishing pass for a list of m matrices (a heterogeneous batch's
// class): matrix k of the list is row rows[k] of A (and of alv, per-matrix exponents when pa); kind normalised as
// gmf_hip_eval (sqrt as pow 1/2 with its table bt).  Launched on st into the workspace work
// (gmf_staged_list_bytes(m, order) bytes); returns where the staged data lie, indexed by k.  With Aw (= A) the nodes
// pass first multiplies each matrix's inputs by its Newton perturbation factors (perturb_factors in philox.hpp:
// sigma, seed, iteration; general, not symmetric) and writes them back over A.
struct GmfStaged;
size_t gmf_staged_list_bytes(int64_t m, int order);
GmfStaged gmf_staged_list(hipStream_t st, int kind, int order, bool pa, int64_t m, const int64_t* rows,
                          const double* A, double al, const double* alv, const GmfCoef& bt, void* work,
                          double* Aw = nullptr, double sigma = 0.0, uint64_t seed = 0, uint32_t iteration = 0);

// the batch through the slot's strategy, in parts of what the workspace holds (the staged evaluation falls back to
// the permuted one when the workspace holds fewer than 2^15 matrices of it); kind normalised as gmf_hip_eval; alv:
// per-matrix exponents or NULL; order from J / H as the C ABI.  Returns 0, -1 (no room) or -2 (launch error).
template <class OT>
int gmf_sorted_launch(hipStream_t st, int slot, int kind, int order, int64_t n, const double* A, OT* F, OT* J, OT* H,
                      double al, const double* alv, const GmfCoef& bt, void* work, size_t bytes);

}  // namespace mf
