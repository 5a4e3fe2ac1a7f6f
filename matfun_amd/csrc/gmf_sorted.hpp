// gmf_sorted.hpp -- host side of the length-sorted general-matrix evaluation (gmf_sorted.hip)
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace mf {

struct GmfCoef;

constexpr int KBINS = 201;           // keys 200 - K for series lengths K = 0 .. 200
constexpr int KSORT_MAXCH = 128;     // histogram chunks of one part at most

// workspace bytes for n matrices in one part; the number of matrices a workspace of `bytes` takes in one part
size_t gmf_sort_workspace(int64_t n);
int64_t gmf_sort_capacity(size_t bytes);

// the batch in parts of at most cap matrices (kind normalised as gmf_hip_eval; alv: per-matrix exponents or NULL;
// order from J / H as the C ABI).  Returns 0 or -2 (launch error).
template <class OT>
int gmf_sorted_launch(hipStream_t st, int kind, int order, int64_t n, const double* A, OT* F, OT* J, OT* H, double al,
                      const double* alv, const GmfCoef& bt, void* work, int64_t cap);

}  // namespace mf
