// matfun_hip.hip -- the C ABI (include/matfun_hip.h): argument checks, exponent normalisation, host-built series
// tables, batch splitting, launch-error reporting.  Semantics of the reference's CUDA API
// (cuda/src/matfun_cuda.cu:119-141 launch, :169-190 glaunch, :194-207 the exported functions).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/matfun_hip.h"
#include "gmf_device.hpp"
#include "gmf_sorted.hpp"
#include "smf_device.hpp"
#include "smf_launch.hpp"

#ifndef MATFUN_HIP_VERSION
#define MATFUN_HIP_VERSION "0.1.3-hip.1"
#endif

namespace {

bool valid(int kind, double alpha, int64_t n, const void* J, const void* H) {
  return !(kind < mf::EXP || kind > mf::SQRT || n < 0 || (H && !J) || (kind == mf::POW && !(alpha > 0)));
}

// the reference's grid limit: more than 2^31 - 1 blocks of 32 (smf) / 64 (gmf) matrices is an error, reported before
// any launch (matfun_cuda.cu:126-131)
constexpr int64_t SMF_MAX_N = (int64_t)INT32_MAX * 32, GMF_MAX_N = (int64_t)INT32_MAX * 64;

int launched() {
  const hipError_t e = hipGetLastError();
  return e == hipSuccess ? 0 : -2;
}

void launch_chunk(hipStream_t st, int kind, bool mat, int64_t n, const double* x, double* F, double* J, double* H,
                  double al, const mf::SmfCoef& c2) {
  mf::smf_launch(st, kind, mat, n, x, F, J, H, al, c2);
}
void launch_chunk(hipStream_t st, int kind, bool mat, int64_t n, const double* x, float* F, float* J, float* H,
                  double al, const mf::SmfCoef& c2) {
  mf::smf_launch_mixed(st, kind, mat, n, x, F, J, H, al, c2);
}
void glaunch_chunk(hipStream_t st, int slot, int kind, int64_t n, const double* A, double* F, double* J, double* H,
                   double al, const mf::GmfCoef& bt) {
  mf::gmf_launch(st, slot, kind, n, A, F, J, H, al, bt);
}
void glaunch_chunk(hipStream_t st, int slot, int kind, int64_t n, const double* A, float* F, float* J, float* H,
                   double al, const mf::GmfCoef& bt) {
  mf::gmf_launch_mixed(st, slot, kind, n, A, F, J, H, al, bt);
}

void pow_chunk(hipStream_t st, bool mat, int64_t n, const double* alv, const double* x, double* F, double* J,
               double* H) {
  mf::smf_launch_pow(st, mat, n, alv, x, F, J, H);
}
void pow_chunk(hipStream_t st, bool mat, int64_t n, const double* alv, const double* x, float* F, float* J, float* H) {
  mf::smf_launch_pow_mixed(st, mat, n, alv, x, F, J, H);
}
void gpow_chunk(hipStream_t st, int64_t n, const double* alv, const double* A, double* F, double* J, double* H,
                const mf::GmfCoef& r) {
  mf::gmf_launch_pow(st, n, alv, A, F, J, H, r);
}
void gpow_chunk(hipStream_t st, int64_t n, const double* alv, const double* A, float* F, float* J, float* H,
                const mf::GmfCoef& r) {
  mf::gmf_launch_pow_mixed(st, n, alv, A, F, J, H, r);
}

// symmetric matrices (x: Voigt vectors or full matrices); splits batches beyond one grid into several launches
template <class OT>
int smf_call(int kind, double alpha, int64_t n, const double* x, bool mat, OT* F, OT* J, OT* H, void* stream) {
  if (!valid(kind, alpha, n, J, H)) return -1;
  if (kind != mf::POW) alpha = kind == mf::SQRT ? 0.5 : 0.0;
  if (n == 0) return 0;
  if (n > SMF_MAX_N) return -2;
  mf::SmfCoef c2;
  mf::smf_coefs(kind, alpha, c2);
  const int nin = mat ? 9 : 6, nf = mat ? 9 : 6;
  for (int64_t lo = 0; lo < n; lo += mf::MAX_PER_LAUNCH) {
    const int64_t m = n - lo < mf::MAX_PER_LAUNCH ? n - lo : mf::MAX_PER_LAUNCH;
    launch_chunk((hipStream_t)stream, kind, mat, m, x + nin * lo, F + nf * lo, J ? J + nf * nf * lo : nullptr,
                 H ? H + nf * nf * nf * lo : nullptr, alpha, c2);
    if (launched()) return -2;
  }
  return 0;
}

// general matrices; sqrt is pow 0.5 (matfun_cuda.cu:172-173)
template <class OT>
int gmf_call(int kind, double alpha, int64_t n, const double* A, OT* F, OT* J, OT* H, void* stream) {
  if (!valid(kind, alpha, n, J, H)) return -1;
  const int slot = kind;
  if (kind == mf::SQRT) kind = mf::POW, alpha = 0.5;
  if (kind != mf::POW) alpha = 0.0;
  if (n == 0) return 0;
  if (n > GMF_MAX_N) return -2;
  mf::GmfCoef bt;
  mf::gmf_coefs(kind, alpha, bt);
  for (int64_t lo = 0; lo < n; lo += mf::MAX_PER_LAUNCH) {
    const int64_t m = n - lo < mf::MAX_PER_LAUNCH ? n - lo : mf::MAX_PER_LAUNCH;
    glaunch_chunk((hipStream_t)stream, slot, kind, m, A + 9 * lo, F + 9 * lo, J ? J + 81 * lo : nullptr,
                  H ? H + 729 * lo : nullptr, alpha, bt);
    if (launched()) return -2;
  }
  return 0;
}

// general matrices ordered by series length through the caller's workspace (gmf_sorted.hip); the same checks and
// normalisation as gmf_call / gmf_pow_call (alv: per-matrix exponents, else kind and alpha), plus -1 for a workspace
// that holds fewer than min(n, 64) matrices
template <class OT>
int gmf_sorted_call(int kind, double alpha, const double* alv, int64_t n, const double* A, OT* F, OT* J, OT* H,
                    void* work, size_t work_bytes, void* stream) {
  if (alv) kind = mf::POW, alpha = 1.0;
  if (!valid(kind, alpha, n, J, H)) return -1;
  const int slot = alv ? 4 : kind;
  if (kind == mf::SQRT) kind = mf::POW, alpha = 0.5;
  if (kind != mf::POW) alpha = 0.0;
  if (n == 0) return 0;
  const int64_t cap = mf::gmf_sort_capacity(work ? work_bytes : 0);
  if (cap < (n < 64 ? n : 64)) return -1;
  if (n > GMF_MAX_N) return -2;
  mf::GmfCoef bt;
  if (alv) mf::gmf_recips(bt);
  else mf::gmf_coefs(kind, alpha, bt);
  const int order = H ? 2 : (J ? 1 : 0);
  return mf::gmf_sorted_launch<OT>((hipStream_t)stream, slot, kind, order, n, A, F, J, H, alpha, alv, bt, work,
                                   work_bytes);
}

void packed_chunk(hipStream_t st, int kind, int64_t n, const double* x, double* F, double* J, double* H, double al,
                  const mf::SmfCoef& c2) {
  mf::smf_launch_packed(st, kind, n, x, F, J, H, al, c2);
}
void packed_chunk(hipStream_t st, int kind, int64_t n, const double* x, float* F, float* J, float* H, double al,
                  const mf::SmfCoef& c2) {
  mf::smf_launch_packed_mixed(st, kind, n, x, F, J, H, al, c2);
}

// symmetric matrices in Voigt form, packed outputs: F 6, J 21, H 126 values per matrix
template <class OT>
int smf_packed_call(int kind, double alpha, int64_t n, const double* v, OT* F, OT* J, OT* H, void* stream) {
  if (!valid(kind, alpha, n, J, H)) return -1;
  if (kind != mf::POW) alpha = kind == mf::SQRT ? 0.5 : 0.0;
  if (n == 0) return 0;
  if (n > SMF_MAX_N) return -2;
  mf::SmfCoef c2;
  mf::smf_coefs(kind, alpha, c2);
  for (int64_t lo = 0; lo < n; lo += mf::MAX_PER_LAUNCH) {
    const int64_t m = n - lo < mf::MAX_PER_LAUNCH ? n - lo : mf::MAX_PER_LAUNCH;
    packed_chunk((hipStream_t)stream, kind, m, v + 6 * lo, F + 6 * lo, J ? J + 21 * lo : nullptr,
                 H ? H + 126 * lo : nullptr, alpha, c2);
    if (launched()) return -2;
  }
  return 0;
}

// powers with one exponent per matrix (device array alpha of n doubles, any finite value)
template <class OT>
int smf_pow_call(const double* alpha, int64_t n, const double* x, bool mat, OT* F, OT* J, OT* H, void* stream) {
  if (n < 0 || (H && !J) || (n > 0 && !alpha)) return -1;
  if (n == 0) return 0;
  if (n > SMF_MAX_N) return -2;
  const int nin = mat ? 9 : 6, nf = mat ? 9 : 6;
  for (int64_t lo = 0; lo < n; lo += mf::MAX_PER_LAUNCH) {
    const int64_t m = n - lo < mf::MAX_PER_LAUNCH ? n - lo : mf::MAX_PER_LAUNCH;
    pow_chunk((hipStream_t)stream, mat, m, alpha + lo, x + nin * lo, F + nf * lo, J ? J + nf * nf * lo : nullptr,
              H ? H + nf * nf * nf * lo : nullptr);
    if (launched()) return -2;
  }
  return 0;
}

template <class OT>
int gmf_pow_call(const double* alpha, int64_t n, const double* A, OT* F, OT* J, OT* H, void* stream) {
  if (n < 0 || (H && !J) || (n > 0 && !alpha)) return -1;
  if (n == 0) return 0;
  if (n > GMF_MAX_N) return -2;
  mf::GmfCoef recips;
  mf::gmf_recips(recips);
  for (int64_t lo = 0; lo < n; lo += mf::MAX_PER_LAUNCH) {
    const int64_t m = n - lo < mf::MAX_PER_LAUNCH ? n - lo : mf::MAX_PER_LAUNCH;
    gpow_chunk((hipStream_t)stream, m, alpha + lo, A + 9 * lo, F + 9 * lo, J ? J + 81 * lo : nullptr,
               H ? H + 729 * lo : nullptr, recips);
    if (launched()) return -2;
  }
  return 0;
}

}  // namespace

extern "C" int smf_hip_eval(int kind, double alpha, int64_t n, const double* v, double* F, double* J, double* H,
                            void* stream) {
  return smf_call(kind, alpha, n, v, false, F, J, H, stream);
}

extern "C" int smf_hip_eval_matrix(int kind, double alpha, int64_t n, const double* A, double* F, double* J, double* H,
                                   void* stream) {
  return smf_call(kind, alpha, n, A, true, F, J, H, stream);
}

extern "C" int gmf_hip_eval(int kind, double alpha, int64_t n, const double* A, double* F, double* J, double* H,
                            void* stream) {
  return gmf_call(kind, alpha, n, A, F, J, H, stream);
}

extern "C" int smf_hip_eval_mixed(int kind, double alpha, int64_t n, const double* v, float* F, float* J, float* H,
                                  void* stream) {
  return smf_call(kind, alpha, n, v, false, F, J, H, stream);
}

extern "C" int smf_hip_eval_matrix_mixed(int kind, double alpha, int64_t n, const double* A, float* F, float* J,
                                         float* H, void* stream) {
  return smf_call(kind, alpha, n, A, true, F, J, H, stream);
}

extern "C" int gmf_hip_eval_mixed(int kind, double alpha, int64_t n, const double* A, float* F, float* J, float* H,
                                  void* stream) {
  return gmf_call(kind, alpha, n, A, F, J, H, stream);
}

extern "C" int smf_hip_eval_pow(const double* alpha, int64_t n, const double* v, double* F, double* J, double* H,
                                void* stream) {
  return smf_pow_call(alpha, n, v, false, F, J, H, stream);
}

extern "C" int smf_hip_eval_matrix_pow(const double* alpha, int64_t n, const double* A, double* F, double* J,
                                       double* H, void* stream) {
  return smf_pow_call(alpha, n, A, true, F, J, H, stream);
}

extern "C" int gmf_hip_eval_pow(const double* alpha, int64_t n, const double* A, double* F, double* J, double* H,
                                void* stream) {
  return gmf_pow_call(alpha, n, A, F, J, H, stream);
}

extern "C" int smf_hip_eval_pow_mixed(const double* alpha, int64_t n, const double* v, float* F, float* J, float* H,
                                      void* stream) {
  return smf_pow_call(alpha, n, v, false, F, J, H, stream);
}

extern "C" int smf_hip_eval_matrix_pow_mixed(const double* alpha, int64_t n, const double* A, float* F, float* J,
                                             float* H, void* stream) {
  return smf_pow_call(alpha, n, A, true, F, J, H, stream);
}

extern "C" int gmf_hip_eval_pow_mixed(const double* alpha, int64_t n, const double* A, float* F, float* J, float* H,
                                      void* stream) {
  return gmf_pow_call(alpha, n, A, F, J, H, stream);
}

extern "C" int smf_hip_eval_packed(int kind, double alpha, int64_t n, const double* v, double* F, double* J,
                                   double* H, void* stream) {
  return smf_packed_call(kind, alpha, n, v, F, J, H, stream);
}

extern "C" int smf_hip_eval_packed_mixed(int kind, double alpha, int64_t n, const double* v, float* F, float* J,
                                         float* H, void* stream) {
  return smf_packed_call(kind, alpha, n, v, F, J, H, stream);
}

extern "C" size_t gmf_hip_sort_workspace(int64_t n) { return mf::gmf_sort_workspace(n); }

extern "C" int gmf_hip_eval_sorted(int kind, double alpha, int64_t n, const double* A, double* F, double* J, double* H,
                                   void* work, size_t work_bytes, void* stream) {
  return gmf_sorted_call(kind, alpha, nullptr, n, A, F, J, H, work, work_bytes, stream);
}

extern "C" int gmf_hip_eval_sorted_mixed(int kind, double alpha, int64_t n, const double* A, float* F, float* J,
                                         float* H, void* work, size_t work_bytes, void* stream) {
  return gmf_sorted_call(kind, alpha, nullptr, n, A, F, J, H, work, work_bytes, stream);
}

extern "C" int gmf_hip_eval_pow_sorted(const double* alpha, int64_t n, const double* A, double* F, double* J,
                                       double* H, void* work, size_t work_bytes, void* stream) {
  if (n > 0 && !alpha) return -1;
  return gmf_sorted_call(mf::POW, 1.0, n > 0 ? alpha : nullptr, n, A, F, J, H, work, work_bytes, stream);
}

extern "C" int gmf_hip_eval_pow_sorted_mixed(const double* alpha, int64_t n, const double* A, float* F, float* J,
                                             float* H, void* work, size_t work_bytes, void* stream) {
  if (n > 0 && !alpha) return -1;
  return gmf_sorted_call(mf::POW, 1.0, n > 0 ? alpha : nullptr, n, A, F, J, H, work, work_bytes, stream);
}

extern "C" int gmf_hip_set_sort_strategy(int kind, int order, int mixed, int strategy) {
  if (kind < mf::EXP || kind > 4 || order < 0 || order > 2 || strategy < mf::GMF_SORT_PLAIN ||
      strategy > mf::GMF_SORT_STAGED)
    return -1;
  mf::gmf_sort_strategy_set(kind, order, mixed != 0, strategy);
  return 0;
}

extern "C" int gmf_hip_get_sort_strategy(int kind, int order, int mixed) {
  if (kind < mf::EXP || kind > 4 || order < 0 || order > 2) return -1;
  return mf::gmf_sort_strategy_get(kind, order, mixed != 0);
}

extern "C" int gmf_hip_set_tiles(int kind, int order, int mixed, int tiles) {
  if (kind < mf::EXP || kind > 4 || order < 0 || order > 2 || !(tiles == 1 || tiles == 2 || tiles == 4)) return -1;
  mf::gmf_tiles_set(kind, order, mixed != 0, tiles);
  return 0;
}

extern "C" int gmf_hip_get_tiles(int kind, int order, int mixed) {
  if (kind < mf::EXP || kind > 4 || order < 0 || order > 2) return -1;
  return mf::gmf_tiles_get(kind, order, mixed != 0);
}

extern "C" int smf_hip_set_tile(int kind, int matrix_layout, int mixed, int hpp) {
  if (kind < mf::EXP || kind > 4 || (hpp != 10 && hpp != 8 && hpp != 5 && hpp != mf::TILE_SEP10)) return -1;
  mf::smf_tile_set(kind, matrix_layout != 0, mixed != 0, hpp);
  return 0;
}

extern "C" int smf_hip_get_tile(int kind, int matrix_layout, int mixed) {
  if (kind < mf::EXP || kind > 4) return -1;
  return mf::smf_tile_get(kind, matrix_layout != 0, mixed != 0);
}

extern "C" const char* smf_hip_version(void) { return MATFUN_HIP_VERSION; }
extern "C" const char* smf_hip_target(void) { return "gfx950"; }
