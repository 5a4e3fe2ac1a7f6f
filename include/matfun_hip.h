/* matfun_hip.h -- C ABI of libmatfun_hip: batched matrix functions exp, log, sqrt, A^alpha of 3x3 matrices with first
 * and second derivatives on an AMD Instinct MI355X (gfx950), FP64.  A drop-in for the reference's CUDA API
 * (BodeTobias/matfun cuda/include/matfun/matfun_cuda.h:15-22): same arguments, layouts and return codes, with
 * "cuda" replaced by "hip" in the names and `stream` a hipStream_t.
 *
 * Conventions (reference c/include/matfun/matfun.h:1-47, README.md:100-105):
 *   kind   0 exp, 1 log, 2 pow (alpha > 0), 3 sqrt; alpha is ignored for the other kinds
 *   layout row-major, matrix index first, every matrix's results contiguous
 *          Voigt  v (n,6) = (A11, A22, A33, A12, A13, A23) -> F (n,6), J (n,6,6), H (n,6,6,6)
 *          matrix A (n,3,3) -> F (n,3,3), J (n,3,3,3,3) = dF_ab/dA_cd, H (n,3,3,3,3,3,3) = d2F_ab/dA_cd dA_ef
 *   J and H may be NULL (F only, or F and J); H without J is invalid
 *   all pointers are DEVICE pointers owned by the caller; nothing is allocated; the work is queued on `stream`
 *   (NULL = the default stream) and runs asynchronously: no synchronisation, errors only from the launch
 * Returns 0, -1 for invalid arguments (unknown kind, n < 0, H without J, pow with !(alpha > 0)), -2 for a launch
 * error or a batch beyond the grid limit (checked before anything is launched).  n == 0 returns 0 without a launch.
 * Invalid matrices give NaN results only for themselves: a non-symmetric input of smf_hip_eval_matrix
 * (|A_ij - A_ji| > 1e-8 max |A_kl|; the CPU API would return -3) and, for gmf log / sqrt / pow, a matrix with a real
 * eigenvalue <= 0.
 */
#ifndef MATFUN_HIP_H
#define MATFUN_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* symmetric matrices, Voigt layout.  Replaces smf_cuda_eval (cuda/include/matfun/matfun_cuda.h:15,
 * cuda/src/matfun_cuda.cu:199-202). */
int smf_hip_eval(int kind, double alpha, int64_t n, const double *v, double *F, double *J, double *H, void *stream);

/* symmetric matrices, full 3x3 in and out.  Replaces smf_cuda_eval_matrix (matfun_cuda.h:17-18,
 * matfun_cuda.cu:204-207). */
int smf_hip_eval_matrix(int kind, double alpha, int64_t n, const double *A, double *F, double *J, double *H,
                        void *stream);

/* general real matrices, principal branch, the full (unsymmetrised) derivatives.  Replaces gmf_cuda_eval
 * (matfun_cuda.h:21, matfun_cuda.cu:194-197). */
int gmf_hip_eval(int kind, double alpha, int64_t n, const double *A, double *F, double *J, double *H, void *stream);

/* ---- mixed-precision form (no counterpart in the reference API): the same inputs (FP64), arguments, layouts and
 * return codes, outputs in FP32.  The eigen-decomposition, f and the divided differences (with their gap tests and
 * Taylor branches) are FP64; the derivative assembly and all stored values are FP32, which halves the output bytes
 * that bound order 1 and 2.  Normwise relative error against the FP64 form about 1e-7 (F) / 1e-6 (J, H) on
 * well-conditioned input; results must lie in the FP32 range. */
int smf_hip_eval_mixed(int kind, double alpha, int64_t n, const double *v, float *F, float *J, float *H, void *stream);
int smf_hip_eval_matrix_mixed(int kind, double alpha, int64_t n, const double *A, float *F, float *J, float *H,
                              void *stream);
int gmf_hip_eval_mixed(int kind, double alpha, int64_t n, const double *A, float *F, float *J, float *H, void *stream);

/* ---- powers with one exponent per matrix (no counterpart in the reference API, which takes one alpha > 0 per call):
 * F_i = A_i^alpha[i] and its derivatives, for batches that mix exponents (e.g. a material parameter per
 * integration point).  alpha is a DEVICE array of n doubles; any finite value is accepted: alpha = 0 gives I with
 * zero derivatives, alpha < 0 the powers of the inverse (the Taylor radii 3/alpha of the reference become 3/|alpha|,
 * identical for alpha > 0, where the results are those of smf_hip_eval* / gmf_hip_eval with kind 2).  A matrix whose
 * exponent is not finite gets NaN results.  Same layouts as above; returns 0, -1 (n < 0, H without J, alpha NULL with
 * n > 0), -2 (launch error / batch too large). */
int smf_hip_eval_pow(const double *alpha, int64_t n, const double *v, double *F, double *J, double *H, void *stream);
int smf_hip_eval_matrix_pow(const double *alpha, int64_t n, const double *A, double *F, double *J, double *H,
                            void *stream);
int gmf_hip_eval_pow(const double *alpha, int64_t n, const double *A, double *F, double *J, double *H, void *stream);
/* ... in the mixed form (FP32 outputs) */
int smf_hip_eval_pow_mixed(const double *alpha, int64_t n, const double *v, float *F, float *J, float *H,
                           void *stream);
int smf_hip_eval_matrix_pow_mixed(const double *alpha, int64_t n, const double *A, float *F, float *J, float *H,
                                  void *stream);
int gmf_hip_eval_pow_mixed(const double *alpha, int64_t n, const double *A, float *F, float *J, float *H,
                           void *stream);

/* library version ("0.1.3"-compatible API level plus this build); replaces smf_version (c/include/matfun/matfun.h:53) */
const char *smf_hip_version(void);

/* the device code's target ("gfx950"); the counterpart of smf_simd_name (matfun.h:51) */
const char *smf_hip_target(void);

#ifdef __cplusplus
}
#endif

#endif /* MATFUN_HIP_H */
