/* oracle.h -- CPU restatement of matfun's batched evaluation (BodeTobias/matfun), in plain scalar C, FP64.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP product (matfun_amd/csrc) and the
 * "port" CPU baseline of bench.py; the product never links, loads or calls it.  Only tests/, the smoke
 * check and bench.py's cpu_baseline leg may use it.
 *
 * Same entry points, layouts and return codes as the reference C API (c/include/matfun/matfun.h:28-47):
 *   kind 0 exp, 1 log, 2 pow (alpha > 0), 3 sqrt; row-major, matrix index first; J and H may be NULL; H needs J.
 *   oracle_smf_eval        Voigt layout:  v (n,6) -> F (n,6), J (n,6,6), H (n,6,6,6)
 *   oracle_smf_eval_matrix matrix layout: A (n,3,3) -> F (n,3,3), J (n,3^4), H (n,3^6); -3 if a matrix is not
 *                          symmetric (its results NaN), as c/src/matfun.cc:34-53, :82
 *   oracle_gmf_eval        general real matrices, matrix layout (c/include/matfun/gmf_kernel.hpp)
 * Returns 0, -1 (invalid arguments, c/src/matfun.cc:23-27) or -3.
 *
 * Extension (no counterpart in the reference API): oracle_smf_eval_alpha / oracle_smf_eval_matrix_alpha /
 * oracle_gmf_eval_alpha take one
 * exponent per matrix (pow only; any finite value -- the reference requires alpha > 0; for alpha <= 0 the series
 * radii 3/alpha become 3/|alpha|, identical for alpha > 0) and never return -3 (NaN results instead, as the GPU
 * API does).
 */
#ifndef MATFUN_ORACLE_H
#define MATFUN_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

int oracle_smf_eval(int kind, double alpha, int64_t n, const double *v, double *F, double *J, double *H);
int oracle_smf_eval_matrix(int kind, double alpha, int64_t n, const double *A, double *F, double *J, double *H);
int oracle_gmf_eval(int kind, double alpha, int64_t n, const double *A, double *F, double *J, double *H);

/* per-matrix exponents (kind must be 2 = pow; every alpha[i] finite) */
int oracle_smf_eval_alpha(int kind, const double *alpha, int64_t n, const double *v, double *F, double *J, double *H);
int oracle_smf_eval_matrix_alpha(int kind, const double *alpha, int64_t n, const double *A, double *F, double *J,
                                 double *H);
int oracle_gmf_eval_alpha(int kind, const double *alpha, int64_t n, const double *A, double *F, double *J, double *H);

/* number of series terms the gmf group series of matrix A used (diagnostics; order 0..2) */
int oracle_gmf_terms(int kind, double alpha, int order, const double *A);

/* OpenMP threads used by the batched calls (0: the OpenMP default) */
void oracle_set_threads(int nthreads);

#ifdef __cplusplus
}
#endif
#endif
