/* oracle_common.c -- batched entry points of the CPU oracle (test infrastructure; see oracle.h).  Argument checks
 * and normalisation as c/src/matfun.cc:23-32, the -3 rule for non-symmetric input as :34-53, :82; matrices are
 * independent, spread over OpenMP threads (static chunks; results do not depend on the thread count). */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "oracle.h"
#include "oracle_internal.h"

static int g_threads = 0;
void oracle_set_threads(int nthreads) { g_threads = nthreads; }
static int nthreads_for(int64_t n) {
#ifdef _OPENMP
  int t = g_threads > 0 ? g_threads : omp_get_max_threads();
  if (n < 64) t = 1;
  return t;
#else
  (void)n;
  return 1;
#endif
}

int oracle_smf_eval(int kind, double alpha, int64_t n, const double *v, double *F, double *J, double *H) {
  if (!oracle_prepare(&kind, &alpha, n, J, H, 0)) return -1;
  double c2[SMF_NTERMS];
  smf_dd2_coefs(kind, alpha, c2);
#pragma omp parallel for schedule(static) num_threads(nthreads_for(n))
  for (int64_t i = 0; i < n; i++)
    smf_eval_one(kind, alpha, v + 6 * i, F + 6 * i, J ? J + 36 * i : NULL, H ? H + 216 * i : NULL, c2);
  return 0;
}

int oracle_smf_eval_alpha(int kind, const double *alpha, int64_t n, const double *v, double *F, double *J, double *H) {
  if (kind != K_POW || n < 0 || (H && !J)) return -1;
  for (int64_t i = 0; i < n; i++)
    if (!isfinite(alpha[i])) return -1;
#pragma omp parallel for schedule(static) num_threads(nthreads_for(n))
  for (int64_t i = 0; i < n; i++) {
    double c2[SMF_NTERMS];
    smf_dd2_coefs(kind, alpha[i], c2);
    smf_eval_one(kind, alpha[i], v + 6 * i, F + 6 * i, J ? J + 36 * i : NULL, H ? H + 216 * i : NULL, c2);
  }
  return 0;
}

static int smf_matrix_batch(int kind, const double *alphas, double alpha, int64_t n, const double *A, double *F,
                            double *J, double *H) {
  int asym = 0;
#pragma omp parallel for schedule(static) reduction(| : asym) num_threads(nthreads_for(n))
  for (int64_t i = 0; i < n; i++) {
    const double al = alphas ? alphas[i] : alpha;
    double c2[SMF_NTERMS], v[6], Fv[6], Jv[36], *Hv = H ? (double *)malloc(216 * sizeof(double)) : NULL;
    smf_dd2_coefs(kind, al, c2);
    asym |= !smf_matrix_to_voigt(A + 9 * i, v);
    smf_eval_one(kind, al, v, Fv, J ? Jv : NULL, Hv, c2);
    smf_voigt_to_matrix(Fv, J ? Jv : NULL, Hv, F + 9 * i, J ? J + 81 * i : NULL, H ? H + 729 * i : NULL);
    free(Hv);
  }
  return asym;
}

int oracle_smf_eval_matrix(int kind, double alpha, int64_t n, const double *A, double *F, double *J, double *H) {
  if (!oracle_prepare(&kind, &alpha, n, J, H, 0)) return -1;
  return smf_matrix_batch(kind, NULL, alpha, n, A, F, J, H) ? -3 : 0;
}

int oracle_smf_eval_matrix_alpha(int kind, const double *alpha, int64_t n, const double *A, double *F, double *J,
                                 double *H) {
  if (kind != K_POW || n < 0 || (H && !J)) return -1;
  for (int64_t i = 0; i < n; i++)
    if (!isfinite(alpha[i])) return -1;
  smf_matrix_batch(kind, alpha, 0.0, n, A, F, J, H);
  return 0;
}

static void gmf_batch(int kind, const double *alphas, double alpha, int64_t n, const double *A, double *F, double *J,
                      double *H) {
  const int order = H ? 2 : (J ? 1 : 0);
  double bt_shared[GMF_KMAX];
  if (!alphas) gmf_series_coefs(kind, alpha, bt_shared);
  else gmf_series_coefs(K_EXP, 0.0, bt_shared);            /* (builds the static tables before the threads start) */
#pragma omp parallel for schedule(static) num_threads(nthreads_for(n))
  for (int64_t i = 0; i < n; i++) {
    double bt_own[GMF_KMAX];
    const double *bt = bt_shared;
    double al = alpha;
    if (alphas) {
      al = alphas[i];
      gmf_series_coefs(kind, al, bt_own);
      bt = bt_own;
    }
    gmf_eval_one(kind, order, al, A + 9 * i, F + 9 * i, J ? J + 81 * i : NULL, H ? H + 729 * i : NULL, bt);
  }
}

int oracle_gmf_eval(int kind, double alpha, int64_t n, const double *A, double *F, double *J, double *H) {
  if (!oracle_prepare(&kind, &alpha, n, J, H, 1)) return -1;
  gmf_batch(kind, NULL, alpha, n, A, F, J, H);
  return 0;
}

int oracle_gmf_eval_alpha(int kind, const double *alpha, int64_t n, const double *A, double *F, double *J, double *H) {
  if (kind != K_POW || n < 0 || (H && !J)) return -1;
  for (int64_t i = 0; i < n; i++)
    if (!isfinite(alpha[i])) return -1;
  gmf_batch(kind, alpha, 0.0, n, A, F, J, H);
  return 0;
}

int oracle_gmf_terms(int kind, double alpha, int order, const double *A) {
  double J[81], H[729], F[9];
  if (!oracle_prepare(&kind, &alpha, 1, J, H, 1) || order < 0 || order > 2) return -1;
  double bt[GMF_KMAX];
  gmf_series_coefs(kind, alpha, bt);
  return gmf_eval_one(kind, order, alpha, A, F, order >= 1 ? J : NULL, order >= 2 ? H : NULL, bt);
}
