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

#include <stddef.h>
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

/* ---- general matrices with a workspace, evaluated in the order of their series lengths (no counterpart in the
 * reference API): gmf_hip_eval / gmf_hip_eval_pow (and their _mixed forms) for batches whose matrices need group series
 * of very different lengths.  A wavefront of 64 matrices runs its series loop as long as its longest series (on the
 * general workload mix the longest of 64 averages 3x the mean), so the library may sort the matrices by series length
 * on the device first.  How, per function (kind 0..3; 4 = per-matrix exponents), order and precision, is a
 * process-wide strategy (gmf_hip_set_sort_strategy; defaults measured on MI355X, matfun_amd.tune times them):
 *   0  no sorting: the kernel of gmf_hip_eval (bitwise its results)
 *   1  the matrices in length order through the same kernel, each writing its own rows (bitwise its results)
 *   2  staged: per matrix its nodes and series inputs (input order), the series in the order of class and length
 *      (real arithmetic for the wavefronts whose group offsets are all real), the rest in input order with contiguous
 *      stores -- gmf_hip_eval's results up to rounding (per matrix normwise <= 1e-13 plus a few times gmf_hip_eval's
 *      own distance to the reference on that matrix; the same error distribution against the reference)
 * `work` is caller-owned device memory of work_bytes bytes; gmf_hip_sort_workspace(n) bytes (about 540 bytes per
 * matrix plus 200 KB) take any batch of n in one pass; a smaller workspace runs the batch in consecutive parts (a
 * staged strategy uses the permuted one when the workspace holds fewer than 2^15 matrices of it).  Same arguments,
 * layouts, return codes and asynchrony as gmf_hip_eval, plus -1 when n > 0 and the workspace is NULL or holds fewer
 * than min(n, 64) matrices of the permuted strategy (gmf_hip_sort_workspace(min(n, 64)) bytes always suffice).  Calls
 * that share a workspace must not overlap (stream order suffices). */
size_t gmf_hip_sort_workspace(int64_t n);
int gmf_hip_eval_sorted(int kind, double alpha, int64_t n, const double *A, double *F, double *J, double *H, void *work,
                        size_t work_bytes, void *stream);
int gmf_hip_eval_sorted_mixed(int kind, double alpha, int64_t n, const double *A, float *F, float *J, float *H,
                              void *work, size_t work_bytes, void *stream);
int gmf_hip_eval_pow_sorted(const double *alpha, int64_t n, const double *A, double *F, double *J, double *H,
                            void *work, size_t work_bytes, void *stream);
int gmf_hip_eval_pow_sorted_mixed(const double *alpha, int64_t n, const double *A, float *F, float *J, float *H,
                                  void *work, size_t work_bytes, void *stream);
/* gmf_hip_eval_sorted without the workspace argument, for callers that cannot allocate device memory (code under
 * hipGraph capture, tight serving loops, harnesses that forbid hipMalloc): no counterpart in the reference.  The
 * workspace is a static device array of the library (21.2 MB on each device the library runs on, held until the
 * process ends); nothing is allocated per call.  At most 2^22 matrices per call (-1 beyond: split the batch); the
 * same strategies, arguments and return codes as gmf_hip_eval_sorted otherwise, FP64.  The array holds a staged
 * evaluation of up to about 39000 matrices at order 2 (56000 at order 1) in one part, larger batches in parts of that
 * size: the same results as gmf_hip_eval_sorted bitwise (a matrix's results do not depend on the parts a batch runs
 * in or on its place in the batch).  Calls that overlap on one device share the array and race: issue them on one
 * stream (or serialise them). */
int gmf_hip_eval_sorted_noalloc(int kind, double alpha, int64_t n, const double *A, double *F, double *J, double *H,
                                void *stream);
/* the strategy of gmf_hip_eval_sorted* for a function (kind 0..3; 4 = per-matrix exponents), order (0..2) and
 * precision (mixed 0 FP64, 1 FP32 outputs): 0, 1, 2 as above.  Returns 0 / -1 (set), the strategy or -1 (get). */
int gmf_hip_set_sort_strategy(int kind, int order, int mixed, int strategy);
int gmf_hip_get_sort_strategy(int kind, int order, int mixed);

/* ---- packed symmetric outputs (no counterpart in the reference API; the H packing is the reference kernel's own
 * HPACK layout, c/include/matfun/smf_kernel.hpp:196-199): Voigt input v (n,6) as smf_hip_eval, outputs of the
 * distinct values only --
 *   F (n,6)    the Voigt vector of f(A)
 *   J (n,21)   Jp[tri(i,k)] = dF_ab / dA_cd, i = (a,b), k = (c,d) Voigt indices, i <= k (symmetric in i <-> k)
 *   H (n,126)  Hp[21 i + tri(p,q)] = d2 F_i / dv_p dv_q, p <= q (symmetric in p <-> q)
 * with tri(p,q) = p (11 - p) / 2 + q the row-major upper triangle of 6x6.  1224 instead of 6624 bytes per matrix at
 * order 2 (FP64).  The API layouts follow with exact weights: Voigt J[6i+k] = Jp[tri(min,max)] w(k), w = 1 (k < 3)
 * or 2; Voigt H[(6i+p)6+q] = Hp[21 i + tri(min(p,q), max(p,q))]; matrix layout J[ab][cd] = Jp[tri(i,k)],
 * H[ab][cd][ef] = Hp[21 i + tri(k,l)] u(cd) u(ef), u = 1 (diagonal) or 1/2 -- bitwise the results of
 * smf_hip_eval / smf_hip_eval_matrix.  Same arguments and return codes as smf_hip_eval. */
int smf_hip_eval_packed(int kind, double alpha, int64_t n, const double *v, double *F, double *J, double *H,
                        void *stream);
int smf_hip_eval_packed_mixed(int kind, double alpha, int64_t n, const double *v, float *F, float *J, float *H,
                              void *stream);

/* ---- heterogeneous batches (no counterpart in the reference API, which evaluates one function per call): one
 * persistent kernel evaluates groups of matrices with different functions and orders -- e.g. the integration points
 * of finite elements, each element with its own constitutive function.  Groups follow one another: the matrices of
 * group g are rows first_g .. first_g + count_g - 1 of A (N x 9, 3x3 matrix layout for both families) and F (N x 9);
 * J has one row of 81 per matrix of order >= 1 and H one row of 729 per matrix of order 2, in the same order (rows
 * of matrices of lower order are skipped, not left empty).  smf groups take symmetric matrices (NaN results for a
 * non-symmetric one, as smf_hip_eval_matrix).  Results are those of the homogeneous entry points on the same
 * matrices: smf_hip_eval_matrix / gmf_hip_eval, and for pow smf_hip_eval_matrix_pow / gmf_hip_eval_pow with the
 * per-matrix exponents (bitwise: the same tile code runs). */
typedef struct mf_hip_group {
  int32_t family;   /* 0 smf (symmetric), 1 gmf (general) */
  int32_t kind;     /* 0 exp, 1 log, 2 pow, 3 sqrt */
  int32_t order;    /* 0: F; 1: F, J; 2: F, J, H */
  int32_t count;    /* matrices in the group, >= 0 */
  double alpha;     /* pow: the group's exponent (finite) when mf_hip_batch_eval gets no per-matrix exponents */
} mf_hip_group;

typedef struct mf_hip_plan mf_hip_plan;          /* opaque */
typedef struct mf_hip_sequence mf_hip_sequence;  /* opaque */

/* Host set-up (allocates device tables: tiles of up to 64 matrices of one class, per-matrix rows, series tables;
 * synchronous).  Returns 0, -1 invalid groups (or 2^31 matrices or more in all), -2 device error.
 * A plan must outlive the sequences captured from it (their graphs use its tables): mf_hip_plan_destroy returns -1
 * and frees nothing while a sequence of the plan is not yet destroyed (0 otherwise; NULL is a no-op).  The caller
 * makes sure no evaluation on the plan is still queued when it is destroyed. */
int mf_hip_plan_create(const mf_hip_group *groups, int64_t ngroups, mf_hip_plan **plan);
int mf_hip_plan_destroy(mf_hip_plan *plan);
/* sizes of the batch: what 0 = matrices N, 1 = J rows, 2 = H rows, 3 = tiles, 4 = classes present */
int64_t mf_hip_plan_size(const mf_hip_plan *plan, int what);

/* Evaluate the batch (asynchronous on `stream`, allocates nothing, capturable in a hipGraph).  alpha: NULL (each pow
 * group uses its alpha) or a device array of N exponents (read for the matrices of pow groups only; any finite value,
 * NaN results for a non-finite one).  J / H must be non-NULL when a group of order >= 1 / 2 is present.  Returns 0,
 * -1 invalid arguments, -2 launch error. */
int mf_hip_batch_eval(const mf_hip_plan *plan, const double *A, const double *alpha, double *F, double *J, double *H,
                      void *stream);
int mf_hip_batch_eval_mixed(const mf_hip_plan *plan, const double *A, const double *alpha, float *F, float *J,
                            float *H, void *stream);

/* Launch strategy (modes 0-2 give bitwise the same results):
 *   mode 0 fused: one kernel over the tiles of all classes, tiles in class order (order 0) or interleaved in
 *          proportion to the classes' sizes (order 1)
 *   mode 1 split: one kernel per class (each compiled for its class alone), one after another on `stream`
 *   mode 2 split, concurrent: the class kernels on `stream` and up to 3 streams of the plan (forked from and joined
 *          back to `stream` with events, so the call stays asynchronous and capturable; calls that share a plan
 *          must not overlap)
 *   mode 3 split, staged: as mode 1, but the general-matrix log, sqrt and pow classes of order 1 and 2 run the
 *          staged evaluation of mf_hip_gmf_eval_sorted (strategy 2) over their matrices: their results agree with
 *          mode 1's to rounding (the series summed in another order; the same bound), the other classes' bitwise.
 *          Uses a workspace of the plan (about 0.6 KB per matrix of the largest such class), allocated by this call
 *          or by mf_hip_plan_tune; calls that share a plan must not overlap.  Returns -2 if the allocation fails.
 *   per_cu > 0: persistent, per_cu workgroups per compute unit loop over the tiles; 0: one workgroup per tile
 *          (ignored by mode 3)
 * mf_hip_plan_tune times the split candidates (modes 1 and 2, one workgroup per tile or 2, 4, 8 per CU, and mode 3
 * when the batch has a staged class; mode 0 is measured slower on MI355X in every form -- its register use and LDS are
 * those of the largest class for every tile -- and is chosen only through mf_hip_plan_set_strategy) on the caller's
 * buffers (the batch is evaluated several times) and keeps the fastest; `result` (optional) receives {mode, per_cu,
 * order, microseconds of the best}.  Synchronous. */
int mf_hip_plan_set_strategy(mf_hip_plan *plan, int mode, int per_cu, int order);
int mf_hip_plan_tune(mf_hip_plan *plan, const double *A, const double *alpha, void *F, void *J, void *H, int mixed,
                     void *stream, double *result);

/* Newton-iteration driver: A <- A o (1 + rel_sigma Z), Z standard normal from Philox4x32-10 (Salmon et al., SC'11)
 * keyed by seed, counter (matrix, iteration, block): entry-wise relative noise, the same for (i, j) and (j, i) of
 * smf matrices so that they stay symmetric.  Asynchronous on `stream`. */
int mf_hip_perturb(const mf_hip_plan *plan, double *A, double rel_sigma, uint64_t seed, int iteration, void *stream);

/* A sequence of Newton iterations captured in one hipGraph: A <- A0 (device copy), then `iterations` times
 * { mf_hip_perturb(iteration i = 1..iterations); mf_hip_batch_eval(_mixed) } with the plan's strategy at creation --
 * the perturbation applied by the evaluation's own kernels as they read A (one pass over A per iteration; A and the
 * results bitwise those of the two calls).  F / J / H hold the last iteration's results.  Replays are independent (each starts from A0).  Creation captures and
 * instantiates (synchronous); launch is asynchronous on `stream`.  The sequence keeps using the plan's tables and the
 * buffers given here: destroy it before the plan (mf_hip_plan_destroy refuses until then) and keep the buffers. */
int mf_hip_sequence_create(const mf_hip_plan *plan, const double *A0, double *A, const double *alpha, void *F, void *J,
                           void *H, int mixed, int iterations, double rel_sigma, uint64_t seed,
                           mf_hip_sequence **seq);
int mf_hip_sequence_launch(mf_hip_sequence *seq, void *stream);
int mf_hip_sequence_destroy(mf_hip_sequence *seq);

/* ---- Newton step of the Hencky-type potential psi(X) = 1/2 |log X|^2 (reference theory.md:190: the Newton method
 * for this potential with the tangent J^T J + log X : H; no counterpart in the reference's API).  From L = log X,
 * J = dL/dX and H = d2L/dX2 in the matrix layout -- as gmf_hip_eval(1, ...) gives them for any real X with positive
 * real eigenvalues, smf_hip_eval_matrix(1, ...) for symmetric positive definite X (the *_mixed forms for the FP32
 * entry point) -- per matrix, summed over a = 0..8:
 *   psi = 1/2 L_a L_a (n),   g_c = L_a J_ac (n,3,3) = dpsi/dX,   K_ce = J_ac J_ae + L_a H_ace (n,3,3,3,3) = dg/dX.
 * psi, g, K may each be NULL (not computed); L is needed for any output, J for g or K, H for K (else -1).  Sums in
 * FP64 for both forms.  Device pointers; asynchronous on `stream`; n == 0 returns 0 without a launch; NaN inputs
 * give NaN outputs for their matrix only.  Returns 0, -1 invalid arguments, -2 launch error. */
int mf_hip_hencky(int64_t n, const double *L, const double *J, const double *H, double *psi, double *g, double *K,
                  void *stream);
int mf_hip_hencky_mixed(int64_t n, const float *L, const float *J, const float *H, float *psi, float *g, float *K,
                        void *stream);
/* The same step fused with the evaluation, for symmetric X (matrix layout, FP64): from X (n,3,3) directly to
 * L = log X (n,3,3), psi, g, K as above, without J and H in memory (FP64 outputs 91 + 9 values per matrix instead of
 * 819 written by smf_hip_eval_matrix and read again).  Equal to smf_hip_eval_matrix(1, 0, ...) followed by
 * mf_hip_hencky up to the order of the sums (per matrix 1e-13 relative).  Any output may be NULL; NaN outputs for a
 * non-symmetric matrix (as smf_hip_eval_matrix).  Returns 0, -1 (n < 0, outputs without A), -2 launch error. */
int smf_hip_hencky_matrix(int64_t n, const double *A, double *L, double *psi, double *g, double *K, void *stream);
/* The same in Voigt coordinates v = (X11, X22, X33, X12, X13, X23) (n,6), as smf_hip_eval takes them: L = log X
 * (n,6), psi, g_p = dpsi/dv_p (n,6), K_pq = d2psi/dv_p dv_q (n,6,6); these are the matrix-layout g, K summed over the
 * entries of each Voigt index (g_p = m_p g_c, K_pq = m_p m_q K_ce with m = 1 diagonal, 2 off-diagonal). */
int smf_hip_hencky(int64_t n, const double *v, double *L, double *psi, double *g, double *K, void *stream);
/* The same for general real X (FP64; log of X with no real eigenvalue <= 0, else NaN outputs for that matrix):
 * gmf_hip_eval(1, 0, ...) followed by mf_hip_hencky in one kernel. */
int gmf_hip_hencky(int64_t n, const double *A, double *L, double *psi, double *g, double *K, void *stream);

/* ---- tile choice of the order-2 symmetric kernels (no counterpart in the reference), per function (kind 0..3;
 * 4 = pow with per-matrix exponents), layout (matrix_layout 0 Voigt, 1 matrix) and precision (mixed 0 FP64, 1 FP32
 * outputs), for the whole process: 1 = 10 matrices per H store pass with the eigenvectors / second divided
 * differences in their own LDS block; 10, 8, 5 = matrices per pass with that block sharing the first phase's LDS
 * region (less LDS per wavefront: more wavefronts per compute unit; fewer matrices per pass: more lanes idle while H
 * is assembled).  Results do not depend on it.  Defaults measured on MI355X; matfun_amd.tune picks them by timing.
 * Returns 0 / -1 (set), the tile or -1 (get). */
int smf_hip_set_tile(int kind, int matrix_layout, int mixed, int hpp);
int smf_hip_get_tile(int kind, int matrix_layout, int mixed);

/* ---- workgroup shape of the general-matrix kernels (gmf_hip_eval*, gmf_hip_eval_pow*, and strategy 0 of the sorted
 * entry points; no counterpart in the reference), per function (kind 0..3; 4 = per-matrix exponents), order (0..2)
 * and precision (mixed 0 FP64, 1 FP32 outputs), for the whole process: tiles = 1, 2 or 4 wavefronts of 64 matrices
 * per workgroup; with 2 or 4 the wavefronts share out their group series by length (each lane then runs a series of
 * about the length of its neighbours').  Results do not depend on it.  Defaults measured on MI355X (4 for the FP32
 * order-1 kernels but exp's and the order-0 log / pow / sqrt kernels, 1 elsewhere); matfun_amd.tune picks them by
 * timing.  Returns 0 / -1 (set), the choice or -1 (get). */
int gmf_hip_set_tiles(int kind, int order, int mixed, int tiles);
int gmf_hip_get_tiles(int kind, int order, int mixed);

/* ---- write-ceiling probe (no counterpart in the reference; a measurement aid): stores a constant pattern into F, J, H
 * in the layout of an order-2 evaluation of n matrices with nf = 9 (matrix) or 6 (Voigt) values per F row -- per
 * wavefront 64 matrices' F, J and H blocks, each one contiguous range, with the evaluation kernels' 16-byte stores
 * and no loads or arithmetic: the rate at which this GPU takes those three output streams.  J / H may be NULL (not
 * written).  Returns 0, -1 (n < 0, nf not 6 or 9, F NULL with n > 0, H without J, a pointer not 16-byte aligned),
 * -2 (launch error).  Asynchronous. */
int mf_hip_store_probe(int64_t n, int nf, double *F, double *J, double *H, void *stream);
int mf_hip_store_probe_f32(int64_t n, int nf, float *F, float *J, float *H, void *stream);
/* the same after reading the input rows x (n x nin doubles, nin = 9 or 6; coalesced through LDS as the evaluation
 * kernels read them), storing values that depend on them: the whole traffic of an evaluation (read x, write F, J, H)
 * with no arithmetic.  -1 also for x NULL with n > 0 or nin not 6 or 9. */
int mf_hip_store_probe_in(int64_t n, const double *x, int nin, int nf, double *F, double *J, double *H, void *stream);
int mf_hip_store_probe_in_f32(int64_t n, const double *x, int nin, int nf, float *F, float *J, float *H,
                              void *stream);

/* library version ("0.1.3"-compatible API level plus this build); replaces smf_version (c/include/matfun/matfun.h:53) */
const char *smf_hip_version(void);

/* the device code's target ("gfx950"); the counterpart of smf_simd_name (matfun.h:51) */
const char *smf_hip_target(void);

#ifdef __cplusplus
}
#endif

#endif /* MATFUN_HIP_H */
