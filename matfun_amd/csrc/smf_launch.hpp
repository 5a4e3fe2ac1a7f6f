// smf_launch.hpp -- host-side entry points of the kernel files (used by the C ABI in matfun_hip.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mf {

struct SmfCoef;
struct GmfCoef;

// one launch over n matrices (n <= MAX_PER_LAUNCH); kind already normalised; J / H select the order
void smf_launch(hipStream_t st, int kind, bool mat, int64_t n, const double* x, double* F, double* J, double* H,
                double al, const SmfCoef& c2);
// (gmf: slot = the function before normalisation, 0 exp .. 3 sqrt, for the tiles-per-workgroup table)
void gmf_launch(hipStream_t st, int slot, int kind, int64_t n, const double* A, double* F, double* J, double* H,
                double al, const GmfCoef& bt);
// mixed form: FP64 inputs, FP32 outputs
void smf_launch_mixed(hipStream_t st, int kind, bool mat, int64_t n, const double* x, float* F, float* J, float* H,
                      double al, const SmfCoef& c2);
void gmf_launch_mixed(hipStream_t st, int slot, int kind, int64_t n, const double* A, float* F, float* J, float* H,
                      double al, const GmfCoef& bt);

// packed outputs (Voigt input): F 6, J 21, H 126 per matrix
void smf_launch_packed(hipStream_t st, int kind, int64_t n, const double* x, double* F, double* J, double* H,
                       double al, const SmfCoef& c2);
void smf_launch_packed_mixed(hipStream_t st, int kind, int64_t n, const double* x, float* F, float* J, float* H,
                             double al, const SmfCoef& c2);
// powers with one exponent per matrix (alv: n device doubles); gmf: recips from gmf_recips
void smf_launch_pow(hipStream_t st, bool mat, int64_t n, const double* alv, const double* x, double* F, double* J,
                    double* H);
void smf_launch_pow_mixed(hipStream_t st, bool mat, int64_t n, const double* alv, const double* x, float* F, float* J,
                          float* H);
void gmf_launch_pow(hipStream_t st, int64_t n, const double* alv, const double* A, double* F, double* J, double* H,
                    const GmfCoef& recips);
void gmf_launch_pow_mixed(hipStream_t st, int64_t n, const double* alv, const double* A, float* F, float* J, float* H,
                          const GmfCoef& recips);

// the H store tile of the order-2 symmetric kernels per (kind 0..3 or 4 = per-matrix pow, layout, precision):
// TILE_SEP10, or 10 / 8 / 5 matrices per pass with V / D2 sharing the phase-A LDS region
constexpr int TILE_SEP10 = 1;
int smf_tile_get(int slot, bool mat, bool mixed);
void smf_tile_set(int slot, bool mat, bool mixed, int hpp);

// wavefronts (tiles of 64 matrices) per workgroup of the homogeneous general-matrix kernels per (slot 0..3 or 4 =
// per-matrix pow, order, precision): 1, 2 or 4, the wavefronts of a workgroup sharing out their group series
int gmf_tiles_get(int slot, int order, bool mixed);
void gmf_tiles_set(int slot, int order, bool mixed, int nw);

constexpr int64_t MAX_PER_LAUNCH = (int64_t)1 << 30;   // matrices per grid (work-items stay below 2^32)

}  // namespace mf
