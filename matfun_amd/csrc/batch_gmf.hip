// batch_gmf.hip -- the per-class kernels of heterogeneous batches for the gmf classes 12 .. 23 (batch_kernels.hpp).
#define MF_BATCH_KERNELS
#include "batch_kernels.hpp"

namespace mf {

int batch_upload_gmf(const BatchTables& t) { return upload_tables(t); }

void batch_launch_gmf(int cls, hipStream_t st, unsigned grid, bool loop, const BatchArgs& a, double* F, double* J,
                      double* H) {
  launch_class_from<12, 24, double>(cls, st, grid, loop, a, F, J, H);
}
void batch_launch_gmf(int cls, hipStream_t st, unsigned grid, bool loop, const BatchArgs& a, float* F, float* J,
                      float* H) {
  launch_class_from<12, 24, float>(cls, st, grid, loop, a, F, J, H);
}

bool batch_gmf_stageable(int cls) { return cls_fam(cls) == 1 && cls_slot(cls) != 0 && cls_order(cls) >= 1; }

template <int C, class OT>
void launch_staged_from(int cls, hipStream_t st, unsigned grid, const BatchArgs& a, int64_t off0, const GmfStaged& sg,
                        OT* F, OT* J, OT* H) {
  if constexpr (C < 24) {
    if constexpr (cls_slot(C) != 0 && cls_order(C) >= 1) {
      if (cls == C) {
        hipLaunchKernelGGL((batch_staged_kernel<C, OT>), dim3(grid), dim3(WAVE), 0, st, a, off0, sg, F, J, H);
        return;
      }
    }
    launch_staged_from<C + 1, OT>(cls, st, grid, a, off0, sg, F, J, H);
  }
}

void batch_launch_gmf_staged(int cls, hipStream_t st, unsigned grid, const BatchArgs& a, int64_t off0,
                             const GmfStaged& sg, double* F, double* J, double* H) {
  launch_staged_from<12, double>(cls, st, grid, a, off0, sg, F, J, H);
}
void batch_launch_gmf_staged(int cls, hipStream_t st, unsigned grid, const BatchArgs& a, int64_t off0,
                             const GmfStaged& sg, float* F, float* J, float* H) {
  launch_staged_from<12, float>(cls, st, grid, a, off0, sg, F, J, H);
}

}  // namespace mf
