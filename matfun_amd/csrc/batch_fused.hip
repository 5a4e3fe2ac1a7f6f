// batch_fused.hip -- the persistent kernel of heterogeneous batches over tiles of any class (batch_kernels.hpp).
#define MF_BATCH_KERNELS
#include "batch_kernels.hpp"

namespace mf {

int batch_upload_fused(const BatchTables& t) { return upload_tables(t); }

template <class OT>
static void launch(hipStream_t st, unsigned grid, bool loop, const BatchArgs& a, OT* F, OT* J, OT* H) {
  if (loop) hipLaunchKernelGGL((batch_kernel<OT, true>), dim3(grid), dim3(WAVE), 0, st, a, F, J, H);
  else hipLaunchKernelGGL((batch_kernel<OT, false>), dim3(grid), dim3(WAVE), 0, st, a, F, J, H);
}

void batch_launch_fused(hipStream_t st, unsigned grid, bool loop, const BatchArgs& a, double* F, double* J,
                        double* H) {
  launch(st, grid, loop, a, F, J, H);
}
void batch_launch_fused(hipStream_t st, unsigned grid, bool loop, const BatchArgs& a, float* F, float* J, float* H) {
  launch(st, grid, loop, a, F, J, H);
}

}  // namespace mf
