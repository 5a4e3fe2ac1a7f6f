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

}  // namespace mf
