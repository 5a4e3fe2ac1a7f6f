// batch_smf.hip -- the per-class kernels of heterogeneous batches for the smf classes 0 .. 11 (batch_kernels.hpp).
#define MF_BATCH_KERNELS
#include "batch_kernels.hpp"

namespace mf {

int batch_upload_smf(const BatchTables& t) { return upload_tables(t); }

void batch_launch_smf(int cls, hipStream_t st, unsigned grid, bool loop, const BatchArgs& a, double* F, double* J,
                      double* H) {
  launch_class_from<0, 12, double>(cls, st, grid, loop, a, F, J, H);
}
void batch_launch_smf(int cls, hipStream_t st, unsigned grid, bool loop, const BatchArgs& a, float* F, float* J,
                      float* H) {
  launch_class_from<0, 12, float>(cls, st, grid, loop, a, F, J, H);
}

}  // namespace mf
