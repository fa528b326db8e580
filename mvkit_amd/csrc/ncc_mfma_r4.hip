// Split-BF16 coarse kernels with four output rows per wave.
#include "ncc_mfma.hip.h"

namespace ncc {
int coarse_launch_r4(int variant, dim3 grid, int threads, size_t lds, hipStream_t st, const MfmaArgs& a) {
  return CoarseLaunchers<4>::grid_launch(variant, grid, threads, lds, st, a);
}
int coarse_queue_r4(int variant, int threads, size_t lds, hipStream_t st, const QueueArgs& q, int n_cus) {
  return CoarseLaunchers<4>::queue_launch(variant, threads, lds, st, q, n_cus);
}
}  // namespace ncc
