// Split-precision GEMM for the Q^T B and Q W products (placeholder until the
// MFMA split kernels land): reports "not supported".
#include "common.h"
#include "kernels.h"

namespace imcom {

size_t gemm_split_work_bytes(int, int, int, int) { return 0; }

hipError_t gemm_split(hipStream_t, int, bool, int, int, int, double, const double*, int64_t,
                      const double*, int64_t, double*, int64_t, void*, size_t) {
  return hipErrorNotSupported;
}

}  // namespace imcom
