// Mixed-precision strided GEMM (FP32 and split-BF16 on the matrix cores).
#include "mh_internal.h"

namespace mh {

int gemm_mixed(mh_handle h, int dtype, int prec, const GemmDesc &g) {
  (void)dtype;
  (void)g;
  return fail(h, MH_ERR_UNSUPPORTED, "precision %d is not built yet", prec);
}

}  // namespace mh
