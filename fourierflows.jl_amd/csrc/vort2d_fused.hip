// Fused 2D vorticity ETDRK4 step (placeholder until the fused kernels land).
#include "ops.h"

namespace ffhip {

int fused_create(ffhip_op* op) {
  op->fused = nullptr;
  return FFHIP_OK;
}

void fused_destroy(ffhip_op* op) { op->fused = nullptr; }

int fused_etdrk4_run(ffhip_op*, ffhip_etdrk4*, void*, int64_t, hipStream_t) {
  return fail(FFHIP_ERR_UNSUPPORTED, "fused path not built");
}

}  // namespace ffhip
