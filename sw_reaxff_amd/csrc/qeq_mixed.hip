// qeq_mixed.hip — mixed-precision QEq solve (FP32 H and inner CG inside an
// FP64 iterative refinement).
#include "swr_internal.h"

int swr_qeq_mixed_solve(swr_system *S) {
  (void)S;
  swr_set_error_msg("mixed-precision QEq is not available in this build");
  return SWR_ERR_STATE;
}
