// adjoint.cpp -- forward tape recording for dBatchStepBackward.
#include "odemi_internal.hpp"

namespace odemi {

int tape_begin(BatchImpl *b, int max_steps) {
  (void)b; (void)max_steps;
  report_error("dBatchTapeBegin: not available in this build");
  return 0;
}
int tape_end(BatchImpl *b) { (void)b; return 1; }
int tape_step(BatchImpl *b, bool sync) { (void)b; (void)sync; return 0; }
int tape_backward(BatchImpl *b, int nsteps, const float *grad_out, float *grad_in, float *grad_mu,
                  float *grad_bounce) {
  (void)b; (void)nsteps; (void)grad_out; (void)grad_in; (void)grad_mu; (void)grad_bounce;
  report_error("dBatchStepBackward: not available in this build");
  return 0;
}

}  // namespace odemi
