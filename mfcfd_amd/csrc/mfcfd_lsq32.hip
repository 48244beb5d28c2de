// The mixed path's LSQ kernel (eval_q_derivatives with the stencil extremes qm and, on a
// run's first stage, func_delta) in a unit of its own, compiled with FMA contraction,
// without the IEEE mode bit and with NaN operands assumed absent (-fno-honor-nans
// -mno-amdgpu-ieee).  The extremes qmax, qmin then take one v_max_f32 / v_min_f32 per
// component and edge instead of a compare and a select (the staged walk's edge: 42 -> 34
// VALU instructions).  They differ from the compare-and-select form only in the sign of
// an extreme equal to zero, which the limiter's differences qm - q do not see.  A NaN
// state still stops the run: the residual norm that carries it is formed and tested by
// kernels of the other units.
#include "lskum_kernels.h"

void launch_lsq_f32(const LaunchEnv& e, const V4f* q, V4f* dq, V4f* qm, const double* cspeed, double* delta,
                    bool with_delta, const Ctl* ctl, int lpp, V4f* qsum) {
  launch_lsq_env<float>(e, q, dq, qm, cspeed, delta, with_delta, ctl, lpp, qsum);
}
