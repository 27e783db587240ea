/* FP32 production instantiation: fast concentrations and buffers in float,
 * the large slow pools (NSR, JSR, CSQN-bound, slow troponin) in double;
 * gating, per-CRU currents, reductions and V stay FP64 in strict-FP regions. */
#include "cru_launch.h"

RYR_DEFINE_LAUNCHERS(f32, float, double)
