/* FP32 production instantiation: every concentration and buffer in float;
 * the slow pools (NSR, JSR, CSQN-bound, slow troponin) are updated in double
 * and rounded stochastically to float; gating, per-CRU currents, reductions
 * and V stay FP64 in strict-FP regions. */
#include "cru_launch.h"

RYR_DEFINE_LAUNCHERS(f32, float, float)
