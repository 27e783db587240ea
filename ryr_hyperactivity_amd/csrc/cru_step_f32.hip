/* FP32 production instantiation: every concentration and buffer in float;
 * the slow pools (NSR, JSR, CSQN-bound, slow troponin) are updated in double
 * and rounded stochastically to float; gating, per-CRU currents, reductions
 * and V stay FP64.  The compiler may fuse a*b+c anywhere here except in the
 * sub-voxel sweep (RYR_EXPR_FP), so the strict-FP regions the exact path
 * relies on are left empty in this unit. */
#define RYR_CONTRACT_FAST
#include "cru_launch.h"

RYR_DEFINE_LAUNCHERS(f32, float, float)
