/* FP64 instantiation: every state field in double, contraction off, so the
 * CRU update reproduces the FP64 oracle bit for bit. */
#include "cru_launch.h"

RYR_DEFINE_LAUNCHERS(f64, double, double)
