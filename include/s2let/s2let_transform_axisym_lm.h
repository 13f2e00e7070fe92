/* s2let_transform_axisym_lm.h: the reference's header name (s2let_transform_axisym_lm.h) for s2let_transform_axisym_lm_*.
 * The declarations are in s2let_hip.h, the library's single C header. */
#ifndef S2LET_TRANSFORM_AXISYM_LM
#define S2LET_TRANSFORM_AXISYM_LM
#include "../s2let_hip.h"
#endif
