/* s2let_transform_axisym_mw.h: the reference's header name (s2let_transform_axisym_mw.h) for s2let_transform_axisym_*_mw*.
 * The declarations are in s2let_hip.h, the library's single C header. */
#ifndef S2LET_TRANSFORM_AXISYM_MW
#define S2LET_TRANSFORM_AXISYM_MW
#include "../s2let_hip.h"
#endif
