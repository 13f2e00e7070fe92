/* s2let_tiling.h: the reference's header name (s2let_tiling.h) for s2let_tiling_*, s2let_j_max, s2let_bandlimit, s2let_L0, s2let_switch_wavtype.
 * The declarations are in s2let_hip.h, the library's single C header. */
#ifndef S2LET_TILING
#define S2LET_TILING
#include "../s2let_hip.h"
#endif
