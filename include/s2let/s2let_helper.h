/* s2let_helper.h: the reference's header name (s2let_helper.h) for s2let_n_*.
 * The declarations are in s2let_hip.h, the library's single C header. */
#ifndef S2LET_HELPER
#define S2LET_HELPER
#include "../s2let_hip.h"
#endif
