/* s2let_synthesis.h: the reference's header name (s2let_synthesis.h) for s2let_synthesis_*.
 * The declarations are in s2let_hip.h, the library's single C header. */
#ifndef S2LET_SYNTHESIS
#define S2LET_SYNTHESIS
#include "../s2let_hip.h"
#endif
