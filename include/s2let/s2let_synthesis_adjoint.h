/* s2let_synthesis_adjoint.h: the reference's header name (s2let_synthesis_adjoint.h) for s2let_synthesis_adjoint_*.
 * The declarations are in s2let_hip.h, the library's single C header. */
#ifndef S2LET_SYNTHESIS_ADJOINT
#define S2LET_SYNTHESIS_ADJOINT
#include "../s2let_hip.h"
#endif
