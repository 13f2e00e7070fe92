/* s2let_analysis_adjoint.h: the reference's header name (s2let_analysis_adjoint.h) for s2let_analysis_adjoint_*.
 * The declarations are in s2let_hip.h, the library's single C header. */
#ifndef S2LET_ANALYSIS_ADJOINT
#define S2LET_ANALYSIS_ADJOINT
#include "../s2let_hip.h"
#endif
