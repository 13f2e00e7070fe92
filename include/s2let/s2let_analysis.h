/* s2let_analysis.h: the reference's header name (s2let_analysis.h) for s2let_analysis_*.
 * The declarations are in s2let_hip.h, the library's single C header. */
#ifndef S2LET_ANALYSIS
#define S2LET_ANALYSIS
#include "../s2let_hip.h"
#endif
