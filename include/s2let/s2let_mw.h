/* s2let_mw.h: the reference's header name (s2let_mw.h) for s2let_mw_alm2map, s2let_mw_map2alm (+ _real), s2let_mw_power(_real).
 * The declarations are in s2let_hip.h, the library's single C header. */
#ifndef S2LET_MW
#define S2LET_MW
#include "../s2let_hip.h"
#endif
