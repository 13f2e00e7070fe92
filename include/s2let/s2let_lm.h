/* s2let_lm.h: the reference's header name (s2let_lm.h) for s2let_lm_power, s2let_lm_random_flm(_real).
 * The declarations are in s2let_hip.h, the library's single C header. */
#ifndef S2LET_LM
#define S2LET_LM
#include "../s2let_hip.h"
#endif
