/* s2let_types.h: the reference's header name (s2let_types.h:18-137) for s2let_parameters_t, s2let_sampling_t, ssht_dl_method_t.
 * The declarations are in s2let_hip.h, the library's single C header. */
#ifndef S2LET_TYPES
#define S2LET_TYPES
#include "../s2let_hip.h"
#endif
