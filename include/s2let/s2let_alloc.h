/* s2let_alloc.h: the reference's header name (s2let_alloc.h) for s2let_allocate_*.
 * The declarations are in s2let_hip.h, the library's single C header. */
#ifndef S2LET_ALLOC
#define S2LET_ALLOC
#include "../s2let_hip.h"
#endif
