/* s2let_math.h: the reference's header name (s2let_math.h) for ran2_dp, maxerr,
 * maxerr_cplx, binomial_coefficient and the s2let_math_* tiling helpers (declared in
 * s2let_hip.h), and its PI / MIN / MAX / ABS macros. */
#ifndef S2LET_MATH
#define S2LET_MATH
#include "../s2let_hip.h"

#ifndef PI
#define PI 3.141592653589793238462643383279502884197
#endif
#ifndef MIN
#define MIN(a, b) ((a) < (b) ? (a) : (b))
#endif
#ifndef MAX
#define MAX(a, b) ((a) > (b) ? (a) : (b))
#endif
#ifndef ABS
#define ABS(a) ((a) > (0.0) ? (a) : (-(a)))
#endif

#endif
