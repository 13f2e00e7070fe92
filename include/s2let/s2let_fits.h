/* s2let_fits.h: the reference's header name (s2let_fits.h) for s2let_fits_*.
 * The declarations are in s2let_hip.h, the library's single C header. */
#ifndef S2LET_FITS
#define S2LET_FITS
#include "../s2let_hip.h"
#endif
