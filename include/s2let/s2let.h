/* s2let.h: the reference's umbrella header (s2let.h:1-26) under its own name, so a
 * caller written against the reference builds unchanged with -I<repo>/include
 * ("s2let/s2let.h") or -I<repo>/include/s2let ("s2let.h") and links
 * libs2let_amd.so in place of libs2let, libso3, libssht and fftw3.  The HEALPix,
 * IDL and MATLAB headers of the reference are not part of this library. */
#ifndef S2LET
#define S2LET

#include "../s2let_hip.h"
#include "s2let_alloc.h"
#include "s2let_analysis.h"
#include "s2let_analysis_adjoint.h"
#include "s2let_error.h"
#include "s2let_fits.h"
#include "s2let_helper.h"
#include "s2let_lm.h"
#include "s2let_math.h"
#include "s2let_mw.h"
#include "s2let_so3.h"
#include "s2let_synthesis.h"
#include "s2let_synthesis_adjoint.h"
#include "s2let_tiling.h"
#include "s2let_transform_axisym_lm.h"
#include "s2let_transform_axisym_mw.h"
#include "s2let_types.h"

#endif
