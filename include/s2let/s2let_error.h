/* s2let_error.h: the reference's error macros (s2let_error.h:13-26).  The library's
 * own entry points report through s2let_hip_last_error(); these are for callers
 * written against the reference, with its behaviour: print the message and where
 * it occurred, then exit(1). */
#ifndef S2LET_ERROR
#define S2LET_ERROR

#include <stdio.h>
#include <stdlib.h>

/* a block, so that it can follow a single-line if */
#define S2LET_ERROR_GENERIC(comment)                                                   \
  {                                                                                  \
    printf("ERROR: %s.\n", comment);                                                  \
    printf("ERROR: %s <%s> %s %s %s %d.\n", "Occurred in function", __func__, "of file", \
           __FILE__, "on line", __LINE__);                                           \
    exit(1);                                                                         \
  }

#define S2LET_ERROR_MEM_ALLOC_CHECK(pointer) \
  if ((pointer) == NULL) {                   \
    S2LET_ERROR_GENERIC("Memory allocation failed") \
  }

#endif
