/*
 * cfd2lcs_inc_dp.h -- C/C++ interface of the MI355X-native libcfd2lcs path,
 * double-precision library variant (link with -lcfd2lcs_dp).
 *
 * Drop-in for the reference header src/cfd2lcs_inc_double.h: the typedef,
 * every flag value and the six C prototypes are identical
 * (cfd2lcs_inc_double.h:17-110, cfd2lcs_c.c:67-142).  The additive entry
 * points of this implementation are declared in cfd2lcs_amd.h, included at
 * the end.
 */
#ifndef CFD2LCS_INC_DP_H
#define CFD2LCS_INC_DP_H

#ifndef CFD2LCS_LCSDATA_DEFINED
#define CFD2LCS_LCSDATA_DEFINED
/* Set the precision here (cfd2lcs_inc_double.h:17) */
typedef double lcsdata_t;
#endif

#include "cfd2lcs_api.h"

#endif
