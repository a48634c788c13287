/*
 * cfd2lcs_inc_sp.h -- C/C++ interface of the MI355X-native libcfd2lcs path,
 * single-precision library variant (link with -lcfd2lcs_sp).
 *
 * Drop-in for the reference header src/cfd2lcs_inc_single.h (identical to
 * the double header except `typedef float lcsdata_t`, cfd2lcs_inc_single.h:17).
 */
#ifndef CFD2LCS_INC_SP_H
#define CFD2LCS_INC_SP_H

#ifndef CFD2LCS_LCSDATA_DEFINED
#define CFD2LCS_LCSDATA_DEFINED
typedef float lcsdata_t;
#endif

#include "cfd2lcs_api.h"

#endif
