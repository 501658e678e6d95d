/* Generated by exoblas_amd/generator.py -- do not edit. */
/* Drop-in replacement for the Exo-generated exo_trmm.h of the reference
 * (library target EXO_BLAS::exo_trmm).  Every function takes host or device
 * pointers and returns once its results are visible to the caller.  */
#ifndef EXOBLAS_AMD_EXO_TRMM_H
#define EXOBLAS_AMD_EXO_TRMM_H

#include "exo_types.h"

#ifdef __cplusplus
extern "C" {
#endif

/* trmm_row_major_Left_Upper_NonTrans: spec src/level3/trmm.py:12-23; called from test/level3/strmm/exo_strmm.h:18-22 */
void exo_strmm_row_major_Left_Upper_NonTrans(void *ctxt, int_fast32_t m, int_fast32_t n, const float *alpha, struct exo_win_2f32c A, struct exo_win_2f32 B, int_fast32_t Diag);

/* trmm_row_major_Left_Lower_NonTrans: spec src/level3/trmm.py:26-46; called from test/level3/strmm/exo_strmm.h:32-36 */
void exo_strmm_row_major_Left_Lower_NonTrans(void *ctxt, int_fast32_t m, int_fast32_t n, const float *alpha, struct exo_win_2f32c A, struct exo_win_2f32 B, int_fast32_t Diag);

/* trmm_row_major_Left_Upper_Trans: spec src/level3/trmm.py:49-69; called from test/level3/strmm/exo_strmm.h:24-28 */
void exo_strmm_row_major_Left_Upper_Trans(void *ctxt, int_fast32_t m, int_fast32_t n, const float *alpha, struct exo_win_2f32c A, struct exo_win_2f32 B, int_fast32_t Diag);

/* trmm_row_major_Left_Lower_Trans: spec src/level3/trmm.py:72-83; called from test/level3/strmm/exo_strmm.h:38-42 */
void exo_strmm_row_major_Left_Lower_Trans(void *ctxt, int_fast32_t m, int_fast32_t n, const float *alpha, struct exo_win_2f32c A, struct exo_win_2f32 B, int_fast32_t Diag);

#ifdef __cplusplus
}
#endif

#endif
