/* Generated by exoblas_amd/generator.py -- do not edit. */
/* Drop-in replacement for the Exo-generated exo_dsdot.h of the reference
 * (library target EXO_BLAS::exo_dsdot).  Every function takes host or device
 * pointers and returns once its results are visible to the caller.  */
#ifndef EXOBLAS_AMD_EXO_DSDOT_H
#define EXOBLAS_AMD_EXO_DSDOT_H

#include "exo_types.h"

#ifdef __cplusplus
extern "C" {
#endif

/* dsdot: spec src/level1/dsdot.py:11-21; called from test/level1/dsdot/exo_dsdot_wrapper.h:9-22 */
void exo_dsdot_stride_any(void *ctxt, int_fast32_t n, struct exo_win_1f32c x, struct exo_win_1f32c y, double *result);
void exo_dsdot_stride_1(void *ctxt, int_fast32_t n, struct exo_win_1f32c x, struct exo_win_1f32c y, double *result);

/* sdsdot: spec src/level1/dsdot.py:24-36; called from test/level1/sdsdot/exo_sdsdot.h:9-23 */
void exo_sdsdot_stride_any(void *ctxt, int_fast32_t n, const float *sb, struct exo_win_1f32c x, struct exo_win_1f32c y, float *result);
void exo_sdsdot_stride_1(void *ctxt, int_fast32_t n, const float *sb, struct exo_win_1f32c x, struct exo_win_1f32c y, float *result);

#ifdef __cplusplus
}
#endif

#endif
