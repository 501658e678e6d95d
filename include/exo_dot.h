/* Generated by exoblas_amd/generator.py -- do not edit. */
/* Drop-in replacement for the Exo-generated exo_dot.h of the reference
 * (library target EXO_BLAS::exo_dot).  Every function takes host or device
 * pointers and returns once its results are visible to the caller.  */
#ifndef EXOBLAS_AMD_EXO_DOT_H
#define EXOBLAS_AMD_EXO_DOT_H

#include "exo_types.h"

#ifdef __cplusplus
extern "C" {
#endif

/* dot: spec src/level1/dot.py:9-13; called from test/level1/dot/exo_dot_wrapper.h:10-24 */
void exo_sdot_stride_any(void *ctxt, int_fast32_t n, struct exo_win_1f32c x, struct exo_win_1f32c y, float *result);
void exo_sdot_stride_1(void *ctxt, int_fast32_t n, struct exo_win_1f32c x, struct exo_win_1f32c y, float *result);
void exo_ddot_stride_any(void *ctxt, int_fast32_t n, struct exo_win_1f64c x, struct exo_win_1f64c y, double *result);
void exo_ddot_stride_1(void *ctxt, int_fast32_t n, struct exo_win_1f64c x, struct exo_win_1f64c y, double *result);

#ifdef __cplusplus
}
#endif

#endif
