/* Generated by exoblas_amd/generator.py -- do not edit. */
/* Drop-in replacement for the Exo-generated exo_rot.h of the reference
 * (library target EXO_BLAS::exo_rot).  Every function takes host or device
 * pointers and returns once its results are visible to the caller.  */
#ifndef EXOBLAS_AMD_EXO_ROT_H
#define EXOBLAS_AMD_EXO_ROT_H

#include "exo_types.h"

#ifdef __cplusplus
extern "C" {
#endif

/* rot: spec src/level1/rot.py:9-15; called from test/level1/rot/exo_rot_wrapper.h:9-21 */
void exo_srot_stride_any(void *ctxt, int_fast32_t n, struct exo_win_1f32 x, struct exo_win_1f32 y, const float *c, const float *s);
void exo_srot_stride_1(void *ctxt, int_fast32_t n, struct exo_win_1f32 x, struct exo_win_1f32 y, const float *c, const float *s);
void exo_drot_stride_any(void *ctxt, int_fast32_t n, struct exo_win_1f64 x, struct exo_win_1f64 y, const double *c, const double *s);
void exo_drot_stride_1(void *ctxt, int_fast32_t n, struct exo_win_1f64 x, struct exo_win_1f64 y, const double *c, const double *s);

#ifdef __cplusplus
}
#endif

#endif
