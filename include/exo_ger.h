/* Generated by exoblas_amd/generator.py -- do not edit. */
/* Drop-in replacement for the Exo-generated exo_ger.h of the reference
 * (library target EXO_BLAS::exo_ger).  Every function takes host or device
 * pointers and returns once its results are visible to the caller.  */
#ifndef EXOBLAS_AMD_EXO_GER_H
#define EXOBLAS_AMD_EXO_GER_H

#include "exo_types.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ger_rm: spec src/level2/ger.py:9-15; called from test/level2/ger/exo_ger_wrapper.h:16-32 */
void exo_sger_rm_stride_any(void *ctxt, int_fast32_t m, int_fast32_t n, const float *alpha, struct exo_win_1f32c x, struct exo_win_1f32c y, struct exo_win_2f32 A);
void exo_sger_rm_stride_1(void *ctxt, int_fast32_t m, int_fast32_t n, const float *alpha, struct exo_win_1f32c x, struct exo_win_1f32c y, struct exo_win_2f32 A);
void exo_dger_rm_stride_any(void *ctxt, int_fast32_t m, int_fast32_t n, const double *alpha, struct exo_win_1f64c x, struct exo_win_1f64c y, struct exo_win_2f64 A);
void exo_dger_rm_stride_1(void *ctxt, int_fast32_t m, int_fast32_t n, const double *alpha, struct exo_win_1f64c x, struct exo_win_1f64c y, struct exo_win_2f64 A);

#ifdef __cplusplus
}
#endif

#endif
