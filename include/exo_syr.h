/* Generated by exoblas_amd/generator.py -- do not edit. */
/* Drop-in replacement for the Exo-generated exo_syr.h of the reference
 * (library target EXO_BLAS::exo_syr).  Every function takes host or device
 * pointers and returns once its results are visible to the caller.  */
#ifndef EXOBLAS_AMD_EXO_SYR_H
#define EXOBLAS_AMD_EXO_SYR_H

#include "exo_types.h"

#ifdef __cplusplus
extern "C" {
#endif

/* syr_rm: spec src/level2/syr.py:10-19; called from test/level2/syr/exo_syr_wrapper.h:17-30 */
void exo_ssyr_rm_stride_any(void *ctxt, int_fast32_t Uplo, int_fast32_t n, const float *alpha, struct exo_win_1f32c x, struct exo_win_1f32c x_copy, struct exo_win_2f32 A);
void exo_ssyr_rm_stride_1(void *ctxt, int_fast32_t Uplo, int_fast32_t n, const float *alpha, struct exo_win_1f32c x, struct exo_win_1f32c x_copy, struct exo_win_2f32 A);
void exo_dsyr_rm_stride_any(void *ctxt, int_fast32_t Uplo, int_fast32_t n, const double *alpha, struct exo_win_1f64c x, struct exo_win_1f64c x_copy, struct exo_win_2f64 A);
void exo_dsyr_rm_stride_1(void *ctxt, int_fast32_t Uplo, int_fast32_t n, const double *alpha, struct exo_win_1f64c x, struct exo_win_1f64c x_copy, struct exo_win_2f64 A);

#ifdef __cplusplus
}
#endif

#endif
