/* Generated by exoblas_amd/generator.py -- do not edit. */
/* Drop-in replacement for the Exo-generated exo_syr2.h of the reference
 * (library target EXO_BLAS::exo_syr2).  Every function takes host or device
 * pointers and returns once its results are visible to the caller.  */
#ifndef EXOBLAS_AMD_EXO_SYR2_H
#define EXOBLAS_AMD_EXO_SYR2_H

#include "exo_types.h"

#ifdef __cplusplus
extern "C" {
#endif

/* syr2_rm: spec src/level2/syr2.py:9-27; called from test/level2/syr2/exo_syr2_wrapper.h:20-34 */
void exo_ssyr2_rm_stride_any(void *ctxt, int_fast32_t Uplo, int_fast32_t n, const float *alpha, struct exo_win_1f32c x, struct exo_win_1f32c x_copy, struct exo_win_1f32c y, struct exo_win_1f32c y_copy, struct exo_win_2f32 A);
void exo_ssyr2_rm_stride_1(void *ctxt, int_fast32_t Uplo, int_fast32_t n, const float *alpha, struct exo_win_1f32c x, struct exo_win_1f32c x_copy, struct exo_win_1f32c y, struct exo_win_1f32c y_copy, struct exo_win_2f32 A);
void exo_dsyr2_rm_stride_any(void *ctxt, int_fast32_t Uplo, int_fast32_t n, const double *alpha, struct exo_win_1f64c x, struct exo_win_1f64c x_copy, struct exo_win_1f64c y, struct exo_win_1f64c y_copy, struct exo_win_2f64 A);
void exo_dsyr2_rm_stride_1(void *ctxt, int_fast32_t Uplo, int_fast32_t n, const double *alpha, struct exo_win_1f64c x, struct exo_win_1f64c x_copy, struct exo_win_1f64c y, struct exo_win_1f64c y_copy, struct exo_win_2f64 A);

#ifdef __cplusplus
}
#endif

#endif
