/* Generated by exoblas_amd/generator.py -- do not edit. */
/* Drop-in replacement for the Exo-generated exo_syrk.h of the reference
 * (library target EXO_BLAS::exo_syrk).  Every function takes host or device
 * pointers and returns once its results are visible to the caller.  */
#ifndef EXOBLAS_AMD_EXO_SYRK_H
#define EXOBLAS_AMD_EXO_SYRK_H

#include "exo_types.h"

#ifdef __cplusplus
extern "C" {
#endif

/* syrk_rm: spec src/level3/syrk.py:17-32; called from test/level3/syrk/exo_syrk_wrapper.h:21-27 */
void exo_ssyrk_rm_stride_any(void *ctxt, int_fast32_t Uplo, int_fast32_t Trans, int_fast32_t N, int_fast32_t K, const float *alpha, struct exo_win_2f32c A, struct exo_win_2f32c A_, struct exo_win_2f32c AT, struct exo_win_2f32c AT_, struct exo_win_2f32 C);
void exo_ssyrk_rm_stride_1(void *ctxt, int_fast32_t Uplo, int_fast32_t Trans, int_fast32_t N, int_fast32_t K, const float *alpha, struct exo_win_2f32c A, struct exo_win_2f32c A_, struct exo_win_2f32c AT, struct exo_win_2f32c AT_, struct exo_win_2f32 C);
void exo_dsyrk_rm_stride_any(void *ctxt, int_fast32_t Uplo, int_fast32_t Trans, int_fast32_t N, int_fast32_t K, const double *alpha, struct exo_win_2f64c A, struct exo_win_2f64c A_, struct exo_win_2f64c AT, struct exo_win_2f64c AT_, struct exo_win_2f64 C);
void exo_dsyrk_rm_stride_1(void *ctxt, int_fast32_t Uplo, int_fast32_t Trans, int_fast32_t N, int_fast32_t K, const double *alpha, struct exo_win_2f64c A, struct exo_win_2f64c A_, struct exo_win_2f64c AT, struct exo_win_2f64c AT_, struct exo_win_2f64 C);

#ifdef __cplusplus
}
#endif

#endif
