/* Generated by exoblas_amd/generator.py -- do not edit. */
/* Drop-in replacement for the Exo-generated exo_syr2k.h of the reference
 * (library target EXO_BLAS::exo_syr2k).  Every function takes host or device
 * pointers and returns once its results are visible to the caller.  */
#ifndef EXOBLAS_AMD_EXO_SYR2K_H
#define EXOBLAS_AMD_EXO_SYR2K_H

#include "exo_types.h"

#ifdef __cplusplus
extern "C" {
#endif

/* syr2k_rm: spec src/level3/syr2k.py:14-39; called from test/level3/syr2k/exo_syr2k_wrapper.h:21-31 */
void exo_ssyr2k_rm_stride_any(void *ctxt, int_fast32_t Uplo, int_fast32_t Trans, int_fast32_t N, int_fast32_t K, const float *alpha, struct exo_win_2f32c A, struct exo_win_2f32c A_, struct exo_win_2f32c B, struct exo_win_2f32c B_, struct exo_win_2f32c AT, struct exo_win_2f32c AT_, struct exo_win_2f32c BT, struct exo_win_2f32c BT_, struct exo_win_2f32 C);
void exo_ssyr2k_rm_stride_1(void *ctxt, int_fast32_t Uplo, int_fast32_t Trans, int_fast32_t N, int_fast32_t K, const float *alpha, struct exo_win_2f32c A, struct exo_win_2f32c A_, struct exo_win_2f32c B, struct exo_win_2f32c B_, struct exo_win_2f32c AT, struct exo_win_2f32c AT_, struct exo_win_2f32c BT, struct exo_win_2f32c BT_, struct exo_win_2f32 C);
void exo_dsyr2k_rm_stride_any(void *ctxt, int_fast32_t Uplo, int_fast32_t Trans, int_fast32_t N, int_fast32_t K, const double *alpha, struct exo_win_2f64c A, struct exo_win_2f64c A_, struct exo_win_2f64c B, struct exo_win_2f64c B_, struct exo_win_2f64c AT, struct exo_win_2f64c AT_, struct exo_win_2f64c BT, struct exo_win_2f64c BT_, struct exo_win_2f64 C);
void exo_dsyr2k_rm_stride_1(void *ctxt, int_fast32_t Uplo, int_fast32_t Trans, int_fast32_t N, int_fast32_t K, const double *alpha, struct exo_win_2f64c A, struct exo_win_2f64c A_, struct exo_win_2f64c B, struct exo_win_2f64c B_, struct exo_win_2f64c AT, struct exo_win_2f64c AT_, struct exo_win_2f64c BT, struct exo_win_2f64c BT_, struct exo_win_2f64 C);

#ifdef __cplusplus
}
#endif

#endif
