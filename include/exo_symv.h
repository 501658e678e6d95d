/* Generated by exoblas_amd/generator.py -- do not edit. */
/* Drop-in replacement for the Exo-generated exo_symv.h of the reference
 * (library target EXO_BLAS::exo_symv).  Every function takes host or device
 * pointers and returns once its results are visible to the caller.  */
#ifndef EXOBLAS_AMD_EXO_SYMV_H
#define EXOBLAS_AMD_EXO_SYMV_H

#include "exo_types.h"

#ifdef __cplusplus
extern "C" {
#endif

/* symv_rm: spec src/level2/symv.py:10-35; called from test/level2/symv/exo_symv_wrapper.h:22-32 */
void exo_ssymv_rm_stride_any(void *ctxt, int_fast32_t Uplo, int_fast32_t n, const float *alpha, struct exo_win_2f32c A, struct exo_win_1f32c x, const float *beta, struct exo_win_1f32 y);
void exo_ssymv_rm_stride_1(void *ctxt, int_fast32_t Uplo, int_fast32_t n, const float *alpha, struct exo_win_2f32c A, struct exo_win_1f32c x, const float *beta, struct exo_win_1f32 y);
void exo_dsymv_rm_stride_any(void *ctxt, int_fast32_t Uplo, int_fast32_t n, const double *alpha, struct exo_win_2f64c A, struct exo_win_1f64c x, const double *beta, struct exo_win_1f64 y);
void exo_dsymv_rm_stride_1(void *ctxt, int_fast32_t Uplo, int_fast32_t n, const double *alpha, struct exo_win_2f64c A, struct exo_win_1f64c x, const double *beta, struct exo_win_1f64 y);

#ifdef __cplusplus
}
#endif

#endif
