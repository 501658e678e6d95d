/* Generated by exoblas_amd/generator.py -- do not edit. */
/* Drop-in replacement for the Exo-generated exo_gemv.h of the reference
 * (library target EXO_BLAS::exo_gemv).  Every function takes host or device
 * pointers and returns once its results are visible to the caller.  */
#ifndef EXOBLAS_AMD_EXO_GEMV_H
#define EXOBLAS_AMD_EXO_GEMV_H

#include "exo_types.h"

#ifdef __cplusplus
extern "C" {
#endif

/* gemv_rm: spec src/level2/gemv.py:10-22; called from test/level2/gemv/exo_gemv_wrapper.h:19-37 */
void exo_sgemv_rm_stride_any(void *ctxt, int_fast32_t TransA, int_fast32_t M, int_fast32_t N, const float *alpha, const float *beta, struct exo_win_2f32c A, struct exo_win_2f32c AT, struct exo_win_1f32c x, struct exo_win_1f32 y);
void exo_sgemv_rm_stride_1(void *ctxt, int_fast32_t TransA, int_fast32_t M, int_fast32_t N, const float *alpha, const float *beta, struct exo_win_2f32c A, struct exo_win_2f32c AT, struct exo_win_1f32c x, struct exo_win_1f32 y);
void exo_dgemv_rm_stride_any(void *ctxt, int_fast32_t TransA, int_fast32_t M, int_fast32_t N, const double *alpha, const double *beta, struct exo_win_2f64c A, struct exo_win_2f64c AT, struct exo_win_1f64c x, struct exo_win_1f64 y);
void exo_dgemv_rm_stride_1(void *ctxt, int_fast32_t TransA, int_fast32_t M, int_fast32_t N, const double *alpha, const double *beta, struct exo_win_2f64c A, struct exo_win_2f64c AT, struct exo_win_1f64c x, struct exo_win_1f64 y);

#ifdef __cplusplus
}
#endif

#endif
