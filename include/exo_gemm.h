/* Generated by exoblas_amd/generator.py -- do not edit. */
/* Drop-in replacement for the Exo-generated exo_gemm.h of the reference
 * (library target EXO_BLAS::exo_gemm).  Every function takes host or device
 * pointers and returns once its results are visible to the caller.  */
#ifndef EXOBLAS_AMD_EXO_GEMM_H
#define EXOBLAS_AMD_EXO_GEMM_H

#include "exo_types.h"

#ifdef __cplusplus
extern "C" {
#endif

/* gemm: spec src/level3/gemm.py:17-47; called from test/level3/gemm/exo_gemm_wrapper.h:21-27 */
void exo_sgemm_stride_any(void *ctxt, int_fast32_t TransA, int_fast32_t TransB, int_fast32_t M, int_fast32_t N, int_fast32_t K, const float *alpha, struct exo_win_2f32c A, struct exo_win_2f32c B, struct exo_win_2f32c AT, struct exo_win_2f32c BT, struct exo_win_2f32 C);
void exo_sgemm_stride_1(void *ctxt, int_fast32_t TransA, int_fast32_t TransB, int_fast32_t M, int_fast32_t N, int_fast32_t K, const float *alpha, struct exo_win_2f32c A, struct exo_win_2f32c B, struct exo_win_2f32c AT, struct exo_win_2f32c BT, struct exo_win_2f32 C);
void exo_dgemm_stride_any(void *ctxt, int_fast32_t TransA, int_fast32_t TransB, int_fast32_t M, int_fast32_t N, int_fast32_t K, const double *alpha, struct exo_win_2f64c A, struct exo_win_2f64c B, struct exo_win_2f64c AT, struct exo_win_2f64c BT, struct exo_win_2f64 C);
void exo_dgemm_stride_1(void *ctxt, int_fast32_t TransA, int_fast32_t TransB, int_fast32_t M, int_fast32_t N, int_fast32_t K, const double *alpha, struct exo_win_2f64c A, struct exo_win_2f64c B, struct exo_win_2f64c AT, struct exo_win_2f64c BT, struct exo_win_2f64 C);

#ifdef __cplusplus
}
#endif

#endif
