/* Generated by exoblas_amd/generator.py -- do not edit. */
/* Drop-in replacement for the Exo-generated exo_mscal.h of the reference
 * (library target EXO_BLAS::exo_mscal).  Every function takes host or device
 * pointers and returns once its results are visible to the caller.  */
#ifndef EXOBLAS_AMD_EXO_MSCAL_H
#define EXOBLAS_AMD_EXO_MSCAL_H

#include "exo_types.h"

#ifdef __cplusplus
extern "C" {
#endif

/* mscal_rm: spec src/level3/mscal.py:12-17; called from test/level3/gemm/exo_gemm_wrapper.h:18-20 */
void exo_smscal_rm_stride_any(void *ctxt, int_fast32_t M, int_fast32_t N, const float *alpha, struct exo_win_2f32 A);
void exo_smscal_rm_stride_1(void *ctxt, int_fast32_t M, int_fast32_t N, const float *alpha, struct exo_win_2f32 A);
void exo_dmscal_rm_stride_any(void *ctxt, int_fast32_t M, int_fast32_t N, const double *alpha, struct exo_win_2f64 A);
void exo_dmscal_rm_stride_1(void *ctxt, int_fast32_t M, int_fast32_t N, const double *alpha, struct exo_win_2f64 A);

/* trmscal_rm: spec src/level3/mscal.py:20-27; called from test/level3/syrk/exo_syrk_wrapper.h:18-20 */
void exo_strmscal_rm_stride_any(void *ctxt, int_fast32_t Uplo, int_fast32_t N, const float *alpha, struct exo_win_2f32 A);
void exo_strmscal_rm_stride_1(void *ctxt, int_fast32_t Uplo, int_fast32_t N, const float *alpha, struct exo_win_2f32 A);
void exo_dtrmscal_rm_stride_any(void *ctxt, int_fast32_t Uplo, int_fast32_t N, const double *alpha, struct exo_win_2f64 A);
void exo_dtrmscal_rm_stride_1(void *ctxt, int_fast32_t Uplo, int_fast32_t N, const double *alpha, struct exo_win_2f64 A);

#ifdef __cplusplus
}
#endif

#endif
