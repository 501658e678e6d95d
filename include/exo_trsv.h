/* Generated by exoblas_amd/generator.py -- do not edit. */
/* Drop-in replacement for the Exo-generated exo_trsv.h of the reference
 * (library target EXO_BLAS::exo_trsv).  Every function takes host or device
 * pointers and returns once its results are visible to the caller.  */
#ifndef EXOBLAS_AMD_EXO_TRSV_H
#define EXOBLAS_AMD_EXO_TRSV_H

#include "exo_types.h"

#ifdef __cplusplus
extern "C" {
#endif

/* trsv_rm: spec src/level2/trsv.py:10-52; called from test/level2/trsv/exo_trsv_wrapper.h:19-27 */
void exo_strsv_rm_stride_any(void *ctxt, int_fast32_t Uplo, int_fast32_t TransA, int_fast32_t Diag, int_fast32_t n, struct exo_win_1f32 x, struct exo_win_2f32c A);
void exo_strsv_rm_stride_1(void *ctxt, int_fast32_t Uplo, int_fast32_t TransA, int_fast32_t Diag, int_fast32_t n, struct exo_win_1f32 x, struct exo_win_2f32c A);
void exo_dtrsv_rm_stride_any(void *ctxt, int_fast32_t Uplo, int_fast32_t TransA, int_fast32_t Diag, int_fast32_t n, struct exo_win_1f64 x, struct exo_win_2f64c A);
void exo_dtrsv_rm_stride_1(void *ctxt, int_fast32_t Uplo, int_fast32_t TransA, int_fast32_t Diag, int_fast32_t n, struct exo_win_1f64 x, struct exo_win_2f64c A);

#ifdef __cplusplus
}
#endif

#endif
