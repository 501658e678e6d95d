/* Generated by exoblas_amd/generator.py -- do not edit. */
/* Drop-in replacement for the Exo-generated exo_symm.h of the reference
 * (library target EXO_BLAS::exo_symm).  Every function takes host or device
 * pointers and returns once its results are visible to the caller.  */
#ifndef EXOBLAS_AMD_EXO_SYMM_H
#define EXOBLAS_AMD_EXO_SYMM_H

#include "exo_types.h"

#ifdef __cplusplus
extern "C" {
#endif

/* symm_rm: spec src/level3/symm.py:16-47; called from test/level3/symm/exo_symm_wrapper.h:21-26 */
void exo_ssymm_rm_stride_any(void *ctxt, int_fast32_t Side, int_fast32_t Uplo, int_fast32_t M, int_fast32_t N, const float *alpha, struct exo_win_2f32c Aleft, struct exo_win_2f32c Aright, struct exo_win_2f32c B, struct exo_win_2f32 C);
void exo_ssymm_rm_stride_1(void *ctxt, int_fast32_t Side, int_fast32_t Uplo, int_fast32_t M, int_fast32_t N, const float *alpha, struct exo_win_2f32c Aleft, struct exo_win_2f32c Aright, struct exo_win_2f32c B, struct exo_win_2f32 C);
void exo_dsymm_rm_stride_any(void *ctxt, int_fast32_t Side, int_fast32_t Uplo, int_fast32_t M, int_fast32_t N, const double *alpha, struct exo_win_2f64c Aleft, struct exo_win_2f64c Aright, struct exo_win_2f64c B, struct exo_win_2f64 C);
void exo_dsymm_rm_stride_1(void *ctxt, int_fast32_t Side, int_fast32_t Uplo, int_fast32_t M, int_fast32_t N, const double *alpha, struct exo_win_2f64c Aleft, struct exo_win_2f64c Aright, struct exo_win_2f64c B, struct exo_win_2f64 C);

#ifdef __cplusplus
}
#endif

#endif
