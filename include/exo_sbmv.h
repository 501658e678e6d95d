/* Generated by exoblas_amd/generator.py -- do not edit. */
/* Drop-in replacement for the Exo-generated exo_sbmv.h of the reference
 * (library target EXO_BLAS::exo_sbmv).  Every function takes host or device
 * pointers and returns once its results are visible to the caller.  */
#ifndef EXOBLAS_AMD_EXO_SBMV_H
#define EXOBLAS_AMD_EXO_SBMV_H

#include "exo_types.h"

#ifdef __cplusplus
extern "C" {
#endif

/* sbmv_scal_y: spec src/level2/sbmv.py:12-16; called from test/level2/ssbmv/exo_ssbmv.h:22-27 */
void exo_ssbmv_scal_y_stride_any(void *ctxt, int_fast32_t n, const float *beta, struct exo_win_1f32 y);
void exo_ssbmv_scal_y_stride_1(void *ctxt, int_fast32_t n, const float *beta, struct exo_win_1f32 y);
void exo_dsbmv_scal_y_stride_any(void *ctxt, int_fast32_t n, const double *beta, struct exo_win_1f64 y);
void exo_dsbmv_scal_y_stride_1(void *ctxt, int_fast32_t n, const double *beta, struct exo_win_1f64 y);

/* sbmv_row_major_Upper: spec src/level2/sbmv.py:19-33; called from test/level2/ssbmv/exo_ssbmv.h:31-39 */
void exo_ssbmv_row_major_Upper_stride_any(void *ctxt, int_fast32_t n, int_fast32_t k, const float *alpha, struct exo_win_2f32c A, struct exo_win_1f32c x, struct exo_win_1f32 y);
void exo_ssbmv_row_major_Upper_stride_1(void *ctxt, int_fast32_t n, int_fast32_t k, const float *alpha, struct exo_win_2f32c A, struct exo_win_1f32c x, struct exo_win_1f32 y);
void exo_dsbmv_row_major_Upper_stride_any(void *ctxt, int_fast32_t n, int_fast32_t k, const double *alpha, struct exo_win_2f64c A, struct exo_win_1f64c x, struct exo_win_1f64 y);
void exo_dsbmv_row_major_Upper_stride_1(void *ctxt, int_fast32_t n, int_fast32_t k, const double *alpha, struct exo_win_2f64c A, struct exo_win_1f64c x, struct exo_win_1f64 y);

/* sbmv_row_major_Lower: spec src/level2/sbmv.py:36-50; called from test/level2/ssbmv/exo_ssbmv.h:43-51 */
void exo_ssbmv_row_major_Lower_stride_any(void *ctxt, int_fast32_t n, int_fast32_t k, const float *alpha, struct exo_win_2f32c A, struct exo_win_1f32c x, struct exo_win_1f32 y);
void exo_ssbmv_row_major_Lower_stride_1(void *ctxt, int_fast32_t n, int_fast32_t k, const float *alpha, struct exo_win_2f32c A, struct exo_win_1f32c x, struct exo_win_1f32 y);
void exo_dsbmv_row_major_Lower_stride_any(void *ctxt, int_fast32_t n, int_fast32_t k, const double *alpha, struct exo_win_2f64c A, struct exo_win_1f64c x, struct exo_win_1f64 y);
void exo_dsbmv_row_major_Lower_stride_1(void *ctxt, int_fast32_t n, int_fast32_t k, const double *alpha, struct exo_win_2f64c A, struct exo_win_1f64c x, struct exo_win_1f64 y);

#ifdef __cplusplus
}
#endif

#endif
