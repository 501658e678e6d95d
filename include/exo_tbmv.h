/* Generated by exoblas_amd/generator.py -- do not edit. */
/* Drop-in replacement for the Exo-generated exo_tbmv.h of the reference
 * (library target EXO_BLAS::exo_tbmv).  Every function takes host or device
 * pointers and returns once its results are visible to the caller.  */
#ifndef EXOBLAS_AMD_EXO_TBMV_H
#define EXOBLAS_AMD_EXO_TBMV_H

#include "exo_types.h"

#ifdef __cplusplus
extern "C" {
#endif

/* tbmv_row_major_Upper_NonTrans: spec src/level2/tbmv.py:12-43; called from test/level2/stbmv/exo_stbmv.h:20-28 */
void exo_stbmv_row_major_Upper_NonTrans_stride_any(void *ctxt, int_fast32_t n, int_fast32_t k, struct exo_win_1f32 x, struct exo_win_2f32c A, int_fast32_t Diag);
void exo_stbmv_row_major_Upper_NonTrans_stride_1(void *ctxt, int_fast32_t n, int_fast32_t k, struct exo_win_1f32 x, struct exo_win_2f32c A, int_fast32_t Diag);
void exo_dtbmv_row_major_Upper_NonTrans_stride_any(void *ctxt, int_fast32_t n, int_fast32_t k, struct exo_win_1f64 x, struct exo_win_2f64c A, int_fast32_t Diag);
void exo_dtbmv_row_major_Upper_NonTrans_stride_1(void *ctxt, int_fast32_t n, int_fast32_t k, struct exo_win_1f64 x, struct exo_win_2f64c A, int_fast32_t Diag);

/* tbmv_row_major_Lower_NonTrans: spec src/level2/tbmv.py:46-77; called from test/level2/stbmv/exo_stbmv.h:46-54 */
void exo_stbmv_row_major_Lower_NonTrans_stride_any(void *ctxt, int_fast32_t n, int_fast32_t k, struct exo_win_1f32 x, struct exo_win_2f32c A, int_fast32_t Diag);
void exo_stbmv_row_major_Lower_NonTrans_stride_1(void *ctxt, int_fast32_t n, int_fast32_t k, struct exo_win_1f32 x, struct exo_win_2f32c A, int_fast32_t Diag);
void exo_dtbmv_row_major_Lower_NonTrans_stride_any(void *ctxt, int_fast32_t n, int_fast32_t k, struct exo_win_1f64 x, struct exo_win_2f64c A, int_fast32_t Diag);
void exo_dtbmv_row_major_Lower_NonTrans_stride_1(void *ctxt, int_fast32_t n, int_fast32_t k, struct exo_win_1f64 x, struct exo_win_2f64c A, int_fast32_t Diag);

/* tbmv_row_major_Upper_Trans: spec src/level2/tbmv.py:80-108; called from test/level2/stbmv/exo_stbmv.h:32-40 */
void exo_stbmv_row_major_Upper_Trans_stride_any(void *ctxt, int_fast32_t n, int_fast32_t k, struct exo_win_1f32 x, struct exo_win_2f32c A, int_fast32_t Diag);
void exo_stbmv_row_major_Upper_Trans_stride_1(void *ctxt, int_fast32_t n, int_fast32_t k, struct exo_win_1f32 x, struct exo_win_2f32c A, int_fast32_t Diag);
void exo_dtbmv_row_major_Upper_Trans_stride_any(void *ctxt, int_fast32_t n, int_fast32_t k, struct exo_win_1f64 x, struct exo_win_2f64c A, int_fast32_t Diag);
void exo_dtbmv_row_major_Upper_Trans_stride_1(void *ctxt, int_fast32_t n, int_fast32_t k, struct exo_win_1f64 x, struct exo_win_2f64c A, int_fast32_t Diag);

/* tbmv_row_major_Lower_Trans: spec src/level2/tbmv.py:111-141; called from test/level2/stbmv/exo_stbmv.h:58-66 */
void exo_stbmv_row_major_Lower_Trans_stride_any(void *ctxt, int_fast32_t n, int_fast32_t k, struct exo_win_1f32 x, struct exo_win_2f32c A, int_fast32_t Diag);
void exo_stbmv_row_major_Lower_Trans_stride_1(void *ctxt, int_fast32_t n, int_fast32_t k, struct exo_win_1f32 x, struct exo_win_2f32c A, int_fast32_t Diag);
void exo_dtbmv_row_major_Lower_Trans_stride_any(void *ctxt, int_fast32_t n, int_fast32_t k, struct exo_win_1f64 x, struct exo_win_2f64c A, int_fast32_t Diag);
void exo_dtbmv_row_major_Lower_Trans_stride_1(void *ctxt, int_fast32_t n, int_fast32_t k, struct exo_win_1f64 x, struct exo_win_2f64c A, int_fast32_t Diag);

#ifdef __cplusplus
}
#endif

#endif
