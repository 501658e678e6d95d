/* Generated by exoblas_amd/generator.py -- do not edit. */
/* Drop-in replacement for the Exo-generated exo_tbsv.h of the reference
 * (library target EXO_BLAS::exo_tbsv).  Every function takes host or device
 * pointers and returns once its results are visible to the caller.  */
#ifndef EXOBLAS_AMD_EXO_TBSV_H
#define EXOBLAS_AMD_EXO_TBSV_H

#include "exo_types.h"

#ifdef __cplusplus
extern "C" {
#endif

/* tbsv_row_major_Upper_NonTrans: spec src/level2/tbsv.py:12-45; called from test/level2/stbsv/exo_stbsv.h:20-28 */
void exo_stbsv_row_major_Upper_NonTrans_stride_any(void *ctxt, int_fast32_t n, int_fast32_t k, struct exo_win_1f32 x, struct exo_win_2f32c A, int_fast32_t Diag);
void exo_stbsv_row_major_Upper_NonTrans_stride_1(void *ctxt, int_fast32_t n, int_fast32_t k, struct exo_win_1f32 x, struct exo_win_2f32c A, int_fast32_t Diag);
void exo_dtbsv_row_major_Upper_NonTrans_stride_any(void *ctxt, int_fast32_t n, int_fast32_t k, struct exo_win_1f64 x, struct exo_win_2f64c A, int_fast32_t Diag);
void exo_dtbsv_row_major_Upper_NonTrans_stride_1(void *ctxt, int_fast32_t n, int_fast32_t k, struct exo_win_1f64 x, struct exo_win_2f64c A, int_fast32_t Diag);

/* tbsv_row_major_Lower_NonTrans: spec src/level2/tbsv.py:48-81; called from test/level2/stbsv/exo_stbsv.h:39-47 */
void exo_stbsv_row_major_Lower_NonTrans_stride_any(void *ctxt, int_fast32_t n, int_fast32_t k, struct exo_win_1f32 x, struct exo_win_2f32c A, int_fast32_t Diag);
void exo_stbsv_row_major_Lower_NonTrans_stride_1(void *ctxt, int_fast32_t n, int_fast32_t k, struct exo_win_1f32 x, struct exo_win_2f32c A, int_fast32_t Diag);
void exo_dtbsv_row_major_Lower_NonTrans_stride_any(void *ctxt, int_fast32_t n, int_fast32_t k, struct exo_win_1f64 x, struct exo_win_2f64c A, int_fast32_t Diag);
void exo_dtbsv_row_major_Lower_NonTrans_stride_1(void *ctxt, int_fast32_t n, int_fast32_t k, struct exo_win_1f64 x, struct exo_win_2f64c A, int_fast32_t Diag);

/* tbsv_row_major_Upper_Trans: spec src/level2/tbsv.py:84-105; called from test/level2/stbsv/exo_stbsv.h:31-34 */
void exo_stbsv_row_major_Upper_Trans_stride_any(void *ctxt, int_fast32_t n, int_fast32_t k, struct exo_win_1f32 x, struct exo_win_2f32c A, int_fast32_t Diag);
void exo_dtbsv_row_major_Upper_Trans_stride_any(void *ctxt, int_fast32_t n, int_fast32_t k, struct exo_win_1f64 x, struct exo_win_2f64c A, int_fast32_t Diag);

/* tbsv_row_major_Lower_Trans: spec src/level2/tbsv.py:108-129; called from test/level2/stbsv/exo_stbsv.h:50-53 */
void exo_stbsv_row_major_Lower_Trans_stride_any(void *ctxt, int_fast32_t n, int_fast32_t k, struct exo_win_1f32 x, struct exo_win_2f32c A, int_fast32_t Diag);
void exo_dtbsv_row_major_Lower_Trans_stride_any(void *ctxt, int_fast32_t n, int_fast32_t k, struct exo_win_1f64 x, struct exo_win_2f64c A, int_fast32_t Diag);

#ifdef __cplusplus
}
#endif

#endif
