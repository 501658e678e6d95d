/* Generated by exoblas_amd/generator.py -- do not edit. */
/* Drop-in replacement for the Exo-generated exo_rotm.h of the reference
 * (library target EXO_BLAS::exo_rotm).  Every function takes host or device
 * pointers and returns once its results are visible to the caller.  */
#ifndef EXOBLAS_AMD_EXO_ROTM_H
#define EXOBLAS_AMD_EXO_ROTM_H

#include "exo_types.h"

#ifdef __cplusplus
extern "C" {
#endif

/* rotm_flag_neg_one: spec src/level1/rotm.py:9-15; called from test/level1/rotm/exo_rotm_wrapper.h:10-13,30-33 */
void exo_srotm_flag_neg_one_stride_any(void *ctxt, int_fast32_t n, struct exo_win_1f32 x, struct exo_win_1f32 y, const float *H);
void exo_srotm_flag_neg_one_stride_1(void *ctxt, int_fast32_t n, struct exo_win_1f32 x, struct exo_win_1f32 y, const float *H);
void exo_drotm_flag_neg_one_stride_any(void *ctxt, int_fast32_t n, struct exo_win_1f64 x, struct exo_win_1f64 y, const double *H);
void exo_drotm_flag_neg_one_stride_1(void *ctxt, int_fast32_t n, struct exo_win_1f64 x, struct exo_win_1f64 y, const double *H);

/* rotm_flag_zero: spec src/level1/rotm.py:18-24; called from test/level1/rotm/exo_rotm_wrapper.h:14-17,34-37 */
void exo_srotm_flag_zero_stride_any(void *ctxt, int_fast32_t n, struct exo_win_1f32 x, struct exo_win_1f32 y, const float *H);
void exo_srotm_flag_zero_stride_1(void *ctxt, int_fast32_t n, struct exo_win_1f32 x, struct exo_win_1f32 y, const float *H);
void exo_drotm_flag_zero_stride_any(void *ctxt, int_fast32_t n, struct exo_win_1f64 x, struct exo_win_1f64 y, const double *H);
void exo_drotm_flag_zero_stride_1(void *ctxt, int_fast32_t n, struct exo_win_1f64 x, struct exo_win_1f64 y, const double *H);

/* rotm_flag_one: spec src/level1/rotm.py:27-33; called from test/level1/rotm/exo_rotm_wrapper.h:18-21,38-41 */
void exo_srotm_flag_one_stride_any(void *ctxt, int_fast32_t n, struct exo_win_1f32 x, struct exo_win_1f32 y, const float *H);
void exo_srotm_flag_one_stride_1(void *ctxt, int_fast32_t n, struct exo_win_1f32 x, struct exo_win_1f32 y, const float *H);
void exo_drotm_flag_one_stride_any(void *ctxt, int_fast32_t n, struct exo_win_1f64 x, struct exo_win_1f64 y, const double *H);
void exo_drotm_flag_one_stride_1(void *ctxt, int_fast32_t n, struct exo_win_1f64 x, struct exo_win_1f64 y, const double *H);

#ifdef __cplusplus
}
#endif

#endif
