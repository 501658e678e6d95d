/* Generated by exoblas_amd/generator.py -- do not edit. */
/* Drop-in replacement for the Exo-generated exo_axpy.h of the reference
 * (library target EXO_BLAS::exo_axpy).  Every function takes host or device
 * pointers and returns once its results are visible to the caller.  */
#ifndef EXOBLAS_AMD_EXO_AXPY_H
#define EXOBLAS_AMD_EXO_AXPY_H

#include "exo_types.h"

#ifdef __cplusplus
extern "C" {
#endif

/* axpy: spec src/level1/axpy.py:9-12; called from test/level1/axpy/exo_axpy_wrapper.h:17-32 */
void exo_saxpy_stride_any(void *ctxt, int_fast32_t n, const float *alpha, struct exo_win_1f32c x, struct exo_win_1f32 y);
void exo_saxpy_stride_1(void *ctxt, int_fast32_t n, const float *alpha, struct exo_win_1f32c x, struct exo_win_1f32 y);
void exo_daxpy_stride_any(void *ctxt, int_fast32_t n, const double *alpha, struct exo_win_1f64c x, struct exo_win_1f64 y);
void exo_daxpy_stride_1(void *ctxt, int_fast32_t n, const double *alpha, struct exo_win_1f64c x, struct exo_win_1f64 y);

/* axpy_alpha_1: spec src/level1/axpy.py:15-18; called from test/level1/axpy/exo_axpy_wrapper.h:13-15 */
void exo_saxpy_alpha_1_stride_any(void *ctxt, int_fast32_t n, struct exo_win_1f32c x, struct exo_win_1f32 y);
void exo_saxpy_alpha_1_stride_1(void *ctxt, int_fast32_t n, struct exo_win_1f32c x, struct exo_win_1f32 y);
void exo_daxpy_alpha_1_stride_any(void *ctxt, int_fast32_t n, struct exo_win_1f64c x, struct exo_win_1f64 y);
void exo_daxpy_alpha_1_stride_1(void *ctxt, int_fast32_t n, struct exo_win_1f64c x, struct exo_win_1f64 y);

#ifdef __cplusplus
}
#endif

#endif
