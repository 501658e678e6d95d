/* Generated by exoblas_amd/generator.py -- do not edit. */
/* Drop-in replacement for the Exo-generated exo_scal.h of the reference
 * (library target EXO_BLAS::exo_scal).  Every function takes host or device
 * pointers and returns once its results are visible to the caller.  */
#ifndef EXOBLAS_AMD_EXO_SCAL_H
#define EXOBLAS_AMD_EXO_SCAL_H

#include "exo_types.h"

#ifdef __cplusplus
extern "C" {
#endif

/* scal: spec src/level1/scal.py:9-12; called from test/level1/scal/exo_scal_wrapper.h:18-20,30-32 */
void exo_sscal_stride_any(void *ctxt, int_fast32_t n, const float *alpha, struct exo_win_1f32 x);
void exo_sscal_stride_1(void *ctxt, int_fast32_t n, const float *alpha, struct exo_win_1f32 x);
void exo_dscal_stride_any(void *ctxt, int_fast32_t n, const double *alpha, struct exo_win_1f64 x);
void exo_dscal_stride_1(void *ctxt, int_fast32_t n, const double *alpha, struct exo_win_1f64 x);

/* scal_alpha_0: spec src/level1/scal.py:15-19; called from test/level1/scal/exo_scal_wrapper.h:15-16,27-28 */
void exo_sscal_alpha_0_stride_any(void *ctxt, int_fast32_t n, struct exo_win_1f32 x);
void exo_sscal_alpha_0_stride_1(void *ctxt, int_fast32_t n, struct exo_win_1f32 x);
void exo_dscal_alpha_0_stride_any(void *ctxt, int_fast32_t n, struct exo_win_1f64 x);
void exo_dscal_alpha_0_stride_1(void *ctxt, int_fast32_t n, struct exo_win_1f64 x);

#ifdef __cplusplus
}
#endif

#endif
