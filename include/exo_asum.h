/* Generated by exoblas_amd/generator.py -- do not edit. */
/* Drop-in replacement for the Exo-generated exo_asum.h of the reference
 * (library target EXO_BLAS::exo_asum).  Every function takes host or device
 * pointers and returns once its results are visible to the caller.  */
#ifndef EXOBLAS_AMD_EXO_ASUM_H
#define EXOBLAS_AMD_EXO_ASUM_H

#include "exo_types.h"

#ifdef __cplusplus
extern "C" {
#endif

/* asum: spec src/level1/asum.py:11-15; called from test/level1/asum/exo_asum_wrapper.h:9-20 */
void exo_sasum_stride_any(void *ctxt, int_fast32_t n, struct exo_win_1f32c x, float *result);
void exo_sasum_stride_1(void *ctxt, int_fast32_t n, struct exo_win_1f32c x, float *result);
void exo_dasum_stride_any(void *ctxt, int_fast32_t n, struct exo_win_1f64c x, double *result);
void exo_dasum_stride_1(void *ctxt, int_fast32_t n, struct exo_win_1f64c x, double *result);

#ifdef __cplusplus
}
#endif

#endif
