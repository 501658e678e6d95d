/* Generated by exoblas_amd/generator.py -- do not edit. */
/* Drop-in replacement for the Exo-generated exo_copy.h of the reference
 * (library target EXO_BLAS::exo_copy).  Every function takes host or device
 * pointers and returns once its results are visible to the caller.  */
#ifndef EXOBLAS_AMD_EXO_COPY_H
#define EXOBLAS_AMD_EXO_COPY_H

#include "exo_types.h"

#ifdef __cplusplus
extern "C" {
#endif

/* copy: spec src/level1/exo_copy.py:9-12; called from test/level1/copy/exo_copy_wrapper.h:9-21 */
void exo_scopy_stride_any(void *ctxt, int_fast32_t n, struct exo_win_1f32c x, struct exo_win_1f32 y);
void exo_scopy_stride_1(void *ctxt, int_fast32_t n, struct exo_win_1f32c x, struct exo_win_1f32 y);
void exo_dcopy_stride_any(void *ctxt, int_fast32_t n, struct exo_win_1f64c x, struct exo_win_1f64 y);
void exo_dcopy_stride_1(void *ctxt, int_fast32_t n, struct exo_win_1f64c x, struct exo_win_1f64 y);

#ifdef __cplusplus
}
#endif

#endif
