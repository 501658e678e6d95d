/* Generated by exoblas_amd/generator.py -- do not edit. */
/* Drop-in replacement for the Exo-generated exo_gbmv.h of the reference
 * (library target EXO_BLAS::exo_gbmv).  Every function takes host or device
 * pointers and returns once its results are visible to the caller.  */
#ifndef EXOBLAS_AMD_EXO_GBMV_H
#define EXOBLAS_AMD_EXO_GBMV_H

#include "exo_types.h"

#ifdef __cplusplus
extern "C" {
#endif

/* gbmv_row_major_NonTrans: spec src/level2/gbmv.py:13-52; called from test/level2/sgbmv/exo_sgbmv.h:9-19 */
void exo_sgbmv_row_major_NonTrans_stride_any(void *ctxt, int_fast32_t m, int_fast32_t n, int_fast32_t kl, int_fast32_t ku, const float *alpha, const float *beta, struct exo_win_2f32c a, struct exo_win_1f32c x, struct exo_win_1f32 y);
void exo_sgbmv_row_major_NonTrans_stride_1(void *ctxt, int_fast32_t m, int_fast32_t n, int_fast32_t kl, int_fast32_t ku, const float *alpha, const float *beta, struct exo_win_2f32c a, struct exo_win_1f32c x, struct exo_win_1f32 y);
void exo_dgbmv_row_major_NonTrans_stride_any(void *ctxt, int_fast32_t m, int_fast32_t n, int_fast32_t kl, int_fast32_t ku, const double *alpha, const double *beta, struct exo_win_2f64c a, struct exo_win_1f64c x, struct exo_win_1f64 y);
void exo_dgbmv_row_major_NonTrans_stride_1(void *ctxt, int_fast32_t m, int_fast32_t n, int_fast32_t kl, int_fast32_t ku, const double *alpha, const double *beta, struct exo_win_2f64c a, struct exo_win_1f64c x, struct exo_win_1f64 y);

#ifdef __cplusplus
}
#endif

#endif
