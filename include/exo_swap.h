/* Generated by exoblas_amd/generator.py -- do not edit. */
/* Drop-in replacement for the Exo-generated exo_swap.h of the reference
 * (library target EXO_BLAS::exo_swap).  Every function takes host or device
 * pointers and returns once its results are visible to the caller.  */
#ifndef EXOBLAS_AMD_EXO_SWAP_H
#define EXOBLAS_AMD_EXO_SWAP_H

#include "exo_types.h"

#ifdef __cplusplus
extern "C" {
#endif

/* swap: spec src/level1/swap.py:9-15; called from test/level1/swap/exo_swap_wrapper.h:9-21 */
void exo_sswap_stride_any(void *ctxt, int_fast32_t n, struct exo_win_1f32 x, struct exo_win_1f32 y);
void exo_sswap_stride_1(void *ctxt, int_fast32_t n, struct exo_win_1f32 x, struct exo_win_1f32 y);
void exo_dswap_stride_any(void *ctxt, int_fast32_t n, struct exo_win_1f64 x, struct exo_win_1f64 y);
void exo_dswap_stride_1(void *ctxt, int_fast32_t n, struct exo_win_1f64 x, struct exo_win_1f64 y);

#ifdef __cplusplus
}
#endif

#endif
