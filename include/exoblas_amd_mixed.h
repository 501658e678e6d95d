/* exoblas_amd: the BF16-update mixed strsv on a copy packed once.
 *
 * No counterpart in the reference.  exoblas_strsv_mixed(_async) with
 * update = 16 converts the whole n x n matrix to BF16 on every call; these two
 * entry points are for callers that solve many right-hand sides (or the same
 * one again) with one matrix: pack once, then solve with the packed copy.
 *
 * exoblas_strsv_mixed_pack_async writes the BF16 copy of A (n x n, row
 * stride n: 2 n^2 bytes of device memory the caller owns, at `packed`).
 * exoblas_strsv_mixed_packed_async is exoblas_strsv_mixed_async with
 * update = 16, reading the off-diagonal blocks from `packed`, which must hold
 * the copy of this A (unchanged since it was packed); A itself still serves
 * the FP64 diagonal-block solves and the FP64 residual.  Device pointers,
 * stream-ordered; windows and checks as exoblas_strsv_mixed_async
 * (include/exoblas_amd_ext.h), which solves the same system with the same
 * kernels, so the results are the same bits.  Returns 0 or an
 * EXOBLAS_STATUS_* code (exoblas_amd.h). */
#ifndef EXOBLAS_AMD_MIXED_H
#define EXOBLAS_AMD_MIXED_H

#include <hip/hip_runtime_api.h>
#include "exo_types.h"

#ifdef __cplusplus
extern "C" {
#endif

int exoblas_strsv_mixed_pack_async(hipStream_t stream, int_fast32_t n, struct exo_win_2f32c A, void *packed);
int exoblas_strsv_mixed_packed_async(hipStream_t stream, int_fast32_t Uplo, int_fast32_t TransA, int_fast32_t Diag,
                                     int_fast32_t n, struct exo_win_1f32 x, struct exo_win_2f32c A,
                                     const void *packed, int_fast32_t refine);

#ifdef __cplusplus
}
#endif

#endif /* EXOBLAS_AMD_MIXED_H */
