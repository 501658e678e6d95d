/* exoblas_amd: an entry point that allocates nothing.
 *
 * exoblas_sdot_noalloc_async has no counterpart in the reference (whose
 * exo_sdot_stride_any, src/level1/dot.py:9-13, is the computation it
 * performs).  It is for callers that cannot allocate device memory: code
 * under hipGraph capture, tight serving loops, harnesses that forbid
 * hipMalloc.  It computes *result = sum_i x_i y_i like
 * exoblas_sdot_stride_any_async (same windows: device pointers at element 0,
 * any stride; the same kernel and launch shape, so the same bits), but its
 * scratch is a fixed 16,400-byte array in the library's device memory instead
 * of the stream's buffers, and it calls no hipMalloc, hipFree or other
 * allocating routine.  Limit: the reduction's grid may use at most 2048
 * partial slots (two workgroups per CU on devices of up to 1024 CUs); a call
 * that would need more is rejected with EXOBLAS_STATUS_INVALID_SIZE.
 * Every call on a device shares that device's scratch and the last call's
 * contents stay there: calls that overlap on one device (different streams,
 * or graph replays running at once) race, so serialise them.  A bounded wait
 * that expires is recorded in the scratch only, not reported.  Returns 0 or
 * an EXOBLAS_STATUS_* code (exoblas_amd.h). */
#ifndef EXOBLAS_AMD_NOALLOC_H
#define EXOBLAS_AMD_NOALLOC_H

#include <hip/hip_runtime_api.h>
#include "exo_types.h"

#ifdef __cplusplus
extern "C" {
#endif

int exoblas_sdot_noalloc_async(hipStream_t stream, int_fast32_t n, struct exo_win_1f32c x, struct exo_win_1f32c y,
                               float *result);

#ifdef __cplusplus
}
#endif

#endif /* EXOBLAS_AMD_NOALLOC_H */
