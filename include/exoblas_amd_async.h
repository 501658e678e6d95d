/* Generated by exoblas_amd/generator.py -- do not edit. */
/* Device-pointer, stream-ordered forms of every exo_* entry point.
 * Same arguments as the exo_* function of the same name, with the
 * hipStream_t the work is enqueued on in place of `void *ctxt`.  All window
 * data and result pointers must be device (or managed) memory; scalar
 * inputs are read on the host before the call returns.  Returns 0 on
 * success, otherwise an EXOBLAS_STATUS_* code (exoblas_amd.h).
 * Values that workgroups exchange through memory (reductions, split
 * gemv, trsv) carry a marker drawn from a per-stream counter on the
 * device, so a call captured in a hipGraph replays correctly.  A call
 * that needs more scratch than its stream holds allocates a larger
 * buffer (without synchronising; the old one stays valid for work
 * already queued or captured); during capture such growth is refused
 * with EXOBLAS_STATUS_NOT_SUPPORTED, so make the largest call once on
 * the stream before capturing.  Calls on one stream, and replays of
 * graphs captured on it, share its counters: serialise them.  Waits
 * between workgroups are bounded; exoblas_stream_check reports one
 * that expired (exoblas_amd.h). */
#ifndef EXOBLAS_AMD_ASYNC_H
#define EXOBLAS_AMD_ASYNC_H

#include <hip/hip_runtime_api.h>
#include "exo_types.h"

#ifdef __cplusplus
extern "C" {
#endif

int exoblas_sasum_stride_any_async(hipStream_t stream, int_fast32_t n, struct exo_win_1f32c x, float *result);
int exoblas_sasum_stride_1_async(hipStream_t stream, int_fast32_t n, struct exo_win_1f32c x, float *result);
int exoblas_dasum_stride_any_async(hipStream_t stream, int_fast32_t n, struct exo_win_1f64c x, double *result);
int exoblas_dasum_stride_1_async(hipStream_t stream, int_fast32_t n, struct exo_win_1f64c x, double *result);
int exoblas_saxpy_stride_any_async(hipStream_t stream, int_fast32_t n, const float *alpha, struct exo_win_1f32c x, struct exo_win_1f32 y);
int exoblas_saxpy_stride_1_async(hipStream_t stream, int_fast32_t n, const float *alpha, struct exo_win_1f32c x, struct exo_win_1f32 y);
int exoblas_daxpy_stride_any_async(hipStream_t stream, int_fast32_t n, const double *alpha, struct exo_win_1f64c x, struct exo_win_1f64 y);
int exoblas_daxpy_stride_1_async(hipStream_t stream, int_fast32_t n, const double *alpha, struct exo_win_1f64c x, struct exo_win_1f64 y);
int exoblas_saxpy_alpha_1_stride_any_async(hipStream_t stream, int_fast32_t n, struct exo_win_1f32c x, struct exo_win_1f32 y);
int exoblas_saxpy_alpha_1_stride_1_async(hipStream_t stream, int_fast32_t n, struct exo_win_1f32c x, struct exo_win_1f32 y);
int exoblas_daxpy_alpha_1_stride_any_async(hipStream_t stream, int_fast32_t n, struct exo_win_1f64c x, struct exo_win_1f64 y);
int exoblas_daxpy_alpha_1_stride_1_async(hipStream_t stream, int_fast32_t n, struct exo_win_1f64c x, struct exo_win_1f64 y);
int exoblas_sdot_stride_any_async(hipStream_t stream, int_fast32_t n, struct exo_win_1f32c x, struct exo_win_1f32c y, float *result);
int exoblas_sdot_stride_1_async(hipStream_t stream, int_fast32_t n, struct exo_win_1f32c x, struct exo_win_1f32c y, float *result);
int exoblas_ddot_stride_any_async(hipStream_t stream, int_fast32_t n, struct exo_win_1f64c x, struct exo_win_1f64c y, double *result);
int exoblas_ddot_stride_1_async(hipStream_t stream, int_fast32_t n, struct exo_win_1f64c x, struct exo_win_1f64c y, double *result);
int exoblas_dsdot_stride_any_async(hipStream_t stream, int_fast32_t n, struct exo_win_1f32c x, struct exo_win_1f32c y, double *result);
int exoblas_dsdot_stride_1_async(hipStream_t stream, int_fast32_t n, struct exo_win_1f32c x, struct exo_win_1f32c y, double *result);
int exoblas_sdsdot_stride_any_async(hipStream_t stream, int_fast32_t n, const float *sb, struct exo_win_1f32c x, struct exo_win_1f32c y, float *result);
int exoblas_sdsdot_stride_1_async(hipStream_t stream, int_fast32_t n, const float *sb, struct exo_win_1f32c x, struct exo_win_1f32c y, float *result);
int exoblas_scopy_stride_any_async(hipStream_t stream, int_fast32_t n, struct exo_win_1f32c x, struct exo_win_1f32 y);
int exoblas_scopy_stride_1_async(hipStream_t stream, int_fast32_t n, struct exo_win_1f32c x, struct exo_win_1f32 y);
int exoblas_dcopy_stride_any_async(hipStream_t stream, int_fast32_t n, struct exo_win_1f64c x, struct exo_win_1f64 y);
int exoblas_dcopy_stride_1_async(hipStream_t stream, int_fast32_t n, struct exo_win_1f64c x, struct exo_win_1f64 y);
int exoblas_sswap_stride_any_async(hipStream_t stream, int_fast32_t n, struct exo_win_1f32 x, struct exo_win_1f32 y);
int exoblas_sswap_stride_1_async(hipStream_t stream, int_fast32_t n, struct exo_win_1f32 x, struct exo_win_1f32 y);
int exoblas_dswap_stride_any_async(hipStream_t stream, int_fast32_t n, struct exo_win_1f64 x, struct exo_win_1f64 y);
int exoblas_dswap_stride_1_async(hipStream_t stream, int_fast32_t n, struct exo_win_1f64 x, struct exo_win_1f64 y);
int exoblas_srot_stride_any_async(hipStream_t stream, int_fast32_t n, struct exo_win_1f32 x, struct exo_win_1f32 y, const float *c, const float *s);
int exoblas_srot_stride_1_async(hipStream_t stream, int_fast32_t n, struct exo_win_1f32 x, struct exo_win_1f32 y, const float *c, const float *s);
int exoblas_drot_stride_any_async(hipStream_t stream, int_fast32_t n, struct exo_win_1f64 x, struct exo_win_1f64 y, const double *c, const double *s);
int exoblas_drot_stride_1_async(hipStream_t stream, int_fast32_t n, struct exo_win_1f64 x, struct exo_win_1f64 y, const double *c, const double *s);
int exoblas_srotm_flag_neg_one_stride_any_async(hipStream_t stream, int_fast32_t n, struct exo_win_1f32 x, struct exo_win_1f32 y, const float *H);
int exoblas_srotm_flag_neg_one_stride_1_async(hipStream_t stream, int_fast32_t n, struct exo_win_1f32 x, struct exo_win_1f32 y, const float *H);
int exoblas_drotm_flag_neg_one_stride_any_async(hipStream_t stream, int_fast32_t n, struct exo_win_1f64 x, struct exo_win_1f64 y, const double *H);
int exoblas_drotm_flag_neg_one_stride_1_async(hipStream_t stream, int_fast32_t n, struct exo_win_1f64 x, struct exo_win_1f64 y, const double *H);
int exoblas_srotm_flag_zero_stride_any_async(hipStream_t stream, int_fast32_t n, struct exo_win_1f32 x, struct exo_win_1f32 y, const float *H);
int exoblas_srotm_flag_zero_stride_1_async(hipStream_t stream, int_fast32_t n, struct exo_win_1f32 x, struct exo_win_1f32 y, const float *H);
int exoblas_drotm_flag_zero_stride_any_async(hipStream_t stream, int_fast32_t n, struct exo_win_1f64 x, struct exo_win_1f64 y, const double *H);
int exoblas_drotm_flag_zero_stride_1_async(hipStream_t stream, int_fast32_t n, struct exo_win_1f64 x, struct exo_win_1f64 y, const double *H);
int exoblas_srotm_flag_one_stride_any_async(hipStream_t stream, int_fast32_t n, struct exo_win_1f32 x, struct exo_win_1f32 y, const float *H);
int exoblas_srotm_flag_one_stride_1_async(hipStream_t stream, int_fast32_t n, struct exo_win_1f32 x, struct exo_win_1f32 y, const float *H);
int exoblas_drotm_flag_one_stride_any_async(hipStream_t stream, int_fast32_t n, struct exo_win_1f64 x, struct exo_win_1f64 y, const double *H);
int exoblas_drotm_flag_one_stride_1_async(hipStream_t stream, int_fast32_t n, struct exo_win_1f64 x, struct exo_win_1f64 y, const double *H);
int exoblas_sscal_stride_any_async(hipStream_t stream, int_fast32_t n, const float *alpha, struct exo_win_1f32 x);
int exoblas_sscal_stride_1_async(hipStream_t stream, int_fast32_t n, const float *alpha, struct exo_win_1f32 x);
int exoblas_dscal_stride_any_async(hipStream_t stream, int_fast32_t n, const double *alpha, struct exo_win_1f64 x);
int exoblas_dscal_stride_1_async(hipStream_t stream, int_fast32_t n, const double *alpha, struct exo_win_1f64 x);
int exoblas_sscal_alpha_0_stride_any_async(hipStream_t stream, int_fast32_t n, struct exo_win_1f32 x);
int exoblas_sscal_alpha_0_stride_1_async(hipStream_t stream, int_fast32_t n, struct exo_win_1f32 x);
int exoblas_dscal_alpha_0_stride_any_async(hipStream_t stream, int_fast32_t n, struct exo_win_1f64 x);
int exoblas_dscal_alpha_0_stride_1_async(hipStream_t stream, int_fast32_t n, struct exo_win_1f64 x);
int exoblas_sgemv_rm_stride_any_async(hipStream_t stream, int_fast32_t TransA, int_fast32_t M, int_fast32_t N, const float *alpha, const float *beta, struct exo_win_2f32c A, struct exo_win_2f32c AT, struct exo_win_1f32c x, struct exo_win_1f32 y);
int exoblas_sgemv_rm_stride_1_async(hipStream_t stream, int_fast32_t TransA, int_fast32_t M, int_fast32_t N, const float *alpha, const float *beta, struct exo_win_2f32c A, struct exo_win_2f32c AT, struct exo_win_1f32c x, struct exo_win_1f32 y);
int exoblas_dgemv_rm_stride_any_async(hipStream_t stream, int_fast32_t TransA, int_fast32_t M, int_fast32_t N, const double *alpha, const double *beta, struct exo_win_2f64c A, struct exo_win_2f64c AT, struct exo_win_1f64c x, struct exo_win_1f64 y);
int exoblas_dgemv_rm_stride_1_async(hipStream_t stream, int_fast32_t TransA, int_fast32_t M, int_fast32_t N, const double *alpha, const double *beta, struct exo_win_2f64c A, struct exo_win_2f64c AT, struct exo_win_1f64c x, struct exo_win_1f64 y);
int exoblas_sger_rm_stride_any_async(hipStream_t stream, int_fast32_t m, int_fast32_t n, const float *alpha, struct exo_win_1f32c x, struct exo_win_1f32c y, struct exo_win_2f32 A);
int exoblas_sger_rm_stride_1_async(hipStream_t stream, int_fast32_t m, int_fast32_t n, const float *alpha, struct exo_win_1f32c x, struct exo_win_1f32c y, struct exo_win_2f32 A);
int exoblas_dger_rm_stride_any_async(hipStream_t stream, int_fast32_t m, int_fast32_t n, const double *alpha, struct exo_win_1f64c x, struct exo_win_1f64c y, struct exo_win_2f64 A);
int exoblas_dger_rm_stride_1_async(hipStream_t stream, int_fast32_t m, int_fast32_t n, const double *alpha, struct exo_win_1f64c x, struct exo_win_1f64c y, struct exo_win_2f64 A);
int exoblas_ssymv_rm_stride_any_async(hipStream_t stream, int_fast32_t Uplo, int_fast32_t n, const float *alpha, struct exo_win_2f32c A, struct exo_win_1f32c x, const float *beta, struct exo_win_1f32 y);
int exoblas_ssymv_rm_stride_1_async(hipStream_t stream, int_fast32_t Uplo, int_fast32_t n, const float *alpha, struct exo_win_2f32c A, struct exo_win_1f32c x, const float *beta, struct exo_win_1f32 y);
int exoblas_dsymv_rm_stride_any_async(hipStream_t stream, int_fast32_t Uplo, int_fast32_t n, const double *alpha, struct exo_win_2f64c A, struct exo_win_1f64c x, const double *beta, struct exo_win_1f64 y);
int exoblas_dsymv_rm_stride_1_async(hipStream_t stream, int_fast32_t Uplo, int_fast32_t n, const double *alpha, struct exo_win_2f64c A, struct exo_win_1f64c x, const double *beta, struct exo_win_1f64 y);
int exoblas_ssyr_rm_stride_any_async(hipStream_t stream, int_fast32_t Uplo, int_fast32_t n, const float *alpha, struct exo_win_1f32c x, struct exo_win_1f32c x_copy, struct exo_win_2f32 A);
int exoblas_ssyr_rm_stride_1_async(hipStream_t stream, int_fast32_t Uplo, int_fast32_t n, const float *alpha, struct exo_win_1f32c x, struct exo_win_1f32c x_copy, struct exo_win_2f32 A);
int exoblas_dsyr_rm_stride_any_async(hipStream_t stream, int_fast32_t Uplo, int_fast32_t n, const double *alpha, struct exo_win_1f64c x, struct exo_win_1f64c x_copy, struct exo_win_2f64 A);
int exoblas_dsyr_rm_stride_1_async(hipStream_t stream, int_fast32_t Uplo, int_fast32_t n, const double *alpha, struct exo_win_1f64c x, struct exo_win_1f64c x_copy, struct exo_win_2f64 A);
int exoblas_ssyr2_rm_stride_any_async(hipStream_t stream, int_fast32_t Uplo, int_fast32_t n, const float *alpha, struct exo_win_1f32c x, struct exo_win_1f32c x_copy, struct exo_win_1f32c y, struct exo_win_1f32c y_copy, struct exo_win_2f32 A);
int exoblas_ssyr2_rm_stride_1_async(hipStream_t stream, int_fast32_t Uplo, int_fast32_t n, const float *alpha, struct exo_win_1f32c x, struct exo_win_1f32c x_copy, struct exo_win_1f32c y, struct exo_win_1f32c y_copy, struct exo_win_2f32 A);
int exoblas_dsyr2_rm_stride_any_async(hipStream_t stream, int_fast32_t Uplo, int_fast32_t n, const double *alpha, struct exo_win_1f64c x, struct exo_win_1f64c x_copy, struct exo_win_1f64c y, struct exo_win_1f64c y_copy, struct exo_win_2f64 A);
int exoblas_dsyr2_rm_stride_1_async(hipStream_t stream, int_fast32_t Uplo, int_fast32_t n, const double *alpha, struct exo_win_1f64c x, struct exo_win_1f64c x_copy, struct exo_win_1f64c y, struct exo_win_1f64c y_copy, struct exo_win_2f64 A);
int exoblas_strmv_rm_stride_any_async(hipStream_t stream, int_fast32_t Uplo, int_fast32_t TransA, int_fast32_t Diag, int_fast32_t n, struct exo_win_1f32 x, struct exo_win_2f32c A);
int exoblas_strmv_rm_stride_1_async(hipStream_t stream, int_fast32_t Uplo, int_fast32_t TransA, int_fast32_t Diag, int_fast32_t n, struct exo_win_1f32 x, struct exo_win_2f32c A);
int exoblas_dtrmv_rm_stride_any_async(hipStream_t stream, int_fast32_t Uplo, int_fast32_t TransA, int_fast32_t Diag, int_fast32_t n, struct exo_win_1f64 x, struct exo_win_2f64c A);
int exoblas_dtrmv_rm_stride_1_async(hipStream_t stream, int_fast32_t Uplo, int_fast32_t TransA, int_fast32_t Diag, int_fast32_t n, struct exo_win_1f64 x, struct exo_win_2f64c A);
int exoblas_strsv_rm_stride_any_async(hipStream_t stream, int_fast32_t Uplo, int_fast32_t TransA, int_fast32_t Diag, int_fast32_t n, struct exo_win_1f32 x, struct exo_win_2f32c A);
int exoblas_strsv_rm_stride_1_async(hipStream_t stream, int_fast32_t Uplo, int_fast32_t TransA, int_fast32_t Diag, int_fast32_t n, struct exo_win_1f32 x, struct exo_win_2f32c A);
int exoblas_dtrsv_rm_stride_any_async(hipStream_t stream, int_fast32_t Uplo, int_fast32_t TransA, int_fast32_t Diag, int_fast32_t n, struct exo_win_1f64 x, struct exo_win_2f64c A);
int exoblas_dtrsv_rm_stride_1_async(hipStream_t stream, int_fast32_t Uplo, int_fast32_t TransA, int_fast32_t Diag, int_fast32_t n, struct exo_win_1f64 x, struct exo_win_2f64c A);
int exoblas_sgbmv_row_major_NonTrans_stride_any_async(hipStream_t stream, int_fast32_t m, int_fast32_t n, int_fast32_t kl, int_fast32_t ku, const float *alpha, const float *beta, struct exo_win_2f32c a, struct exo_win_1f32c x, struct exo_win_1f32 y);
int exoblas_sgbmv_row_major_NonTrans_stride_1_async(hipStream_t stream, int_fast32_t m, int_fast32_t n, int_fast32_t kl, int_fast32_t ku, const float *alpha, const float *beta, struct exo_win_2f32c a, struct exo_win_1f32c x, struct exo_win_1f32 y);
int exoblas_dgbmv_row_major_NonTrans_stride_any_async(hipStream_t stream, int_fast32_t m, int_fast32_t n, int_fast32_t kl, int_fast32_t ku, const double *alpha, const double *beta, struct exo_win_2f64c a, struct exo_win_1f64c x, struct exo_win_1f64 y);
int exoblas_dgbmv_row_major_NonTrans_stride_1_async(hipStream_t stream, int_fast32_t m, int_fast32_t n, int_fast32_t kl, int_fast32_t ku, const double *alpha, const double *beta, struct exo_win_2f64c a, struct exo_win_1f64c x, struct exo_win_1f64 y);
int exoblas_ssbmv_scal_y_stride_any_async(hipStream_t stream, int_fast32_t n, const float *beta, struct exo_win_1f32 y);
int exoblas_ssbmv_scal_y_stride_1_async(hipStream_t stream, int_fast32_t n, const float *beta, struct exo_win_1f32 y);
int exoblas_dsbmv_scal_y_stride_any_async(hipStream_t stream, int_fast32_t n, const double *beta, struct exo_win_1f64 y);
int exoblas_dsbmv_scal_y_stride_1_async(hipStream_t stream, int_fast32_t n, const double *beta, struct exo_win_1f64 y);
int exoblas_ssbmv_row_major_Upper_stride_any_async(hipStream_t stream, int_fast32_t n, int_fast32_t k, const float *alpha, struct exo_win_2f32c A, struct exo_win_1f32c x, struct exo_win_1f32 y);
int exoblas_ssbmv_row_major_Upper_stride_1_async(hipStream_t stream, int_fast32_t n, int_fast32_t k, const float *alpha, struct exo_win_2f32c A, struct exo_win_1f32c x, struct exo_win_1f32 y);
int exoblas_dsbmv_row_major_Upper_stride_any_async(hipStream_t stream, int_fast32_t n, int_fast32_t k, const double *alpha, struct exo_win_2f64c A, struct exo_win_1f64c x, struct exo_win_1f64 y);
int exoblas_dsbmv_row_major_Upper_stride_1_async(hipStream_t stream, int_fast32_t n, int_fast32_t k, const double *alpha, struct exo_win_2f64c A, struct exo_win_1f64c x, struct exo_win_1f64 y);
int exoblas_ssbmv_row_major_Lower_stride_any_async(hipStream_t stream, int_fast32_t n, int_fast32_t k, const float *alpha, struct exo_win_2f32c A, struct exo_win_1f32c x, struct exo_win_1f32 y);
int exoblas_ssbmv_row_major_Lower_stride_1_async(hipStream_t stream, int_fast32_t n, int_fast32_t k, const float *alpha, struct exo_win_2f32c A, struct exo_win_1f32c x, struct exo_win_1f32 y);
int exoblas_dsbmv_row_major_Lower_stride_any_async(hipStream_t stream, int_fast32_t n, int_fast32_t k, const double *alpha, struct exo_win_2f64c A, struct exo_win_1f64c x, struct exo_win_1f64 y);
int exoblas_dsbmv_row_major_Lower_stride_1_async(hipStream_t stream, int_fast32_t n, int_fast32_t k, const double *alpha, struct exo_win_2f64c A, struct exo_win_1f64c x, struct exo_win_1f64 y);
int exoblas_stbmv_row_major_Upper_NonTrans_stride_any_async(hipStream_t stream, int_fast32_t n, int_fast32_t k, struct exo_win_1f32 x, struct exo_win_2f32c A, int_fast32_t Diag);
int exoblas_stbmv_row_major_Upper_NonTrans_stride_1_async(hipStream_t stream, int_fast32_t n, int_fast32_t k, struct exo_win_1f32 x, struct exo_win_2f32c A, int_fast32_t Diag);
int exoblas_dtbmv_row_major_Upper_NonTrans_stride_any_async(hipStream_t stream, int_fast32_t n, int_fast32_t k, struct exo_win_1f64 x, struct exo_win_2f64c A, int_fast32_t Diag);
int exoblas_dtbmv_row_major_Upper_NonTrans_stride_1_async(hipStream_t stream, int_fast32_t n, int_fast32_t k, struct exo_win_1f64 x, struct exo_win_2f64c A, int_fast32_t Diag);
int exoblas_stbmv_row_major_Lower_NonTrans_stride_any_async(hipStream_t stream, int_fast32_t n, int_fast32_t k, struct exo_win_1f32 x, struct exo_win_2f32c A, int_fast32_t Diag);
int exoblas_stbmv_row_major_Lower_NonTrans_stride_1_async(hipStream_t stream, int_fast32_t n, int_fast32_t k, struct exo_win_1f32 x, struct exo_win_2f32c A, int_fast32_t Diag);
int exoblas_dtbmv_row_major_Lower_NonTrans_stride_any_async(hipStream_t stream, int_fast32_t n, int_fast32_t k, struct exo_win_1f64 x, struct exo_win_2f64c A, int_fast32_t Diag);
int exoblas_dtbmv_row_major_Lower_NonTrans_stride_1_async(hipStream_t stream, int_fast32_t n, int_fast32_t k, struct exo_win_1f64 x, struct exo_win_2f64c A, int_fast32_t Diag);
int exoblas_stbmv_row_major_Upper_Trans_stride_any_async(hipStream_t stream, int_fast32_t n, int_fast32_t k, struct exo_win_1f32 x, struct exo_win_2f32c A, int_fast32_t Diag);
int exoblas_stbmv_row_major_Upper_Trans_stride_1_async(hipStream_t stream, int_fast32_t n, int_fast32_t k, struct exo_win_1f32 x, struct exo_win_2f32c A, int_fast32_t Diag);
int exoblas_dtbmv_row_major_Upper_Trans_stride_any_async(hipStream_t stream, int_fast32_t n, int_fast32_t k, struct exo_win_1f64 x, struct exo_win_2f64c A, int_fast32_t Diag);
int exoblas_dtbmv_row_major_Upper_Trans_stride_1_async(hipStream_t stream, int_fast32_t n, int_fast32_t k, struct exo_win_1f64 x, struct exo_win_2f64c A, int_fast32_t Diag);
int exoblas_stbmv_row_major_Lower_Trans_stride_any_async(hipStream_t stream, int_fast32_t n, int_fast32_t k, struct exo_win_1f32 x, struct exo_win_2f32c A, int_fast32_t Diag);
int exoblas_stbmv_row_major_Lower_Trans_stride_1_async(hipStream_t stream, int_fast32_t n, int_fast32_t k, struct exo_win_1f32 x, struct exo_win_2f32c A, int_fast32_t Diag);
int exoblas_dtbmv_row_major_Lower_Trans_stride_any_async(hipStream_t stream, int_fast32_t n, int_fast32_t k, struct exo_win_1f64 x, struct exo_win_2f64c A, int_fast32_t Diag);
int exoblas_dtbmv_row_major_Lower_Trans_stride_1_async(hipStream_t stream, int_fast32_t n, int_fast32_t k, struct exo_win_1f64 x, struct exo_win_2f64c A, int_fast32_t Diag);
int exoblas_stbsv_row_major_Upper_NonTrans_stride_any_async(hipStream_t stream, int_fast32_t n, int_fast32_t k, struct exo_win_1f32 x, struct exo_win_2f32c A, int_fast32_t Diag);
int exoblas_stbsv_row_major_Upper_NonTrans_stride_1_async(hipStream_t stream, int_fast32_t n, int_fast32_t k, struct exo_win_1f32 x, struct exo_win_2f32c A, int_fast32_t Diag);
int exoblas_dtbsv_row_major_Upper_NonTrans_stride_any_async(hipStream_t stream, int_fast32_t n, int_fast32_t k, struct exo_win_1f64 x, struct exo_win_2f64c A, int_fast32_t Diag);
int exoblas_dtbsv_row_major_Upper_NonTrans_stride_1_async(hipStream_t stream, int_fast32_t n, int_fast32_t k, struct exo_win_1f64 x, struct exo_win_2f64c A, int_fast32_t Diag);
int exoblas_stbsv_row_major_Lower_NonTrans_stride_any_async(hipStream_t stream, int_fast32_t n, int_fast32_t k, struct exo_win_1f32 x, struct exo_win_2f32c A, int_fast32_t Diag);
int exoblas_stbsv_row_major_Lower_NonTrans_stride_1_async(hipStream_t stream, int_fast32_t n, int_fast32_t k, struct exo_win_1f32 x, struct exo_win_2f32c A, int_fast32_t Diag);
int exoblas_dtbsv_row_major_Lower_NonTrans_stride_any_async(hipStream_t stream, int_fast32_t n, int_fast32_t k, struct exo_win_1f64 x, struct exo_win_2f64c A, int_fast32_t Diag);
int exoblas_dtbsv_row_major_Lower_NonTrans_stride_1_async(hipStream_t stream, int_fast32_t n, int_fast32_t k, struct exo_win_1f64 x, struct exo_win_2f64c A, int_fast32_t Diag);
int exoblas_stbsv_row_major_Upper_Trans_stride_any_async(hipStream_t stream, int_fast32_t n, int_fast32_t k, struct exo_win_1f32 x, struct exo_win_2f32c A, int_fast32_t Diag);
int exoblas_dtbsv_row_major_Upper_Trans_stride_any_async(hipStream_t stream, int_fast32_t n, int_fast32_t k, struct exo_win_1f64 x, struct exo_win_2f64c A, int_fast32_t Diag);
int exoblas_stbsv_row_major_Lower_Trans_stride_any_async(hipStream_t stream, int_fast32_t n, int_fast32_t k, struct exo_win_1f32 x, struct exo_win_2f32c A, int_fast32_t Diag);
int exoblas_dtbsv_row_major_Lower_Trans_stride_any_async(hipStream_t stream, int_fast32_t n, int_fast32_t k, struct exo_win_1f64 x, struct exo_win_2f64c A, int_fast32_t Diag);
int exoblas_sgemm_stride_any_async(hipStream_t stream, int_fast32_t TransA, int_fast32_t TransB, int_fast32_t M, int_fast32_t N, int_fast32_t K, const float *alpha, struct exo_win_2f32c A, struct exo_win_2f32c B, struct exo_win_2f32c AT, struct exo_win_2f32c BT, struct exo_win_2f32 C);
int exoblas_sgemm_stride_1_async(hipStream_t stream, int_fast32_t TransA, int_fast32_t TransB, int_fast32_t M, int_fast32_t N, int_fast32_t K, const float *alpha, struct exo_win_2f32c A, struct exo_win_2f32c B, struct exo_win_2f32c AT, struct exo_win_2f32c BT, struct exo_win_2f32 C);
int exoblas_dgemm_stride_any_async(hipStream_t stream, int_fast32_t TransA, int_fast32_t TransB, int_fast32_t M, int_fast32_t N, int_fast32_t K, const double *alpha, struct exo_win_2f64c A, struct exo_win_2f64c B, struct exo_win_2f64c AT, struct exo_win_2f64c BT, struct exo_win_2f64 C);
int exoblas_dgemm_stride_1_async(hipStream_t stream, int_fast32_t TransA, int_fast32_t TransB, int_fast32_t M, int_fast32_t N, int_fast32_t K, const double *alpha, struct exo_win_2f64c A, struct exo_win_2f64c B, struct exo_win_2f64c AT, struct exo_win_2f64c BT, struct exo_win_2f64 C);
int exoblas_smscal_rm_stride_any_async(hipStream_t stream, int_fast32_t M, int_fast32_t N, const float *alpha, struct exo_win_2f32 A);
int exoblas_smscal_rm_stride_1_async(hipStream_t stream, int_fast32_t M, int_fast32_t N, const float *alpha, struct exo_win_2f32 A);
int exoblas_dmscal_rm_stride_any_async(hipStream_t stream, int_fast32_t M, int_fast32_t N, const double *alpha, struct exo_win_2f64 A);
int exoblas_dmscal_rm_stride_1_async(hipStream_t stream, int_fast32_t M, int_fast32_t N, const double *alpha, struct exo_win_2f64 A);
int exoblas_strmscal_rm_stride_any_async(hipStream_t stream, int_fast32_t Uplo, int_fast32_t N, const float *alpha, struct exo_win_2f32 A);
int exoblas_strmscal_rm_stride_1_async(hipStream_t stream, int_fast32_t Uplo, int_fast32_t N, const float *alpha, struct exo_win_2f32 A);
int exoblas_dtrmscal_rm_stride_any_async(hipStream_t stream, int_fast32_t Uplo, int_fast32_t N, const double *alpha, struct exo_win_2f64 A);
int exoblas_dtrmscal_rm_stride_1_async(hipStream_t stream, int_fast32_t Uplo, int_fast32_t N, const double *alpha, struct exo_win_2f64 A);
int exoblas_ssymm_rm_stride_any_async(hipStream_t stream, int_fast32_t Side, int_fast32_t Uplo, int_fast32_t M, int_fast32_t N, const float *alpha, struct exo_win_2f32c Aleft, struct exo_win_2f32c Aright, struct exo_win_2f32c B, struct exo_win_2f32 C);
int exoblas_ssymm_rm_stride_1_async(hipStream_t stream, int_fast32_t Side, int_fast32_t Uplo, int_fast32_t M, int_fast32_t N, const float *alpha, struct exo_win_2f32c Aleft, struct exo_win_2f32c Aright, struct exo_win_2f32c B, struct exo_win_2f32 C);
int exoblas_dsymm_rm_stride_any_async(hipStream_t stream, int_fast32_t Side, int_fast32_t Uplo, int_fast32_t M, int_fast32_t N, const double *alpha, struct exo_win_2f64c Aleft, struct exo_win_2f64c Aright, struct exo_win_2f64c B, struct exo_win_2f64 C);
int exoblas_dsymm_rm_stride_1_async(hipStream_t stream, int_fast32_t Side, int_fast32_t Uplo, int_fast32_t M, int_fast32_t N, const double *alpha, struct exo_win_2f64c Aleft, struct exo_win_2f64c Aright, struct exo_win_2f64c B, struct exo_win_2f64 C);
int exoblas_ssyrk_rm_stride_any_async(hipStream_t stream, int_fast32_t Uplo, int_fast32_t Trans, int_fast32_t N, int_fast32_t K, const float *alpha, struct exo_win_2f32c A, struct exo_win_2f32c A_, struct exo_win_2f32c AT, struct exo_win_2f32c AT_, struct exo_win_2f32 C);
int exoblas_ssyrk_rm_stride_1_async(hipStream_t stream, int_fast32_t Uplo, int_fast32_t Trans, int_fast32_t N, int_fast32_t K, const float *alpha, struct exo_win_2f32c A, struct exo_win_2f32c A_, struct exo_win_2f32c AT, struct exo_win_2f32c AT_, struct exo_win_2f32 C);
int exoblas_dsyrk_rm_stride_any_async(hipStream_t stream, int_fast32_t Uplo, int_fast32_t Trans, int_fast32_t N, int_fast32_t K, const double *alpha, struct exo_win_2f64c A, struct exo_win_2f64c A_, struct exo_win_2f64c AT, struct exo_win_2f64c AT_, struct exo_win_2f64 C);
int exoblas_dsyrk_rm_stride_1_async(hipStream_t stream, int_fast32_t Uplo, int_fast32_t Trans, int_fast32_t N, int_fast32_t K, const double *alpha, struct exo_win_2f64c A, struct exo_win_2f64c A_, struct exo_win_2f64c AT, struct exo_win_2f64c AT_, struct exo_win_2f64 C);
int exoblas_ssyr2k_rm_stride_any_async(hipStream_t stream, int_fast32_t Uplo, int_fast32_t Trans, int_fast32_t N, int_fast32_t K, const float *alpha, struct exo_win_2f32c A, struct exo_win_2f32c A_, struct exo_win_2f32c B, struct exo_win_2f32c B_, struct exo_win_2f32c AT, struct exo_win_2f32c AT_, struct exo_win_2f32c BT, struct exo_win_2f32c BT_, struct exo_win_2f32 C);
int exoblas_ssyr2k_rm_stride_1_async(hipStream_t stream, int_fast32_t Uplo, int_fast32_t Trans, int_fast32_t N, int_fast32_t K, const float *alpha, struct exo_win_2f32c A, struct exo_win_2f32c A_, struct exo_win_2f32c B, struct exo_win_2f32c B_, struct exo_win_2f32c AT, struct exo_win_2f32c AT_, struct exo_win_2f32c BT, struct exo_win_2f32c BT_, struct exo_win_2f32 C);
int exoblas_dsyr2k_rm_stride_any_async(hipStream_t stream, int_fast32_t Uplo, int_fast32_t Trans, int_fast32_t N, int_fast32_t K, const double *alpha, struct exo_win_2f64c A, struct exo_win_2f64c A_, struct exo_win_2f64c B, struct exo_win_2f64c B_, struct exo_win_2f64c AT, struct exo_win_2f64c AT_, struct exo_win_2f64c BT, struct exo_win_2f64c BT_, struct exo_win_2f64 C);
int exoblas_dsyr2k_rm_stride_1_async(hipStream_t stream, int_fast32_t Uplo, int_fast32_t Trans, int_fast32_t N, int_fast32_t K, const double *alpha, struct exo_win_2f64c A, struct exo_win_2f64c A_, struct exo_win_2f64c B, struct exo_win_2f64c B_, struct exo_win_2f64c AT, struct exo_win_2f64c AT_, struct exo_win_2f64c BT, struct exo_win_2f64c BT_, struct exo_win_2f64 C);
int exoblas_strmm_row_major_Left_Upper_NonTrans_async(hipStream_t stream, int_fast32_t m, int_fast32_t n, const float *alpha, struct exo_win_2f32c A, struct exo_win_2f32 B, int_fast32_t Diag);
int exoblas_strmm_row_major_Left_Lower_NonTrans_async(hipStream_t stream, int_fast32_t m, int_fast32_t n, const float *alpha, struct exo_win_2f32c A, struct exo_win_2f32 B, int_fast32_t Diag);
int exoblas_strmm_row_major_Left_Upper_Trans_async(hipStream_t stream, int_fast32_t m, int_fast32_t n, const float *alpha, struct exo_win_2f32c A, struct exo_win_2f32 B, int_fast32_t Diag);
int exoblas_strmm_row_major_Left_Lower_Trans_async(hipStream_t stream, int_fast32_t m, int_fast32_t n, const float *alpha, struct exo_win_2f32c A, struct exo_win_2f32 B, int_fast32_t Diag);

#ifdef __cplusplus
}
#endif

#endif
