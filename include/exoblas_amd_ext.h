/* Generated by exoblas_amd/generator.py -- do not edit. */
/* Native entry points of exoblas_amd with no counterpart in the reference
 * (the reference's exo_* functions are plain-precision only).  The plain
 * form takes host or device pointers and returns once results are
 * visible; the _async form takes device pointers and enqueues on
 * `stream`.  Both return 0 or an EXOBLAS_STATUS_* code (exoblas_amd.h). */
#ifndef EXOBLAS_AMD_EXT_H
#define EXOBLAS_AMD_EXT_H

#include <hip/hip_runtime_api.h>
#include "exo_types.h"

#ifdef __cplusplus
extern "C" {
#endif

/* C += alpha op(A) op(B) with FP32 operands split into BF16 terms on the BF16 MFMA; arguments as exo_sgemm_stride_1 (src/level3/gemm.py:17-47) plus the number of BF16 products */
int exoblas_sgemm_split(void *ctxt, int_fast32_t TransA, int_fast32_t TransB, int_fast32_t M, int_fast32_t N, int_fast32_t K, const float *alpha, struct exo_win_2f32c A, struct exo_win_2f32c B, struct exo_win_2f32c AT, struct exo_win_2f32c BT, struct exo_win_2f32 C, int_fast32_t terms);
int exoblas_sgemm_split_async(hipStream_t stream, int_fast32_t TransA, int_fast32_t TransB, int_fast32_t M, int_fast32_t N, int_fast32_t K, const float *alpha, struct exo_win_2f32c A, struct exo_win_2f32c B, struct exo_win_2f32c AT, struct exo_win_2f32c BT, struct exo_win_2f32 C, int_fast32_t terms);

/* one triangle of C += alpha A A^T (or A^T A) in the split BF16 form; arguments as exo_ssyrk_rm_stride_1 (src/level3/syrk.py:17-32) without the duplicate windows */
int exoblas_ssyrk_split(void *ctxt, int_fast32_t Uplo, int_fast32_t Trans, int_fast32_t N, int_fast32_t K, const float *alpha, struct exo_win_2f32c A, struct exo_win_2f32c AT, struct exo_win_2f32 C, int_fast32_t terms);
int exoblas_ssyrk_split_async(hipStream_t stream, int_fast32_t Uplo, int_fast32_t Trans, int_fast32_t N, int_fast32_t K, const float *alpha, struct exo_win_2f32c A, struct exo_win_2f32c AT, struct exo_win_2f32 C, int_fast32_t terms);

/* C = alpha op(A) op(B) + beta C with CBLAS dimensions in row-major (Order 101) or column-major (102) storage, C not read when beta == 0, beta applied in the product's epilogue; each window is the stored array (rows x leading dimension); the product as exo_<p>gemm_stride_1 (src/level3/gemm.py:17-47) */
int exoblas_sgemm_ex(void *ctxt, int_fast32_t Order, int_fast32_t TransA, int_fast32_t TransB, int_fast32_t M, int_fast32_t N, int_fast32_t K, const float *alpha, struct exo_win_2f32c A, struct exo_win_2f32c B, const float *beta, struct exo_win_2f32 C);
int exoblas_sgemm_ex_async(hipStream_t stream, int_fast32_t Order, int_fast32_t TransA, int_fast32_t TransB, int_fast32_t M, int_fast32_t N, int_fast32_t K, const float *alpha, struct exo_win_2f32c A, struct exo_win_2f32c B, const float *beta, struct exo_win_2f32 C);

/* C = alpha op(A) op(B) + beta C with CBLAS dimensions in row-major (Order 101) or column-major (102) storage, C not read when beta == 0, beta applied in the product's epilogue; each window is the stored array (rows x leading dimension); the product as exo_<p>gemm_stride_1 (src/level3/gemm.py:17-47) */
int exoblas_dgemm_ex(void *ctxt, int_fast32_t Order, int_fast32_t TransA, int_fast32_t TransB, int_fast32_t M, int_fast32_t N, int_fast32_t K, const double *alpha, struct exo_win_2f64c A, struct exo_win_2f64c B, const double *beta, struct exo_win_2f64 C);
int exoblas_dgemm_ex_async(hipStream_t stream, int_fast32_t Order, int_fast32_t TransA, int_fast32_t TransB, int_fast32_t M, int_fast32_t N, int_fast32_t K, const double *alpha, struct exo_win_2f64c A, struct exo_win_2f64c B, const double *beta, struct exo_win_2f64 C);

/* y = alpha op(A) x + beta y for an M x N matrix A in row-major (Order 101) or column-major (102) storage, y not read when beta == 0; element i of x at x.data + i * x.strides[0]; the product as exo_<p>gemv_rm_stride_any (src/level2/gemv.py:10-22) */
int exoblas_sgemv_ex(void *ctxt, int_fast32_t Order, int_fast32_t TransA, int_fast32_t M, int_fast32_t N, const float *alpha, struct exo_win_2f32c A, struct exo_win_1f32c x, const float *beta, struct exo_win_1f32 y);
int exoblas_sgemv_ex_async(hipStream_t stream, int_fast32_t Order, int_fast32_t TransA, int_fast32_t M, int_fast32_t N, const float *alpha, struct exo_win_2f32c A, struct exo_win_1f32c x, const float *beta, struct exo_win_1f32 y);

/* y = alpha op(A) x + beta y for an M x N matrix A in row-major (Order 101) or column-major (102) storage, y not read when beta == 0; element i of x at x.data + i * x.strides[0]; the product as exo_<p>gemv_rm_stride_any (src/level2/gemv.py:10-22) */
int exoblas_dgemv_ex(void *ctxt, int_fast32_t Order, int_fast32_t TransA, int_fast32_t M, int_fast32_t N, const double *alpha, struct exo_win_2f64c A, struct exo_win_1f64c x, const double *beta, struct exo_win_1f64 y);
int exoblas_dgemv_ex_async(hipStream_t stream, int_fast32_t Order, int_fast32_t TransA, int_fast32_t M, int_fast32_t N, const double *alpha, struct exo_win_2f64c A, struct exo_win_1f64c x, const double *beta, struct exo_win_1f64 y);

/* op(A) x = b with BF16 (update 16) or FP32 (update 32) off-diagonal updates, FP64 diagonal-block solves and `refine` steps of iterative refinement on an FP64 residual; arguments as exo_strsv_rm_stride_any (src/level2/trsv.py:10-52) plus update and refine */
int exoblas_strsv_mixed(void *ctxt, int_fast32_t Uplo, int_fast32_t TransA, int_fast32_t Diag, int_fast32_t n, struct exo_win_1f32 x, struct exo_win_2f32c A, int_fast32_t update, int_fast32_t refine);
int exoblas_strsv_mixed_async(hipStream_t stream, int_fast32_t Uplo, int_fast32_t TransA, int_fast32_t Diag, int_fast32_t n, struct exo_win_1f32 x, struct exo_win_2f32c A, int_fast32_t update, int_fast32_t refine);

#ifdef __cplusplus
}
#endif

#endif
