// Kernel adapters behind the exo_* C-ABI.  One adapter per reference proc:
// plain device pointers, element strides (negative allowed for vectors,
// pointer at element 0), scalars by value.  Each validates nothing the
// generated export already checked and picks a HIP kernel variant.
#pragma once

#include <cstdint>

#include "runtime.hpp"

namespace exoblas {
namespace ops {

// ---------------------------------------------------------------- level 1
template <class T> void asum(Ctx&, int64_t n, const T* x, int64_t incx, T* result);
template <class T> void axpy(Ctx&, int64_t n, T alpha, const T* x, int64_t incx, T* y, int64_t incy);
template <class T> void axpy_alpha_1(Ctx&, int64_t n, const T* x, int64_t incx, T* y, int64_t incy);
template <class T> void dot(Ctx&, int64_t n, const T* x, int64_t incx, const T* y, int64_t incy, T* result);
void dsdot(Ctx&, int64_t n, const float* x, int64_t incx, const float* y, int64_t incy, double* result);
void sdsdot(Ctx&, int64_t n, float sb, const float* x, int64_t incx, const float* y, int64_t incy,
            float* result);
template <class T> void copy(Ctx&, int64_t n, const T* x, int64_t incx, T* y, int64_t incy);
template <class T> void swap(Ctx&, int64_t n, T* x, int64_t incx, T* y, int64_t incy);
template <class T> void rot(Ctx&, int64_t n, T* x, int64_t incx, T* y, int64_t incy, T c, T s);
template <class T> void rotm_flag_neg_one(Ctx&, int64_t n, T* x, int64_t incx, T* y, int64_t incy, const T* H);
template <class T> void rotm_flag_zero(Ctx&, int64_t n, T* x, int64_t incx, T* y, int64_t incy, const T* H);
template <class T> void rotm_flag_one(Ctx&, int64_t n, T* x, int64_t incx, T* y, int64_t incy, const T* H);
template <class T> void scal(Ctx&, int64_t n, T alpha, T* x, int64_t incx);
template <class T> void scal_alpha_0(Ctx&, int64_t n, T* x, int64_t incx);

// ---------------------------------------------------------------- level 2
template <class T>
void gemv_rm(Ctx&, int64_t TransA, int64_t M, int64_t N, T alpha, T beta, const T* A, int64_t lda,
             const T* AT, int64_t ldat, const T* x, int64_t incx, T* y, int64_t incy);
template <class T>
void ger_rm(Ctx&, int64_t m, int64_t n, T alpha, const T* x, int64_t incx, const T* y, int64_t incy, T* A,
            int64_t lda);
template <class T>
void symv_rm(Ctx&, int64_t Uplo, int64_t n, T alpha, const T* A, int64_t lda, const T* x, int64_t incx,
             T beta, T* y, int64_t incy);
template <class T>
void syr_rm(Ctx&, int64_t Uplo, int64_t n, T alpha, const T* x, int64_t incx, const T* x_copy,
            int64_t incxc, T* A, int64_t lda);
template <class T>
void syr2_rm(Ctx&, int64_t Uplo, int64_t n, T alpha, const T* x, int64_t incx, const T* x_copy,
             int64_t incxc, const T* y, int64_t incy, const T* y_copy, int64_t incyc, T* A, int64_t lda);
template <class T>
void trmv_rm(Ctx&, int64_t Uplo, int64_t TransA, int64_t Diag, int64_t n, T* x, int64_t incx, const T* A,
             int64_t lda);
template <class T>
void trsv_rm(Ctx&, int64_t Uplo, int64_t TransA, int64_t Diag, int64_t n, T* x, int64_t incx, const T* A,
             int64_t lda);
// Band solve through the block solver (trsv.hip): kb off-diagonals, row-major
// band storage with the diagonal in column c0.
template <class T>
void trsv_band_rm(Ctx&, bool fwd, bool trans, bool unit, int64_t n, int64_t kb, const T* A, int64_t lda, int64_t c0,
                  T* x, int64_t incx);

// ---------------------------------------------------------------- banded level 2
template <class T>
void gbmv_row_major_NonTrans(Ctx&, int64_t m, int64_t n, int64_t kl, int64_t ku, T alpha, T beta, const T* a,
                             int64_t lda, const T* x, int64_t incx, T* y, int64_t incy);
template <class T> void sbmv_scal_y(Ctx&, int64_t n, T beta, T* y, int64_t incy);
template <class T>
void sbmv_row_major_Upper(Ctx&, int64_t n, int64_t k, T alpha, const T* A, int64_t lda, const T* x, int64_t incx,
                          T* y, int64_t incy);
template <class T>
void sbmv_row_major_Lower(Ctx&, int64_t n, int64_t k, T alpha, const T* A, int64_t lda, const T* x, int64_t incx,
                          T* y, int64_t incy);
#define EXO_BAND_TRI_DECL(NAME) \
  template <class T>            \
  void NAME(Ctx&, int64_t n, int64_t k, T* x, int64_t incx, const T* A, int64_t lda, int64_t Diag);
EXO_BAND_TRI_DECL(tbmv_row_major_Upper_NonTrans)
EXO_BAND_TRI_DECL(tbmv_row_major_Lower_NonTrans)
EXO_BAND_TRI_DECL(tbmv_row_major_Upper_Trans)
EXO_BAND_TRI_DECL(tbmv_row_major_Lower_Trans)
EXO_BAND_TRI_DECL(tbsv_row_major_Upper_NonTrans)
EXO_BAND_TRI_DECL(tbsv_row_major_Lower_NonTrans)
EXO_BAND_TRI_DECL(tbsv_row_major_Upper_Trans)
EXO_BAND_TRI_DECL(tbsv_row_major_Lower_Trans)
#undef EXO_BAND_TRI_DECL
// ---------------------------------------------------------------- level 3
template <class T>
void gemm(Ctx&, int64_t TransA, int64_t TransB, int64_t M, int64_t N, int64_t K, T alpha, const T* A,
          int64_t lda, const T* B, int64_t ldb, const T* AT, int64_t ldat, const T* BT, int64_t ldbt, T* C,
          int64_t ldc);
template <class T> void mscal_rm(Ctx&, int64_t M, int64_t N, T alpha, T* A, int64_t lda);
template <class T> void trmscal_rm(Ctx&, int64_t Uplo, int64_t N, T alpha, T* A, int64_t lda);
template <class T>
void symm_rm(Ctx&, int64_t Side, int64_t Uplo, int64_t M, int64_t N, T alpha, const T* Aleft,
             int64_t ldal, const T* Aright, int64_t ldar, const T* B, int64_t ldb, T* C, int64_t ldc);
template <class T>
void syrk_rm(Ctx&, int64_t Uplo, int64_t Trans, int64_t N, int64_t K, T alpha, const T* A, int64_t lda,
             const T* A_, int64_t lda_, const T* AT, int64_t ldat, const T* AT_, int64_t ldat_, T* C,
             int64_t ldc);

template <class T>
void syr2k_rm(Ctx&, int64_t Uplo, int64_t Trans, int64_t N, int64_t K, T alpha, const T* A, int64_t lda,
              const T* A_, int64_t lda_, const T* B, int64_t ldb, const T* B_, int64_t ldb_, const T* AT,
              int64_t ldat, const T* AT_, int64_t ldat_, const T* BT, int64_t ldbt, const T* BT_, int64_t ldbt_,
              T* C, int64_t ldc);
// B <- alpha op(A) B, A triangular (unit diagonal when Diag != 0), in place.
template <class T>
void trmm_row_major_Left_Upper_NonTrans(Ctx&, int64_t m, int64_t n, T alpha, const T* A, int64_t lda, T* B,
                                        int64_t ldb, int64_t Diag);
template <class T>
void trmm_row_major_Left_Lower_NonTrans(Ctx&, int64_t m, int64_t n, T alpha, const T* A, int64_t lda, T* B,
                                        int64_t ldb, int64_t Diag);
template <class T>
void trmm_row_major_Left_Upper_Trans(Ctx&, int64_t m, int64_t n, T alpha, const T* A, int64_t lda, T* B,
                                     int64_t ldb, int64_t Diag);
template <class T>
void trmm_row_major_Left_Lower_Trans(Ctx&, int64_t m, int64_t n, T alpha, const T* A, int64_t lda, T* B,
                                     int64_t ldb, int64_t Diag);
// ---------------------------------------------------------------- CBLAS-shaped forms
// Order 101 (row-major) or 102 (column-major) operands, CBLAS dimensions and
// beta: C = alpha op(A) op(B) + beta C and y = alpha op(A) x + beta y, with C
// (y) not read when beta == 0.
void gemm_ex(Ctx&, int64_t Order, int64_t TransA, int64_t TransB, int64_t M, int64_t N, int64_t K, float alpha,
             const float* A, int64_t lda, const float* B, int64_t ldb, float beta, float* C, int64_t ldc);
void gemm_ex(Ctx&, int64_t Order, int64_t TransA, int64_t TransB, int64_t M, int64_t N, int64_t K, double alpha,
             const double* A, int64_t lda, const double* B, int64_t ldb, double beta, double* C, int64_t ldc);
void gemv_ex(Ctx&, int64_t Order, int64_t TransA, int64_t M, int64_t N, float alpha, const float* A, int64_t lda,
             const float* x, int64_t incx, float beta, float* y, int64_t incy);
void gemv_ex(Ctx&, int64_t Order, int64_t TransA, int64_t M, int64_t N, double alpha, const double* A,
             int64_t lda, const double* x, int64_t incx, double beta, double* y, int64_t incy);
// ---------------------------------------------------------------- split precision
// C += alpha op(A) op(B) (or one triangle of alpha A A^T / A^T A) with each
// FP32 operand split into `terms`-product BF16 pieces on the BF16 MFMA.
void gemm_split(Ctx&, int64_t TransA, int64_t TransB, int64_t M, int64_t N, int64_t K, float alpha,
                const float* A, int64_t lda, const float* B, int64_t ldb, const float* AT, int64_t ldat,
                const float* BT, int64_t ldbt, float* C, int64_t ldc, int64_t terms);
void syrk_split(Ctx&, int64_t Uplo, int64_t Trans, int64_t N, int64_t K, float alpha, const float* A, int64_t lda,
                const float* AT, int64_t ldat, float* C, int64_t ldc, int64_t terms);

// op(A) x = b in FP32 with BF16 (update 16) / FP32 (update 32) off-diagonal
// updates, FP64 diagonal blocks and `refine` FP64-residual refinement steps.
// With update 16 and `packed` (trsv_mixed_pack of the same A), the BF16
// copy is read from there instead of being made by the call.
void trsv_mixed(Ctx&, int64_t Uplo, int64_t TransA, int64_t Diag, int64_t n, float* x, int64_t incx, const float* A,
                int64_t lda, int64_t update, int64_t refine, const void* packed = nullptr);
// The BF16 copy the update-16 solve reads: n x n, row stride n, at `packed`.
void trsv_mixed_pack(Ctx&, int64_t n, const float* A, int64_t lda, void* packed);

}  // namespace ops
}  // namespace exoblas
