// CPU oracle for the exo_* BLAS entry points.
//
// TEST INFRASTRUCTURE ONLY.  This file restates, loop nest for loop nest, the
// Exo `@proc` specifications of the reference BLAS (src/level{1,2,3}/*.py).
// It is the checker that the HIP kernels are compared against and the CPU
// baseline that bench.py times; the shipped library never links or calls it.
//
// Every routine is a template over
//   T : the storage type of the caller's arrays (float or double), and
//   W : the working type every arithmetic operation is carried out in.
// W == T gives the strict restatement (same operation order and precision as
// the spec).  W == long double applied to double-converted inputs gives the
// "truth" oracle (64-bit significand) used to state errors against FP64.
//
// Arrays are addressed exactly as the callee of the reference C-ABI sees
// them: a pointer to element 0 and a signed element stride (negative strides
// walk downwards from element 0, which is what the reference wrappers pass
// after re-basing, e.g. test/level1/axpy/exo_axpy_wrapper.h:23-28).
// Matrices are row-major with unit column stride and leading dimension ld.

#include <cmath>
#include <cstdint>
#include <cstring>
#include <type_traits>
#include <vector>

// The all-cores CPU baseline build (-fopenmp -DEXO_ORACLE_OMP) runs the outer
// loops whose iterations are independent (elements, rows of the output) on
// every host thread, and the level-1 reductions as OpenMP + reductions; loops
// that carry a dependence or scatter (symv, trmv, trsv) stay sequential.
// The checker build defines neither, so its loops are the spec's.
#if defined(EXO_ORACLE_OMP)
#include <omp.h>
#define EXO_PRAGMA(x) _Pragma(#x)
#define EXO_PAR_FOR EXO_PRAGMA(omp parallel for schedule(static))
#define EXO_PAR_SUM(v) EXO_PRAGMA(omp parallel for schedule(static) reduction(+ : v))
#else
#define EXO_PAR_FOR
#define EXO_PAR_SUM(v)
#endif

namespace {

constexpr int64_t kNoTrans = 111;
constexpr int64_t kUpper = 121;
constexpr int64_t kNonUnit = 131;
constexpr int64_t kUnit = 132;
constexpr int64_t kLeft = 141;

// 1-D view.  When W == T it aliases the caller's memory; otherwise it holds a
// W-typed copy which is written back (rounded to T) on destruction if `out`.
template <class T, class W>
struct V1 {
  T* orig;
  int64_t n, os;
  bool out;
  std::vector<W> buf;
  W* p;
  int64_t s;
  V1(const T* data, int64_t n_, int64_t stride, bool out_)
      : orig(const_cast<T*>(data)), n(n_), os(stride), out(out_) {
    if constexpr (std::is_same<T, W>::value) {
      p = orig;
      s = stride;
    } else {
      buf.resize(n > 0 ? n : 1);
      for (int64_t i = 0; i < n; ++i) buf[i] = (W)orig[i * os];
      p = buf.data();
      s = 1;
    }
  }
  ~V1() {
    if constexpr (!std::is_same<T, W>::value) {
      if (out)
        for (int64_t i = 0; i < n; ++i) orig[i * os] = (T)buf[i];
    }
  }
  W& operator[](int64_t i) { return p[i * s]; }
};

// 2-D row-major view (unit column stride, leading dimension ld).
template <class T, class W>
struct M2 {
  T* orig;
  int64_t r, c, old;
  bool out;
  std::vector<W> buf;
  W* p;
  int64_t ld;
  M2(const T* data, int64_t rows, int64_t cols, int64_t ld_, bool out_)
      : orig(const_cast<T*>(data)), r(rows), c(cols), old(ld_), out(out_) {
    if constexpr (std::is_same<T, W>::value) {
      p = orig;
      ld = ld_;
    } else {
      buf.resize(r * c > 0 ? r * c : 1);
      for (int64_t i = 0; i < r; ++i)
        for (int64_t j = 0; j < c; ++j) buf[i * c + j] = (W)orig[i * old + j];
      p = buf.data();
      ld = c;
    }
  }
  ~M2() {
    if constexpr (!std::is_same<T, W>::value) {
      if (out)
        for (int64_t i = 0; i < r; ++i)
          for (int64_t j = 0; j < c; ++j) orig[i * old + j] = (T)buf[i * c + j];
    }
  }
  W& operator()(int64_t i, int64_t j) { return p[i * ld + j]; }
};

// ---------------------------------------------------------------- level 1

// src/level1/asum.py:11-15  result = sum select(0, x, x, -x)
template <class T, class W>
W asum(int64_t n, const T* x_, int64_t incx) {
  V1<T, W> x(x_, n, incx, false);
  W result = 0;
  EXO_PAR_SUM(result)
  for (int64_t i = 0; i < n; ++i) result += (x[i] > 0 ? x[i] : -x[i]);
  return result;
}

// src/level1/axpy.py:9-12
template <class T, class W>
void axpy(int64_t n, T alpha_, const T* x_, int64_t incx, T* y_, int64_t incy) {
  V1<T, W> x(x_, n, incx, false), y(y_, n, incy, true);
  W alpha = alpha_;
  EXO_PAR_FOR
  for (int64_t i = 0; i < n; ++i) y[i] += alpha * x[i];
}

// src/level1/axpy.py:15-18
template <class T, class W>
void axpy_alpha_1(int64_t n, const T* x_, int64_t incx, T* y_, int64_t incy) {
  V1<T, W> x(x_, n, incx, false), y(y_, n, incy, true);
  EXO_PAR_FOR
  for (int64_t i = 0; i < n; ++i) y[i] += x[i];
}

// src/level1/dot.py:9-13
template <class T, class W>
W dot(int64_t n, const T* x_, int64_t incx, const T* y_, int64_t incy) {
  V1<T, W> x(x_, n, incx, false), y(y_, n, incy, false);
  W result = 0;
  EXO_PAR_SUM(result)
  for (int64_t i = 0; i < n; ++i) result += x[i] * y[i];
  return result;
}

// src/level1/dsdot.py:11-21 (W is at least double: d_x, d_y, d_dot are f64)
template <class W>
W dsdot(int64_t n, const float* x, int64_t incx, const float* y, int64_t incy) {
  W d_dot = 0;
  EXO_PAR_SUM(d_dot)
  for (int64_t i = 0; i < n; ++i) {
    W d_x = (W)x[i * incx];
    W d_y = (W)y[i * incy];
    d_dot += d_x * d_y;
  }
  return d_dot;
}

// src/level1/dsdot.py:24-36: f64 dot, plus sb in f64, rounded once to f32
template <class W>
float sdsdot(int64_t n, float sb, const float* x, int64_t incx, const float* y, int64_t incy) {
  W d_result = dsdot<W>(n, x, incx, y, incy);
  d_result += (W)sb;
  return (float)d_result;
}

// src/level1/exo_copy.py:9-12
template <class T>
void copy(int64_t n, const T* x, int64_t incx, T* y, int64_t incy) {
  EXO_PAR_FOR
  for (int64_t i = 0; i < n; ++i) y[i * incy] = x[i * incx];
}

// src/level1/swap.py:9-15
template <class T>
void swap(int64_t n, T* x, int64_t incx, T* y, int64_t incy) {
  EXO_PAR_FOR
  for (int64_t i = 0; i < n; ++i) {
    T tmp = x[i * incx];
    x[i * incx] = y[i * incy];
    y[i * incy] = tmp;
  }
}

// src/level1/rot.py:9-15
template <class T, class W>
void rot(int64_t n, T* x_, int64_t incx, T* y_, int64_t incy, T c_, T s_) {
  V1<T, W> x(x_, n, incx, true), y(y_, n, incy, true);
  W c = c_, s = s_;
  EXO_PAR_FOR
  for (int64_t i = 0; i < n; ++i) {
    W xReg = x[i];
    x[i] = c * xReg + s * y[i];
    y[i] = -s * xReg + c * y[i];
  }
}

// src/level1/rotm.py:9-34; H is the row-major 2x2 {H00, H01, H10, H11}
// (the wrapper builds it from P as {P[1], P[3], P[2], P[4]},
// test/level1/rotm/exo_rotm_wrapper.h:8).
template <class T, class W>
void rotm(int flag, int64_t n, T* x_, int64_t incx, T* y_, int64_t incy, const T* H_) {
  V1<T, W> x(x_, n, incx, true), y(y_, n, incy, true);
  W h00 = H_[0], h01 = H_[1], h10 = H_[2], h11 = H_[3];
  EXO_PAR_FOR
  for (int64_t i = 0; i < n; ++i) {
    W xReg = x[i];
    if (flag == -1) {
      x[i] = h00 * xReg + h01 * y[i];
      y[i] = h10 * xReg + h11 * y[i];
    } else if (flag == 0) {
      x[i] = xReg + h01 * y[i];
      y[i] = h10 * xReg + y[i];
    } else {
      x[i] = h00 * xReg + y[i];
      y[i] = -xReg + h11 * y[i];
    }
  }
}

// src/level1/scal.py:9-19
template <class T, class W>
void scal(int64_t n, T alpha_, T* x_, int64_t incx) {
  V1<T, W> x(x_, n, incx, true);
  W alpha = alpha_;
  EXO_PAR_FOR
  for (int64_t i = 0; i < n; ++i) x[i] = alpha * x[i];
}

template <class T>
void scal_alpha_0(int64_t n, T* x, int64_t incx) {
  EXO_PAR_FOR
  for (int64_t i = 0; i < n; ++i) x[i * incx] = 0;
}

// ---------------------------------------------------------------- level 2

// src/level2/gemv.py:10-22.  m, n are the callee's M, N: for NoTrans A is
// m x n; for Trans the matrix read is AT, n x m (same data and lda).
template <class T, class W>
void gemv(int64_t trans, int64_t m, int64_t n, T alpha_, T beta_, const T* A_, int64_t lda,
          const T* x_, int64_t incx, T* y_, int64_t incy) {
  V1<T, W> x(x_, n, incx, false), y(y_, m, incy, true);
  W alpha = alpha_, beta = beta_;
  EXO_PAR_FOR
  for (int64_t i = 0; i < m; ++i) y[i] = y[i] * beta;
  if (trans == kNoTrans) {
    M2<T, W> A(A_, m, n, lda, false);
    EXO_PAR_FOR
    for (int64_t i = 0; i < m; ++i)
      for (int64_t j = 0; j < n; ++j) y[i] += alpha * (x[j] * A(i, j));
  } else {
    M2<T, W> AT(A_, n, m, lda, false);
#if defined(EXO_ORACLE_OMP)
    // each thread owns a contiguous range of y and walks the rows of AT over it
#pragma omp parallel
    {
      const int64_t nt = omp_get_num_threads(), t = omp_get_thread_num();
      const int64_t j0 = m * t / nt, j1 = m * (t + 1) / nt;
      for (int64_t i = 0; i < n; ++i)
        for (int64_t j = j0; j < j1; ++j) y[j] += alpha * x[i] * AT(i, j);
    }
#else
    for (int64_t i = 0; i < n; ++i)
      for (int64_t j = 0; j < m; ++j) y[j] += alpha * x[i] * AT(i, j);
#endif
  }
}

// src/level2/ger.py:9-15
template <class T, class W>
void ger(int64_t m, int64_t n, T alpha_, const T* x_, int64_t incx, const T* y_, int64_t incy,
         T* A_, int64_t lda) {
  V1<T, W> x(x_, m, incx, false), y(y_, n, incy, false);
  M2<T, W> A(A_, m, n, lda, true);
  W alpha = alpha_;
  EXO_PAR_FOR
  for (int64_t i = 0; i < m; ++i)
    for (int64_t j = 0; j < n; ++j) A(i, j) += alpha * x[i] * y[j];
}

// src/level2/symv.py:10-35
template <class T, class W>
void symv(int64_t uplo, int64_t n, T alpha_, const T* A_, int64_t lda, const T* x_, int64_t incx,
          T beta_, T* y_, int64_t incy) {
  V1<T, W> x(x_, n, incx, false), y(y_, n, incy, true);
  M2<T, W> A(A_, n, n, lda, false);
  W alpha = alpha_, beta = beta_;
  for (int64_t i = 0; i < n; ++i) y[i] = beta * y[i];
  for (int64_t i = 0; i < n; ++i) {
    if (uplo == kUpper) {
      W temp = alpha * x[n - i - 1];
      W dot = 0;
      for (int64_t j = 0; j < i + 1; ++j) {
        y[n - j - 1] += temp * A(n - i - 1, n - j - 1);
        dot += A(n - i - 1, n - j - 1) * x[n - j - 1];
      }
      y[n - i - 1] += alpha * (dot - A(n - i - 1, n - i - 1) * x[n - i - 1]);
    } else {
      W temp = alpha * x[i];
      W dot = 0;
      for (int64_t j = 0; j < i + 1; ++j) {
        y[j] += temp * A(i, j);
        dot += A(i, j) * x[j];
      }
      y[i] += alpha * (dot - A(i, i) * x[i]);
    }
  }
}

// src/level2/syr.py:10-19
template <class T, class W>
void syr(int64_t uplo, int64_t n, T alpha_, const T* x_, int64_t incx, T* A_, int64_t lda) {
  V1<T, W> x(x_, n, incx, false);
  M2<T, W> A(A_, n, n, lda, true);
  W alpha = alpha_;
  EXO_PAR_FOR
  for (int64_t i = 0; i < n; ++i)
    for (int64_t j = 0; j < i + 1; ++j) {
      if (uplo == kUpper)
        A(n - i - 1, n - j - 1) += (alpha * x[n - i - 1]) * x[n - j - 1];
      else
        A(i, j) += alpha * x[i] * x[j];
    }
}

// src/level2/syr2.py:9-27
template <class T, class W>
void syr2(int64_t uplo, int64_t n, T alpha_, const T* x_, int64_t incx, const T* y_, int64_t incy,
          T* A_, int64_t lda) {
  V1<T, W> x(x_, n, incx, false), y(y_, n, incy, false);
  M2<T, W> A(A_, n, n, lda, true);
  W alpha = alpha_;
  EXO_PAR_FOR
  for (int64_t i = 0; i < n; ++i)
    for (int64_t j = 0; j < i + 1; ++j) {
      if (uplo == kUpper)
        A(n - i - 1, n - j - 1) +=
            (alpha * x[n - i - 1]) * y[n - j - 1] + (alpha * y[n - i - 1]) * x[n - j - 1];
      else
        A(i, j) += (alpha * x[i]) * y[j] + (alpha * y[i]) * x[j];
    }
}

// src/level2/trmv.py:9-43 (xCopy temporary, unit diagonal via +(1 - A_ii) x_i)
template <class T, class W>
void trmv(int64_t uplo, int64_t trans, int64_t diag, int64_t n, T* x_, int64_t incx, const T* A_,
          int64_t lda) {
  V1<T, W> x(x_, n, incx, true);
  M2<T, W> A(A_, n, n, lda, false);
  std::vector<W> xCopy(n > 0 ? n : 1, W(0));
  for (int64_t i = 0; i < n; ++i) {
    if (trans == kNoTrans) {
      if (uplo == kUpper) {
        for (int64_t j = 0; j < i + 1; ++j) xCopy[n - i - 1] += A(n - i - 1, n - j - 1) * x[n - j - 1];
        if (diag == kUnit) xCopy[n - i - 1] += (1 - A(n - i - 1, n - i - 1)) * x[n - i - 1];
      } else {
        for (int64_t j = 0; j < i + 1; ++j) xCopy[i] += A(i, j) * x[j];
        if (diag == kUnit) xCopy[i] += (1 - A(i, i)) * x[i];
      }
    } else {
      if (uplo == kUpper) {
        for (int64_t j = 0; j < i + 1; ++j) xCopy[n - j - 1] += A(n - i - 1, n - j - 1) * x[n - i - 1];
        if (diag == kUnit) xCopy[n - i - 1] += (1 - A(n - i - 1, n - i - 1)) * x[n - i - 1];
      } else {
        for (int64_t j = 0; j < i + 1; ++j) xCopy[j] += A(i, j) * x[i];
        if (diag == kUnit) xCopy[i] += (1 - A(i, i)) * x[i];
      }
    }
  }
  for (int64_t l = 0; l < n; ++l) x[l] = xCopy[l];
}

// src/level2/trsv.py:10-52 (NoTrans: dot-form substitution; Trans: column
// updates accumulated in Dot[])
template <class T, class W>
void trsv(int64_t uplo, int64_t trans, int64_t diag, int64_t n, T* x_, int64_t incx, const T* A_,
          int64_t lda) {
  V1<T, W> x(x_, n, incx, true);
  M2<T, W> A(A_, n, n, lda, false);
  std::vector<W> Dot(n > 0 ? n : 1, W(0));
  for (int64_t i = 0; i < n; ++i) {
    if (trans == kNoTrans) {
      W dot = 0;
      if (uplo == kUpper) {
        for (int64_t j = 0; j < i; ++j) dot += A(n - i - 1, n - j - 1) * x[n - j - 1];
        x[n - i - 1] = x[n - i - 1] - dot;
        if (diag == kNonUnit) x[n - i - 1] = x[n - i - 1] / A(n - i - 1, n - i - 1);
      } else {
        for (int64_t j = 0; j < i; ++j) dot += A(i, j) * x[j];
        x[i] = x[i] - dot;
        if (diag == kNonUnit) x[i] = x[i] / A(i, i);
      }
    } else {
      if (uplo == kUpper) {
        x[i] = x[i] - Dot[i];
        if (diag == kNonUnit) x[i] = x[i] / A(i, i);
        for (int64_t j = 0; j < n - i - 1; ++j) Dot[n - j - 1] += A(i, n - j - 1) * x[i];
      } else {
        x[n - i - 1] = x[n - i - 1] - Dot[n - i - 1];
        if (diag == kNonUnit) x[n - i - 1] = x[n - i - 1] / A(n - i - 1, n - i - 1);
        for (int64_t j = 0; j < n - i - 1; ++j) Dot[j] += A(n - i - 1, j) * x[n - i - 1];
      }
    }
  }
}

// ---------------------------------------------------------------- level 3

// src/level3/gemm.py:17-47 (C += alpha * op(A) op(B); beta is the separate
// mscal pass).  A is M x K (NoTrans) or K x M (Trans); B is K x N or N x K.
template <class T, class W>
void gemm(int64_t ta, int64_t tb, int64_t M, int64_t N, int64_t K, T alpha_, const T* A_,
          int64_t lda, const T* B_, int64_t ldb, T* C_, int64_t ldc) {
  M2<T, W> A(A_, ta == kNoTrans ? M : K, ta == kNoTrans ? K : M, lda, false);
  M2<T, W> B(B_, tb == kNoTrans ? K : N, tb == kNoTrans ? N : K, ldb, false);
  M2<T, W> C(C_, M, N, ldc, true);
  W alpha = alpha_;
  EXO_PAR_FOR
  for (int64_t i = 0; i < M; ++i)
    for (int64_t j = 0; j < N; ++j)
      for (int64_t k = 0; k < K; ++k) {
        W a = ta == kNoTrans ? A(i, k) : A(k, i);
        W b = tb == kNoTrans ? B(k, j) : B(j, k);
        C(i, j) += alpha * (a * b);
      }
}

// src/level3/mscal.py:12-17
template <class T, class W>
void mscal(int64_t M, int64_t N, T alpha_, T* A_, int64_t lda) {
  M2<T, W> A(A_, M, N, lda, true);
  W alpha = alpha_;
  EXO_PAR_FOR
  for (int64_t i = 0; i < M; ++i)
    for (int64_t j = 0; j < N; ++j) A(i, j) = A(i, j) * alpha;
}

// src/level3/mscal.py:20-27
template <class T, class W>
void trmscal(int64_t uplo, int64_t N, T alpha_, T* A_, int64_t lda) {
  M2<T, W> A(A_, N, N, lda, true);
  W alpha = alpha_;
  EXO_PAR_FOR
  for (int64_t i = 0; i < N; ++i)
    for (int64_t j = 0; j < N; ++j)
      if ((uplo == kUpper && j >= i) || (uplo != kUpper && j < i + 1)) A(i, j) = A(i, j) * alpha;
}

// src/level3/symm.py:16-47.  A is M x M (Left) or N x N (Right), one
// triangle read according to uplo.
template <class T, class W>
void symm(int64_t side, int64_t uplo, int64_t M, int64_t N, T alpha_, const T* A_, int64_t lda,
          const T* B_, int64_t ldb, T* C_, int64_t ldc) {
  int64_t na = side == kLeft ? M : N;
  M2<T, W> A(A_, na, na, lda, false);
  M2<T, W> B(B_, M, N, ldb, false);
  M2<T, W> C(C_, M, N, ldc, true);
  W alpha = alpha_;
  EXO_PAR_FOR
  for (int64_t i = 0; i < M; ++i)
    for (int64_t j = 0; j < N; ++j) {
      if (side == kLeft) {
        for (int64_t k = 0; k < M; ++k) {
          W a_val;
          if (uplo == kUpper)
            a_val = k < i + 1 ? A(k, i) : A(i, k);
          else
            a_val = k < i + 1 ? A(i, k) : A(k, i);
          C(i, j) += alpha * (a_val * B(k, j));
        }
      } else {
        for (int64_t k = 0; k < N; ++k) {
          W a_val;
          if (uplo == kUpper)
            a_val = j < k + 1 ? A(j, k) : A(k, j);
          else
            a_val = j < k + 1 ? A(k, j) : A(j, k);
          C(i, j) += alpha * (B(i, k) * a_val);
        }
      }
    }
}

// src/level3/syrk.py:17-32.  A is N x K (NoTrans) or K x N (Trans).
template <class T, class W>
void syrk(int64_t uplo, int64_t trans, int64_t N, int64_t K, T alpha_, const T* A_, int64_t lda,
          T* C_, int64_t ldc) {
  M2<T, W> A(A_, trans == kNoTrans ? N : K, trans == kNoTrans ? K : N, lda, false);
  M2<T, W> C(C_, N, N, ldc, true);
  W alpha = alpha_;
  EXO_PAR_FOR
  for (int64_t i = 0; i < N; ++i)
    for (int64_t j = 0; j < N; ++j)
      if ((uplo == kUpper && j >= i) || (uplo != kUpper && j < i + 1))
        for (int64_t k = 0; k < K; ++k) {
          if (trans == kNoTrans)
            C(i, j) += alpha * (A(i, k) * A(j, k));
          else
            C(i, j) += alpha * (A(k, i) * A(k, j));
        }
}

}  // namespace

// ------------------------------------------------------------ banded level 2

// src/level2/gbmv.py:13-52.  a[i, kl + (j - i)] holds A(i, j); the four
// branches of the spec all sum the in-band columns of row i in increasing
// order.  A row with no in-band column (possible only when m > n + kl, where
// the spec leaves its sum unset) contributes 0.
template <class T, class W>
void gbmv(int64_t m, int64_t n, int64_t kl, int64_t ku, T alpha_, T beta_, const T* a_, int64_t lda,
          const T* x_, int64_t incx, T* y_, int64_t incy) {
  M2<T, W> a(a_, m, kl + ku + 1, lda, false);
  V1<T, W> x(x_, n, incx, false);
  V1<T, W> y(y_, m, incy, true);
  W alpha = alpha_, beta = beta_;
  for (int64_t i = 0; i < m; ++i) {
    W result = 0;
    const int64_t c0 = i - kl > 0 ? i - kl : 0, c1 = i + ku < n - 1 ? i + ku : n - 1;
    for (int64_t c = c0; c <= c1; ++c) result += a(i, kl + c - i) * x[c];
    y[i] = beta * y[i] + alpha * result;
  }
}

// src/level2/sbmv.py:12-16
template <class T, class W>
void sbmv_scal_y(int64_t n, T beta_, T* y_, int64_t incy) {
  V1<T, W> y(y_, n, incy, true);
  W beta = beta_;
  for (int64_t i = 0; i < n; ++i) y[i] = beta * y[i];
}

// src/level2/sbmv.py:19-33 (upper) and :36-50 (lower): y += alpha S x, S
// symmetric with k off-diagonals, one triangle stored in band form.
template <class T, class W>
void sbmv(int64_t upper, int64_t n, int64_t k, T alpha_, const T* A_, int64_t lda, const T* x_, int64_t incx, T* y_,
          int64_t incy) {
  M2<T, W> A(A_, n, k + 1, lda, false);
  V1<T, W> x(x_, n, incx, false);
  V1<T, W> y(y_, n, incy, true);
  W alpha = alpha_;
  for (int64_t i = 0; i < n; ++i) {
    const W temp = alpha * x[i];
    W dot = 0;
    for (int64_t j = 0; j < k; ++j) {
      if (upper) {
        if (i + j + 1 < n) {
          y[i + j + 1] += temp * A(i, j + 1);
          dot += A(i, j + 1) * x[i + j + 1];
        }
      } else if (i - j - 1 >= 0) {
        y[i - j - 1] += temp * A(i, k - j - 1);
        dot += A(i, k - j - 1) * x[i - j - 1];
      }
    }
    y[i] += temp * A(i, upper ? 0 : k) + alpha * dot;
  }
}

// src/level2/tbmv.py:12-141: x = op(A) x in place for a triangular band; the
// diagonal (A[i, 0] upper, A[i, k] lower) is read only when diag == 0.
template <class T, class W>
void tbmv(int64_t upper, int64_t trans, int64_t n, int64_t k, T* x_, int64_t incx, const T* A_, int64_t lda,
          int64_t diag) {
  M2<T, W> A(A_, n, k + 1, lda, false);
  V1<T, W> x(x_, n, incx, true);
  if (!trans && upper) {  // :12-43, rows upwards from 0; reads rows below, not yet updated
    for (int64_t i = 0; i < n - k; ++i) {
      W dot = diag == 0 ? x[i] * A(i, 0) : x[i];
      for (int64_t j = 0; j < k; ++j) dot += A(i, j + 1) * x[i + j + 1];
      x[i] = dot;
    }
    for (int64_t i = 0; i < k; ++i) {
      const int64_t r = n - k + i;
      W dot = diag == 0 ? x[r] * A(r, 0) : x[r];
      for (int64_t j = 0; j < k - i - 1; ++j) dot += A(r, j + 1) * x[r + j + 1];
      x[r] = dot;
    }
  } else if (!trans) {  // :46-77, rows from n-1 down
    for (int64_t i = 0; i < n; ++i) {
      const int64_t r = n - i - 1;
      const int64_t cnt = i < n - k ? k : k - (i - (n - k)) - 1;
      W dot = 0;
      for (int64_t j = 0; j < cnt; ++j) dot += A(r, k - j - 1) * x[r - j - 1];
      dot += diag == 0 ? x[r] * A(r, k) : x[r];
      x[r] = dot;
    }
  } else {  // :80-141, scatter into a result vector, then copy back
    std::vector<W> res(n > 0 ? n : 1, W(0));
    if (upper) {
      for (int64_t i = 0; i < n; ++i) {
        const int64_t cnt = i < n - k ? k : k - (i - (n - k)) - 1;
        res[i] += diag == 0 ? x[i] * A(i, 0) : x[i];
        for (int64_t j = 0; j < cnt; ++j) res[i + j + 1] += A(i, j + 1) * x[i];
      }
    } else {
      for (int64_t i = 0; i < n; ++i) {
        const int64_t r = n - i - 1;
        const int64_t cnt = i < n - k ? k : k - (i - (n - k)) - 1;
        for (int64_t j = 0; j < cnt; ++j) res[r - j - 1] += A(r, k - j - 1) * x[r];
        res[r] += diag == 0 ? x[r] * A(r, k) : x[r];
      }
    }
    for (int64_t i = 0; i < n; ++i) x[i] = res[i];
  }
}

// src/level2/tbsv.py:12-129: op(A) x = b in place for a triangular band.
template <class T, class W>
void tbsv(int64_t upper, int64_t trans, int64_t n, int64_t k, T* x_, int64_t incx, const T* A_, int64_t lda,
          int64_t diag) {
  M2<T, W> A(A_, n, k + 1, lda, false);
  V1<T, W> x(x_, n, incx, true);
  if (!trans && upper) {  // :12-45, rows n-1 down to 0
    for (int64_t s = 0; s < n; ++s) {
      const int64_t r = n - s - 1;
      const int64_t cnt = s < k ? s : k;
      const W pivot = diag == 0 ? A(r, 0) : W(1);
      W dot = 0;
      for (int64_t j = 0; j < cnt; ++j) dot += A(r, j + 1) * x[r + 1 + j];
      x[r] = (x[r] - dot) / pivot;
    }
  } else if (!trans) {  // :48-81, rows 0 up to n-1
    for (int64_t i = 0; i < n; ++i) {
      const int64_t cnt = i < k ? i : k;
      W dot = 0;
      for (int64_t j = 0; j < cnt; ++j) dot += A(i, k - j - 1) * x[i - j - 1];
      const W pivot = diag == 0 ? A(i, k) : W(1);
      x[i] = (x[i] - dot) / pivot;
    }
  } else if (upper) {  // :84-105, rows 0 up to n-1
    for (int64_t i = 0; i < n; ++i) {
      W dot = 0;
      for (int64_t j = 0; j < k; ++j)
        if (i - j - 1 >= 0) dot += A(i - j - 1, j + 1) * x[i - j - 1];
      const W pivot = diag == 0 ? A(i, 0) : W(1);
      x[i] = (x[i] - dot) / pivot;
    }
  } else {  // :108-129, rows n-1 down to 0
    for (int64_t i = 0; i < n; ++i) {
      const int64_t r = n - i - 1;
      const W pivot = diag == 0 ? A(r, k) : W(1);
      W dot = 0;
      for (int64_t j = 0; j < k; ++j)
        if (n - i + j < n) dot += A(n - i + j, k - j - 1) * x[n - i + j];
      x[r] = (x[r] - dot) / pivot;
    }
  }
}

// src/level3/syr2k.py:14-39.  A and B are N x K (NoTrans) or K x N (Trans).
template <class T, class W>
void syr2k(int64_t uplo, int64_t trans, int64_t N, int64_t K, T alpha_, const T* A_, int64_t lda, const T* B_,
           int64_t ldb, T* C_, int64_t ldc) {
  const int64_t r = trans == kNoTrans ? N : K, c = trans == kNoTrans ? K : N;
  M2<T, W> A(A_, r, c, lda, false);
  M2<T, W> B(B_, r, c, ldb, false);
  M2<T, W> C(C_, N, N, ldc, true);
  W alpha = alpha_;
  for (int64_t i = 0; i < N; ++i)
    for (int64_t j = 0; j < N; ++j)
      if ((uplo == kUpper && j >= i) || (uplo != kUpper && j < i + 1))
        for (int64_t k = 0; k < K; ++k) {
          if (trans == kNoTrans)
            C(i, j) += alpha * (A(i, k) * B(j, k) + B(i, k) * A(j, k));
          else
            C(i, j) += alpha * (A(k, i) * B(k, j) + B(k, i) * A(k, j));
        }
}

// src/level3/trmm.py:12-83, the four Left procs, in place on B (m x n); the
// diagonal of A is read only when diag == 0 (the wrapper passes Diag == Unit).
// Upper NoTrans (:12-23) and Lower Trans (:72-83) walk the rows downwards and
// read rows below i, which are not yet updated; Lower NoTrans (:26-46) and
// Upper Trans (:49-69) walk upwards from the last row and read rows above it.
template <class T, class W>
void trmm(int64_t upper, int64_t trans, int64_t m, int64_t n, T alpha_, const T* A_, int64_t lda, T* B_,
          int64_t ldb, int64_t diag) {
  M2<T, W> A(A_, m, m, lda, false);
  M2<T, W> B(B_, m, n, ldb, true);
  W alpha = alpha_;
  const bool down = (upper != 0) != (trans != 0);  // Upper NoTrans or Lower Trans
  for (int64_t j = 0; j < n; ++j)
    for (int64_t i = 0; i < m; ++i) {
      if (down) {
        if (diag == 0) B(i, j) = A(i, i) * B(i, j);
        for (int64_t k = 0; k < m - i - 1; ++k)
          B(i, j) += (trans ? A(i + k + 1, i) : A(i, i + k + 1)) * B(i + k + 1, j);
        B(i, j) = alpha * B(i, j);
      } else {
        const int64_t r = m - i - 1;
        const W org = B(r, j);
        B(r, j) = 0;
        for (int64_t k = 0; k < m - i - 1; ++k) B(r, j) += (trans ? A(k, r) : A(r, k)) * B(k, j);
        if (diag == 0)
          B(r, j) += A(r, r) * org;
        else
          B(r, j) += org;
        B(r, j) = alpha * B(r, j);
      }
    }
}

// ------------------------------------------------------------ C exports
// oracle_s*: float strict, oracle_d*: double strict, oracle_x*: double
// storage with long-double arithmetic (the truth oracle).

#define ORACLE_TYPED(P, T, W)                                                                       \
  extern "C" T oracle_##P##asum(int64_t n, const T* x, int64_t incx) {                              \
    return (T)asum<T, W>(n, x, incx);                                                               \
  }                                                                                                 \
  extern "C" void oracle_##P##axpy(int64_t n, T a, const T* x, int64_t ix, T* y, int64_t iy) {      \
    axpy<T, W>(n, a, x, ix, y, iy);                                                                 \
  }                                                                                                 \
  extern "C" void oracle_##P##axpy_alpha_1(int64_t n, const T* x, int64_t ix, T* y, int64_t iy) {   \
    axpy_alpha_1<T, W>(n, x, ix, y, iy);                                                            \
  }                                                                                                 \
  extern "C" T oracle_##P##dot(int64_t n, const T* x, int64_t ix, const T* y, int64_t iy) {         \
    return (T)dot<T, W>(n, x, ix, y, iy);                                                           \
  }                                                                                                 \
  extern "C" void oracle_##P##copy(int64_t n, const T* x, int64_t ix, T* y, int64_t iy) {           \
    copy<T>(n, x, ix, y, iy);                                                                       \
  }                                                                                                 \
  extern "C" void oracle_##P##swap(int64_t n, T* x, int64_t ix, T* y, int64_t iy) {                 \
    swap<T>(n, x, ix, y, iy);                                                                       \
  }                                                                                                 \
  extern "C" void oracle_##P##rot(int64_t n, T* x, int64_t ix, T* y, int64_t iy, T c, T s) {        \
    rot<T, W>(n, x, ix, y, iy, c, s);                                                               \
  }                                                                                                 \
  extern "C" void oracle_##P##rotm(int flag, int64_t n, T* x, int64_t ix, T* y, int64_t iy,         \
                                   const T* H) {                                                    \
    rotm<T, W>(flag, n, x, ix, y, iy, H);                                                           \
  }                                                                                                 \
  extern "C" void oracle_##P##scal(int64_t n, T a, T* x, int64_t ix) { scal<T, W>(n, a, x, ix); }   \
  extern "C" void oracle_##P##scal_alpha_0(int64_t n, T* x, int64_t ix) { scal_alpha_0<T>(n, x, ix); } \
  extern "C" void oracle_##P##gemv(int64_t tr, int64_t m, int64_t n, T a, T b, const T* A,          \
                                   int64_t lda, const T* x, int64_t ix, T* y, int64_t iy) {         \
    gemv<T, W>(tr, m, n, a, b, A, lda, x, ix, y, iy);                                               \
  }                                                                                                 \
  extern "C" void oracle_##P##ger(int64_t m, int64_t n, T a, const T* x, int64_t ix, const T* y,    \
                                  int64_t iy, T* A, int64_t lda) {                                  \
    ger<T, W>(m, n, a, x, ix, y, iy, A, lda);                                                       \
  }                                                                                                 \
  extern "C" void oracle_##P##symv(int64_t up, int64_t n, T a, const T* A, int64_t lda,             \
                                   const T* x, int64_t ix, T b, T* y, int64_t iy) {                 \
    symv<T, W>(up, n, a, A, lda, x, ix, b, y, iy);                                                  \
  }                                                                                                 \
  extern "C" void oracle_##P##syr(int64_t up, int64_t n, T a, const T* x, int64_t ix, T* A,         \
                                  int64_t lda) {                                                    \
    syr<T, W>(up, n, a, x, ix, A, lda);                                                             \
  }                                                                                                 \
  extern "C" void oracle_##P##syr2(int64_t up, int64_t n, T a, const T* x, int64_t ix, const T* y,  \
                                   int64_t iy, T* A, int64_t lda) {                                 \
    syr2<T, W>(up, n, a, x, ix, y, iy, A, lda);                                                     \
  }                                                                                                 \
  extern "C" void oracle_##P##trmv(int64_t up, int64_t tr, int64_t dg, int64_t n, T* x, int64_t ix, \
                                   const T* A, int64_t lda) {                                       \
    trmv<T, W>(up, tr, dg, n, x, ix, A, lda);                                                       \
  }                                                                                                 \
  extern "C" void oracle_##P##trsv(int64_t up, int64_t tr, int64_t dg, int64_t n, T* x, int64_t ix, \
                                   const T* A, int64_t lda) {                                       \
    trsv<T, W>(up, tr, dg, n, x, ix, A, lda);                                                       \
  }                                                                                                 \
  extern "C" void oracle_##P##gemm(int64_t ta, int64_t tb, int64_t M, int64_t N, int64_t K, T a,    \
                                   const T* A, int64_t lda, const T* B, int64_t ldb, T* C,          \
                                   int64_t ldc) {                                                   \
    gemm<T, W>(ta, tb, M, N, K, a, A, lda, B, ldb, C, ldc);                                         \
  }                                                                                                 \
  extern "C" void oracle_##P##mscal(int64_t M, int64_t N, T a, T* A, int64_t lda) {                 \
    mscal<T, W>(M, N, a, A, lda);                                                                   \
  }                                                                                                 \
  extern "C" void oracle_##P##trmscal(int64_t up, int64_t N, T a, T* A, int64_t lda) {              \
    trmscal<T, W>(up, N, a, A, lda);                                                                \
  }                                                                                                 \
  extern "C" void oracle_##P##symm(int64_t sd, int64_t up, int64_t M, int64_t N, T a, const T* A,   \
                                   int64_t lda, const T* B, int64_t ldb, T* C, int64_t ldc) {       \
    symm<T, W>(sd, up, M, N, a, A, lda, B, ldb, C, ldc);                                            \
  }                                                                                                 \
  extern "C" void oracle_##P##syrk(int64_t up, int64_t tr, int64_t N, int64_t K, T a, const T* A,   \
                                   int64_t lda, T* C, int64_t ldc) {                                \
    syrk<T, W>(up, tr, N, K, a, A, lda, C, ldc);                                                    \
  }                                                                                                 \
  extern "C" void oracle_##P##syr2k(int64_t up, int64_t tr, int64_t N, int64_t K, T a, const T* A,  \
                                    int64_t lda, const T* B, int64_t ldb, T* C, int64_t ldc) {      \
    syr2k<T, W>(up, tr, N, K, a, A, lda, B, ldb, C, ldc);                                           \
  }                                                                                                 \
  extern "C" void oracle_##P##gbmv(int64_t m, int64_t n, int64_t kl, int64_t ku, T a, T b,          \
                                   const T* A, int64_t lda, const T* x, int64_t ix, T* y, int64_t iy) { \
    gbmv<T, W>(m, n, kl, ku, a, b, A, lda, x, ix, y, iy);                                            \
  }                                                                                                 \
  extern "C" void oracle_##P##sbmv_scal_y(int64_t n, T b, T* y, int64_t iy) {                       \
    sbmv_scal_y<T, W>(n, b, y, iy);                                                                 \
  }                                                                                                 \
  extern "C" void oracle_##P##sbmv(int64_t up, int64_t n, int64_t k, T a, const T* A, int64_t lda,  \
                                   const T* x, int64_t ix, T* y, int64_t iy) {                      \
    sbmv<T, W>(up, n, k, a, A, lda, x, ix, y, iy);                                                  \
  }                                                                                                 \
  extern "C" void oracle_##P##tbmv(int64_t up, int64_t tr, int64_t n, int64_t k, T* x, int64_t ix,  \
                                   const T* A, int64_t lda, int64_t diag) {                         \
    tbmv<T, W>(up, tr, n, k, x, ix, A, lda, diag);                                                  \
  }                                                                                                 \
  extern "C" void oracle_##P##tbsv(int64_t up, int64_t tr, int64_t n, int64_t k, T* x, int64_t ix,  \
                                   const T* A, int64_t lda, int64_t diag) {                         \
    tbsv<T, W>(up, tr, n, k, x, ix, A, lda, diag);                                                  \
  }                                                                                                 \
  extern "C" void oracle_##P##trmm(int64_t upper, int64_t trans, int64_t m, int64_t n, T a,         \
                                   const T* A, int64_t lda, T* B, int64_t ldb, int64_t diag) {      \
    trmm<T, W>(upper, trans, m, n, a, A, lda, B, ldb, diag);                                        \
  }

ORACLE_TYPED(s, float, float)
ORACLE_TYPED(d, double, double)
ORACLE_TYPED(x, double, long double)

extern "C" double oracle_dsdot(int64_t n, const float* x, int64_t ix, const float* y, int64_t iy) {
  return dsdot<double>(n, x, ix, y, iy);
}
extern "C" float oracle_sdsdot(int64_t n, float sb, const float* x, int64_t ix, const float* y,
                               int64_t iy) {
  return sdsdot<double>(n, sb, x, ix, y, iy);
}
extern "C" double oracle_xdsdot(int64_t n, const float* x, int64_t ix, const float* y, int64_t iy) {
  return (double)dsdot<long double>(n, x, ix, y, iy);
}
