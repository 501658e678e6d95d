// The level-3 product entry shared by every kernel family built on the MFMA
// tile engine (plain level-3, batched, split precision).
#pragma once

#include "gemm_core.hpp"
#include "runtime.hpp"

namespace exoblas {
namespace gemm {

// C[i, j] += alpha * sum_k X_A(i, k) X_B(j, k),  i < M, j < N, k < K, on the
// `tri` part of C (kFull, kLower, kUpper).  Enqueued on c.stream.
template <class T>
void product(Ctx& c, int tri, int64_t M, int64_t N, int64_t K, T alpha, const Operand<T>& A, const Operand<T>& B,
             T* Cp, int64_t ldc);

// Operand descriptor with the 16-byte-load flag filled in.
template <class T>
Operand<T> operand(const T* p, int64_t ld, int form);

}  // namespace gemm
}  // namespace exoblas
