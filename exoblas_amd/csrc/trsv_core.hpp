// Block step of the triangular solve op(A) x = b, shared by the standalone
// trsv kernel and the batched kernel.
//
// The four (Uplo, TransA) combinations reduce to a forward (lower/NoTrans,
// upper/Trans) or a backward (upper/NoTrans, lower/Trans) substitution over
// M(i, j) = TransA ? A[j, i] : A[i, j], in 64-row blocks taken in processing
// order p = 0 .. nb-1 (block p for forward, nb-1-p for backward).  The
// workgroup that runs step p streams the blocks of its row panel on the
// solved side of the diagonal, each after the step that owns it has published
// its part of x (a release store of `marker` into done[block]), so every
// workgroup reads its panel while the chain of diagonal solves is still
// running; then wave 0 solves the 64 x 64 diagonal block by substitution and
// publishes its x block.  Waits are bounded: past the spin limit the step
// records `marker` in *err and finishes.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace exoblas {
namespace trsv {

constexpr int kB = 64;    // rows per block
constexpr int kNT = 256;  // threads per workgroup
constexpr long long kSpinLimit = 1ll << 22;

template <class T>
struct Smem {
  T xs[kB];
  T red[kNT / kB][kB];
  T diag[kB][kB + 1];
};

template <class T, bool TRANS>
__device__ __forceinline__ T mat(const T* A, int64_t lda, int64_t i, int64_t j) {
  return TRANS ? A[j * lda + i] : A[i * lda + j];
}

__device__ __forceinline__ void wait_marker(const unsigned* f, unsigned marker, unsigned* err) {
  long long it = 0;
  while (__hip_atomic_load(f, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) != marker) {
    if (++it > kSpinLimit) {
      __hip_atomic_store(err, marker, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      return;
    }
    __builtin_amdgcn_s_sleep(1);
  }
}

// One processing step p of the solve; called by all kNT threads of a workgroup.
template <class T, bool FWD, bool TRANS, bool UNIT>
__device__ __forceinline__ void step(int p, int64_t n, const T* __restrict__ A, int64_t lda, T* x, unsigned* done,
                                     unsigned marker, unsigned* err, Smem<T>& sm) {
  const int nb = (int)((n + kB - 1) / kB);
  const int t = threadIdx.x;
  const int64_t bi = FWD ? p : nb - 1 - p;
  const int64_t r0 = bi * kB;
  // NoTrans: 4 threads per row, 16 consecutive columns each (contiguous
  // loads); Trans: one column of 64 rows per wave, 16 columns per wave
  // (consecutive lanes read consecutive addresses).
  const int r = TRANS ? t % kB : t / 4;
  const int cq = TRANS ? t / kB : t % 4;
  const int64_t gi = r0 + r;
  T acc = T(0);
  for (int q = 0; q < p; ++q) {
    const int64_t bj = FWD ? q : nb - 1 - q;
    const int64_t c0 = bj * kB + cq * 16;
    T m[16];
#pragma unroll
    for (int e = 0; e < 16; ++e) {
      const int64_t gj = c0 + e;
      m[e] = (gi < n && gj < n) ? mat<T, TRANS>(A, lda, gi, gj) : T(0);
    }
    if (t == 0) wait_marker(done + bj, marker, err);
    __syncthreads();
    if (t < kB) sm.xs[t] = (bj * kB + t < n) ? x[bj * kB + t] : T(0);
    __syncthreads();
#pragma unroll
    for (int e = 0; e < 16; ++e) acc += m[e] * sm.xs[cq * 16 + e];
  }
  sm.red[cq][r] = acc;
  for (int idx = t; idx < kB * kB; idx += kNT) {
    // NoTrans reads rows of A, Trans reads columns: keep lanes on consecutive addresses
    const int i = TRANS ? idx % kB : idx / kB;
    const int j = TRANS ? idx / kB : idx % kB;
    const int64_t ai = r0 + i, aj = r0 + j;
    sm.diag[i][j] = (ai < n && aj < n) ? mat<T, TRANS>(A, lda, ai, aj) : T(i == j ? 1 : 0);
  }
  __syncthreads();
  if (t < kB) {
    const int lane = t;
    T v = T(0);
    if (r0 + lane < n) {
      T s = T(0);
#pragma unroll
      for (int w = 0; w < kNT / kB; ++w) s += sm.red[w][lane];
      v = x[r0 + lane] - s;
    }
    for (int s = 0; s < kB; ++s) {
      const int k = FWD ? s : kB - 1 - s;
      T xk = __shfl(v, k, kB);
      if (!UNIT) xk = xk / sm.diag[k][k];
      if (lane == k)
        v = xk;
      else if (FWD ? lane > k : lane < k)
        v -= sm.diag[lane][k] * xk;
    }
    if (r0 + lane < n) x[r0 + lane] = v;
    __threadfence();
    if (lane == 0) __hip_atomic_store(done + bi, marker, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
  }
  __syncthreads();  // the shared buffers are reused by the caller's next step
}

}  // namespace trsv
}  // namespace exoblas
