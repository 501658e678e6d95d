// Triangular solve op(A) x = b for gfx950: one kernel launch, 64-row blocks,
// decoupled look-back between workgroups.
//
// The four (Uplo, TransA) combinations reduce to a forward (lower/NoTrans,
// upper/Trans) or a backward (upper/NoTrans, lower/Trans) substitution over
// M(i, j) = TransA ? A[j, i] : A[i, j].  Workgroups take a ticket in start
// order; ticket p owns block p of the processing order.  It streams the
// blocks of its row panel left of the diagonal (forward) or right of it
// (backward), each after the owning workgroup has published that block of x,
// so every workgroup reads its panel while the chain of diagonal solves is
// still running; then wave 0 solves the 64 x 64 diagonal block by
// substitution and publishes its x block (release store of the call's epoch).
// A workgroup waits only on smaller tickets, which were taken by workgroups
// already running, so the grid cannot deadlock even when it does not fit on
// the chip at once.  The wait is bounded: past the spin limit the kernel
// records a timeout and finishes.
#include "device.hpp"
#include "ops.hpp"

namespace exoblas {
namespace {

constexpr int kB = 64;    // rows per block
constexpr int kNT = 256;  // threads per workgroup
constexpr int64_t kNoTrans = 111;
constexpr int64_t kLowerV = 122;
constexpr int64_t kUnitV = 132;
constexpr long long kSpinLimit = 1ll << 22;

// flags[0]: ticket, flags[1]: timeout (epoch on timeout), flags[2 + b]: block b done
constexpr int kFlagBase = 2;

template <class T, bool TRANS>
__device__ __forceinline__ T mat(const T* A, int64_t lda, int64_t i, int64_t j) {
  return TRANS ? A[j * lda + i] : A[i * lda + j];
}

__device__ __forceinline__ void wait_epoch(unsigned* f, unsigned epoch, unsigned* err) {
  long long it = 0;
  while (__hip_atomic_load(f, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) != epoch) {
    if (++it > kSpinLimit) {
      __hip_atomic_store(err, epoch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      return;
    }
    __builtin_amdgcn_s_sleep(1);
  }
}

template <class T, bool FWD, bool TRANS, bool UNIT>
__global__ __launch_bounds__(kNT) void trsv_kernel(int64_t n, const T* __restrict__ A, int64_t lda, T* x,
                                                   unsigned* flags, unsigned epoch) {
  __shared__ T xs[kB];
  __shared__ T red[kNT / kB][kB];
  __shared__ T diag[kB][kB + 1];
  __shared__ int s_p;
  const int nb = (int)((n + kB - 1) / kB);
  const int t = threadIdx.x;
  if (t == 0) {
    const unsigned p = atomicAdd(flags, 1u);
    if ((int)p == nb - 1) atomicExch(flags, 0u);  // the last ticket: every ticket is taken
    s_p = (int)p;
  }
  __syncthreads();
  const int p = s_p;
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
    if (t == 0) wait_epoch(flags + kFlagBase + bj, epoch, flags + 1);
    __syncthreads();
    if (t < kB) xs[t] = (bj * kB + t < n) ? x[bj * kB + t] : T(0);
    __syncthreads();
#pragma unroll
    for (int e = 0; e < 16; ++e) acc += m[e] * xs[cq * 16 + e];
  }
  red[cq][r] = acc;
  for (int idx = t; idx < kB * kB; idx += kNT) {
    // NoTrans reads rows of A, Trans reads columns: keep lanes on consecutive addresses
    const int i = TRANS ? idx % kB : idx / kB;
    const int j = TRANS ? idx / kB : idx % kB;
    const int64_t ai = r0 + i, aj = r0 + j;
    T v = (ai < n && aj < n) ? mat<T, TRANS>(A, lda, ai, aj) : T(i == j ? 1 : 0);
    diag[i][j] = v;
  }
  __syncthreads();
  if (t < kB) {
    const int lane = t;
    T v = T(0);
    if (r0 + lane < n) {
      T s = T(0);
#pragma unroll
      for (int w = 0; w < kNT / kB; ++w) s += red[w][lane];
      v = x[r0 + lane] - s;
    }
    for (int s = 0; s < kB; ++s) {
      const int k = FWD ? s : kB - 1 - s;
      T xk = __shfl(v, k, kB);
      if (!UNIT) xk = xk / diag[k][k];
      if (lane == k)
        v = xk;
      else if (FWD ? lane > k : lane < k)
        v -= diag[lane][k] * xk;
    }
    if (r0 + lane < n) x[r0 + lane] = v;
    __threadfence();
    if (lane == 0) __hip_atomic_store(flags + kFlagBase + bi, epoch, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
  }
}

template <class T, bool FWD, bool TRANS>
void launch_trsv(Ctx& c, bool unit, int64_t n, const T* A, int64_t lda, T* x) {
  const int64_t nb = ceil_div(n, kB);
  if (nb > 0x7fffffff) throw Error(kInvalidSize, "trsv: n too large");
  unsigned epoch = 0;
  unsigned* flags = c.epoch_flags((size_t)nb + kFlagBase, &epoch);
  if (unit)
    hipLaunchKernelGGL((trsv_kernel<T, FWD, TRANS, true>), dim3((unsigned)nb), dim3(kNT), 0, c.stream, n, A, lda, x,
                       flags, epoch);
  else
    hipLaunchKernelGGL((trsv_kernel<T, FWD, TRANS, false>), dim3((unsigned)nb), dim3(kNT), 0, c.stream, n, A, lda,
                       x, flags, epoch);
  EXO_LAUNCH_CHECK();
}

}  // namespace

namespace ops {

// src/level2/trsv.py:10-52
template <class T>
void trsv_rm(Ctx& c, int64_t Uplo, int64_t TransA, int64_t Diag, int64_t n, T* x, int64_t incx, const T* A,
             int64_t lda) {
  if (n <= 0) return;
  T* xs = x;
  if (incx != 1) {  // solve on a contiguous copy (scratch slot 1), then scatter back
    xs = static_cast<T*>(c.scratch(sizeof(T) * n, 1));
    ops::copy<T>(c, n, x, incx, xs, 1);
  }
  const bool trans = TransA != kNoTrans;
  const bool fwd = (Uplo == kLowerV) != trans;
  const bool unit = Diag == kUnitV;
  if (fwd && !trans) launch_trsv<T, true, false>(c, unit, n, A, lda, xs);
  if (fwd && trans) launch_trsv<T, true, true>(c, unit, n, A, lda, xs);
  if (!fwd && !trans) launch_trsv<T, false, false>(c, unit, n, A, lda, xs);
  if (!fwd && trans) launch_trsv<T, false, true>(c, unit, n, A, lda, xs);
  if (incx != 1) ops::copy<T>(c, n, xs, 1, x, incx);
}

template void trsv_rm<float>(Ctx&, int64_t, int64_t, int64_t, int64_t, float*, int64_t, const float*, int64_t);
template void trsv_rm<double>(Ctx&, int64_t, int64_t, int64_t, int64_t, double*, int64_t, const double*, int64_t);

}  // namespace ops
}  // namespace exoblas
