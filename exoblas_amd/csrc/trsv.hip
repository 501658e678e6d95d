// Triangular solve op(A) x = b for gfx950: one kernel launch, one workgroup
// per 64-row block, decoupled look-back between workgroups (trsv_core.hpp).
// Workgroups take a ticket in start order and ticket p runs processing step
// p; a step waits only on smaller tickets, which were taken by workgroups
// already running, so the grid cannot deadlock even when it does not fit on
// the chip at once.
#include "device.hpp"
#include "ops.hpp"
#include "trsv_core.hpp"

namespace exoblas {
namespace {

constexpr int64_t kNoTrans = 111;
constexpr int64_t kLowerV = 122;
constexpr int64_t kUnitV = 132;
using trsv::kB;
using trsv::kNT;

// flags[0]: ticket, flags[2 ..]: the publication words of trsv_core.hpp
// (8-byte aligned).  The marker is the stream's launch generation
// (tagged.hpp), so a captured solve replays.  err: the stream's timeout word.
constexpr int kFlagBase = 2;

// BAND: A is a band of kb off-diagonals (trsv_core.hpp, mat).
template <class T, bool FWD, bool TRANS, bool UNIT, int KB, bool BAND>
__global__ __launch_bounds__(4 * KB) void trsv_kernel(int64_t n, const T* __restrict__ A, int64_t lda, T* x,
                                                      unsigned* flags, unsigned* gen, unsigned* err, int64_t kb) {
  __shared__ __align__(16) trsv::Smem<T, T, KB> sm;
  __shared__ __align__(16) T mb[KB][trsv::mb_stride<T, KB>()];
  __shared__ __align__(16) T mb2[KB][trsv::mb_stride<T, KB>()];
  __shared__ int s_p;
  __shared__ unsigned s_marker;
  const int nb = (int)((n + KB - 1) / KB);
  if (threadIdx.x == 0) {
    unsigned marker;
    s_p = (int)tagged::generation_ticket(flags, (unsigned)nb, gen, &marker);
    s_marker = marker;
  }
  __syncthreads();
  // tickets come from the stream's counter: a ticket past the last block
  // (only if launches sharing the counter overlapped) runs no step
  if (s_p >= nb) return;
  trsv::step<T, FWD, TRANS, UNIT, KB, true, BAND>(s_p, n, A, lda, x, flags + kFlagBase, s_marker, err, sm, mb, mb2,
                                                  kb);
}

// 64-row blocks.  (128-row blocks halve the number of dependent steps but
// measured 1.8x slower per solve on MI355X: each step's critical path grows
// by more than the chain shrinks.)
template <class T, bool FWD, bool TRANS, int KB, bool BAND>
void launch_trsv_kb(Ctx& c, bool unit, int64_t n, const T* A, int64_t lda, T* x, int64_t kb) {
  const int64_t nb = ceil_div(n, KB);
  if (nb > 0x7fffffff) throw Error(kInvalidSize, "trsv: n too large");
  unsigned epoch = 0;  // (not used: the marker comes from the launch generation)
  unsigned* flags = c.epoch_flags((size_t)trsv::pub_words<T>(n) + kFlagBase, &epoch);
  unsigned* gen = c.counters() + Ctx::kGeneration;
  if (unit)
    hipLaunchKernelGGL((trsv_kernel<T, FWD, TRANS, true, KB, BAND>), dim3((unsigned)nb), dim3(4 * KB), 0, c.stream,
                       n, A, lda, x, flags, gen, c.timeout_word(), kb);
  else
    hipLaunchKernelGGL((trsv_kernel<T, FWD, TRANS, false, KB, BAND>), dim3((unsigned)nb), dim3(4 * KB), 0, c.stream,
                       n, A, lda, x, flags, gen, c.timeout_word(), kb);
  EXO_LAUNCH_CHECK();
}

template <class T, bool FWD, bool TRANS, bool BAND = false>
void launch_trsv(Ctx& c, bool unit, int64_t n, const T* A, int64_t lda, T* x, int64_t kb = 0) {
  launch_trsv_kb<T, FWD, TRANS, kB, BAND>(c, unit, n, A, lda, x, kb);
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

// Triangular band solve on the block solver: a band of kb off-diagonals in
// row-major band storage with the diagonal in column c0 (src/level2/tbsv.py:
// 12-129) is, entry for entry, the dense triangle at A + c0 with leading
// dimension lda - 1 (entry (i, j) at i * lda + c0 + (j - i)), restricted to
// |i - j| <= kb.  fwd: rows in increasing order; trans: op(A) = A^T.
template <class T>
void trsv_band_rm(Ctx& c, bool fwd, bool trans, bool unit, int64_t n, int64_t kb, const T* A, int64_t lda,
                  int64_t c0, T* x, int64_t incx) {
  if (n <= 0) return;
  if (lda < 2) throw Error(kInvalidSize, "trsv_band: lda must exceed 1");
  T* xs = x;
  if (incx != 1) {  // solve on a contiguous copy (scratch slot 1), then scatter back
    xs = static_cast<T*>(c.scratch(sizeof(T) * n, 1));
    ops::copy<T>(c, n, x, incx, xs, 1);
  }
  const T* Ab = A + c0;
  const int64_t ld = lda - 1;
  if (fwd && !trans) launch_trsv<T, true, false, true>(c, unit, n, Ab, ld, xs, kb);
  if (fwd && trans) launch_trsv<T, true, true, true>(c, unit, n, Ab, ld, xs, kb);
  if (!fwd && !trans) launch_trsv<T, false, false, true>(c, unit, n, Ab, ld, xs, kb);
  if (!fwd && trans) launch_trsv<T, false, true, true>(c, unit, n, Ab, ld, xs, kb);
  if (incx != 1) ops::copy<T>(c, n, xs, 1, x, incx);
}

template void trsv_rm<float>(Ctx&, int64_t, int64_t, int64_t, int64_t, float*, int64_t, const float*, int64_t);
template void trsv_band_rm<float>(Ctx&, bool, bool, bool, int64_t, int64_t, const float*, int64_t, int64_t, float*,
                                  int64_t);
template void trsv_band_rm<double>(Ctx&, bool, bool, bool, int64_t, int64_t, const double*, int64_t, int64_t,
                                   double*, int64_t);
template void trsv_rm<double>(Ctx&, int64_t, int64_t, int64_t, int64_t, double*, int64_t, const double*, int64_t);

}  // namespace ops
}  // namespace exoblas
