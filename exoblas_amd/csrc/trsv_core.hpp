// Block step of the triangular solve op(A) x = b, shared by the standalone
// trsv kernel and the batched kernel.
//
// The four (Uplo, TransA) combinations reduce to a forward (lower/NoTrans,
// upper/Trans) or a backward (upper/NoTrans, lower/Trans) substitution over
// M(i, j) = TransA ? A[j, i] : A[i, j], in 64-row blocks taken in processing
// order p = 0 .. nb-1 (block p for forward, nb-1-p for backward).  The
// workgroup that runs step p
//   1. loads its diagonal block (in processing order it is lower triangular)
//      and inverts it in LDS, before waiting on anything, so the inversion
//      overlaps the earlier steps of the chain;
//   2. streams the blocks of its row panel on the solved side of the
//      diagonal, each after the step that owns it has published its part of x;
//   3. forms v = b - (panel . x) and x_block = inv(D) v with all its threads,
//      so the serial part of the chain is one small mat-vec per block.
// The standalone solve (PF) also folds the two previous blocks into the
// inverse ahead of time, G1 = inv(D) M(p, p-1) and G2 = inv(D) M(p, p-2)
// (matrix-core products, fold_g): the panel covers blocks 0 .. p-3,
// y0 = inv(D)(b - panel . x), and the step ends x_block = y0 - G2 x_{p-2}
// - G1 x_{p-1}, so once x_{p-1} appears the chain's critical path is one
// block mat-vec.  The diagonal-block inverse takes 16 x 16 blocks by
// right-looking substitution, then two pairwise merges on the matrix cores
// (merge_mfma); every load of the step's own blocks is issued before any is
// used.
// Publication of x: the tagged words of tagged.hpp, one value per row.
// Waits are bounded: past the spin limit the step records `marker` in *err
// and finishes.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "gemm_core.hpp"
#include "tagged.hpp"

namespace exoblas {
namespace trsv {

constexpr int kB = 64;    // default rows per block (batched and mixed solves)
constexpr int kNT = 256;  // threads per workgroup at kB

// TI: precision of the diagonal-block inverse and of the block mat-vec.
// KB: rows per block; a block's workgroup has 4 * KB threads (4 per row).
template <class T, int N>
struct alignas(N * sizeof(T)) VecN {
  T v[N];
};

// Row stride of the previous-block staging buffer `mb` (and of G, stored
// over it): a multiple of 16 bytes, so a thread's 16 consecutive entries of
// a row are 16-byte LDS reads.
template <class T, int KB>
constexpr int mb_stride() {
  return KB + 16 / (int)sizeof(T);
}

template <class T, class TI = T, int KB = kB>
struct Smem {
  TI inv[KB][KB + 1];
  T red[4][KB];
  T xs[2][KB];
  TI vs[KB];
};

// 32-bit words of publication space a solve of order n needs (see above).
template <class T>
__host__ __device__ inline int64_t pub_words(int64_t n) {
  return tagged::words<T>(n);
}

// Wave 0 of the workgroup fetches the published x block bj into xs.
template <class T, int KB = kB>
__device__ __forceinline__ void fetch_block(int64_t n, int64_t bj, unsigned* pub, unsigned marker, unsigned* err,
                                            T* xs) {
  const int t = threadIdx.x;
  if (t < KB) {
    const int64_t g = bj * KB + t;
    xs[t] = g < n ? tagged::read<T>(pub, g, marker, err) : T(0);
  }
}

// Publishes x[g] = v (called by the thread that owns row g).
template <class T>
__device__ __forceinline__ void publish_value(int64_t g, T v, T* x, unsigned* pub, unsigned marker) {
  x[g] = v;
  tagged::publish<T>(pub, g, v, marker);
}

// M(i, j) of a dense triangle; with BAND, of a band of kb off-diagonals
// whose storage the dense addressing reaches (row-major band storage seen
// with leading dimension lda - 1 from the diagonal's column): entries
// farther than kb from the diagonal are zero and not read.
template <class T, bool TRANS, bool BAND = false>
__device__ __forceinline__ T mat(const T* A, int64_t lda, int64_t i, int64_t j, int64_t kb = 0) {
  if constexpr (BAND) {
    if (i - j > kb || j - i > kb) return T(0);
  }
  return TRANS ? A[j * lda + i] : A[i * lda + j];
}


// Natural row of local index i of block bi (local order = processing order).
template <bool FWD, int KB = kB>
__device__ __forceinline__ int64_t row_of(int64_t r0, int i) {
  return FWD ? r0 + i : r0 + (KB - 1 - i);
}

// One level of the pairwise merges of invert_diag (block size SZ -> 2 SZ),
// then the next level.
template <class TI, int KB, int SZ>
__device__ __forceinline__ void merge_level(TI (&L)[KB][KB + 1], int t) {
  if constexpr (SZ < KB) {
    constexpr int NT = 4 * KB;
    constexpr int PER = (KB / (2 * SZ)) * SZ * SZ / NT;  // outputs per thread
    static_assert(PER >= 1 && (KB / (2 * SZ)) * SZ * SZ % NT == 0, "merge layout");
    int base[PER], r[PER], c[PER];
#pragma unroll
    for (int e = 0; e < PER; ++e) {
      const int idx = t + NT * e;
      base[e] = (idx / (SZ * SZ)) * 2 * SZ;
      r[e] = (idx % (SZ * SZ)) / SZ;
      c[e] = idx % SZ;
    }
    TI pv[PER];
#pragma unroll
    for (int e = 0; e < PER; ++e) {  // (B inv(A))[r][c]
      TI acc = TI(0);
#pragma unroll
      for (int k = 0; k < SZ; ++k) acc += L[base[e] + SZ + r[e]][base[e] + k] * L[base[e] + k][base[e] + c[e]];
      pv[e] = acc;
    }
    __syncthreads();
#pragma unroll
    for (int e = 0; e < PER; ++e) L[base[e] + SZ + r[e]][base[e] + c[e]] = pv[e];
    __syncthreads();
#pragma unroll
    for (int e = 0; e < PER; ++e) {  // -(inv(C) P)[r][c]
      TI acc = TI(0);
#pragma unroll
      for (int k = 0; k < SZ; ++k)
        acc += L[base[e] + SZ + r[e]][base[e] + SZ + k] * L[base[e] + SZ + k][base[e] + c[e]];
      pv[e] = -acc;
    }
    __syncthreads();
#pragma unroll
    for (int e = 0; e < PER; ++e) L[base[e] + SZ + r[e]][base[e] + c[e]] = pv[e];
    __syncthreads();
    merge_level<TI, KB, 2 * SZ>(L, t);
  }
}

// The merges of invert_diag on the matrix cores (v_mfma_*_16x16x4; KB = 64
// or 128 with 4 KB threads, i.e. KB / 16 waves).  Level SZ (16 -> 32 -> ..
// -> KB) merges KB / (2 SZ) pairs [A 0; B C] of inverted SZ x SZ blocks:
// P = B inv(A) (inv(A) lower triangular: K tiles J .. SZ/16 - 1), P goes
// over B, then -(inv(C) P) (K tiles 0 .. I) goes over P.  The 16 x 16
// output tiles of a level are dealt to the waves round-robin.  Step j of a
// 16-deep product lets lane group lg take k = out_row(lg, j): the row of
// the accumulator layout this lane holds in register j, so the tile of P a
// wave formed itself is the B operand of its diagonal K tile as it stands.
template <class TI, int KB, int SZ>
__device__ __forceinline__ void merge_mfma(TI (&L)[KB][KB + 1], int t) {
  if constexpr (SZ < KB) {
    using MF = gemm::Mfma<TI>;
    using Acc = typename MF::acc;
    constexpr int NW = KB / 16;                  // waves
    constexpr int TS = SZ / 16;                  // tiles per side of a block
    constexpr int NTILE = (KB / (2 * SZ)) * TS * TS;
    constexpr int PER = (NTILE + NW - 1) / NW;   // tiles per wave
    const int w = t / 64, lane = t % 64, lm = lane % 16, lg = lane / 16;
    Acc P[PER];
#pragma unroll
    for (int e = 0; e < PER; ++e) {
#pragma unroll
      for (int q = 0; q < 4; ++q) P[e][q] = TI(0);
      const int tau = w + NW * e;
      if (tau < NTILE) {
        const int o = 2 * SZ * (tau / (TS * TS)), I = (tau % (TS * TS)) / TS, J = tau % TS;
#pragma unroll
        for (int K = 0; K < TS; ++K) {
          if (K < J) continue;
#pragma unroll
          for (int j = 0; j < 4; ++j) {
            const int k = 16 * K + MF::out_row(lg, j);
            P[e] = MF::step(L[o + SZ + 16 * I + lm][o + k], L[o + k][o + 16 * J + lm], P[e]);
          }
        }
      }
    }
    __syncthreads();  // every wave has read B
#pragma unroll
    for (int e = 0; e < PER; ++e) {
      const int tau = w + NW * e;
      if (tau < NTILE) {
        const int o = 2 * SZ * (tau / (TS * TS)), I = (tau % (TS * TS)) / TS, J = tau % TS;
#pragma unroll
        for (int q = 0; q < 4; ++q) L[o + SZ + 16 * I + MF::out_row(lg, q)][o + 16 * J + lm] = P[e][q];
      }
    }
    __syncthreads();
    Acc Q[PER];
#pragma unroll
    for (int e = 0; e < PER; ++e) {
#pragma unroll
      for (int q = 0; q < 4; ++q) Q[e][q] = TI(0);
      const int tau = w + NW * e;
      if (tau < NTILE) {
        const int o = 2 * SZ * (tau / (TS * TS)), I = (tau % (TS * TS)) / TS, J = tau % TS;
#pragma unroll
        for (int K = 0; K < TS; ++K) {
          if (K > I) continue;
#pragma unroll
          for (int j = 0; j < 4; ++j) {
            const int k = 16 * K + MF::out_row(lg, j);
            const TI bv = K == I ? P[e][j] : L[o + SZ + k][o + 16 * J + lm];
            Q[e] = MF::step(L[o + SZ + 16 * I + lm][o + SZ + k], bv, Q[e]);
          }
        }
      }
    }
    __syncthreads();  // every wave has read P
#pragma unroll
    for (int e = 0; e < PER; ++e) {
      const int tau = w + NW * e;
      if (tau < NTILE) {
        const int o = 2 * SZ * (tau / (TS * TS)), I = (tau % (TS * TS)) / TS, J = tau % TS;
#pragma unroll
        for (int q = 0; q < 4; ++q) L[o + SZ + 16 * I + MF::out_row(lg, q)][o + 16 * J + lm] = -Q[e][q];
      }
    }
    __syncthreads();
    merge_mfma<TI, KB, 2 * SZ>(L, t);
  }
}

// Entries of a KB x KB block each of the 4 * KB threads of a block step
// loads (KB / 4).
template <int KB>
constexpr int per_thread() {
  return KB * KB / (4 * KB);
}

// This thread's entries of the diagonal block of block bi (rows r0 ..) in
// local order, converted to TI: entry e is local (i, j) of idx = t + 4 KB e
// (NoTrans: i = idx / KB, j = idx % KB; Trans the other way round, so lanes
// read consecutive addresses either way).  Every load is issued before any
// is used (one memory round trip for the whole block, not one per entry).
template <class T, class TI, bool FWD, bool TRANS, bool UNIT, int KB = kB, bool BAND = false>
__device__ __forceinline__ void load_diag(int64_t n, const T* A, int64_t lda, int64_t r0, int t,
                                          TI (&v)[per_thread<KB>()], int64_t kb = 0) {
  constexpr int NT = 4 * KB;
#pragma unroll
  for (int e = 0; e < per_thread<KB>(); ++e) {
    const int idx = t + NT * e;
    const int a = idx / KB, b = idx % KB;
    const int i = TRANS ? b : a;  // local row
    const int j = TRANS ? a : b;  // local column
    const int64_t gi = row_of<FWD, KB>(r0, i), gj = row_of<FWD, KB>(r0, j);
    TI w = TI(0);
    if (j == i)
      w = (UNIT || gi >= n) ? TI(1) : (TI)mat<T, TRANS, BAND>(A, lda, gi, gj, kb);
    else if (j < i && gi < n && gj < n)
      w = (TI)mat<T, TRANS, BAND>(A, lda, gi, gj, kb);
    v[e] = w;
  }
}

template <class TI, bool TRANS, int KB>
__device__ __forceinline__ void invert_loaded(TI (&L)[KB][KB + 1], int t, const TI (&v)[per_thread<KB>()]);

// Loads the diagonal block of block bi (rows r0 ..) in local order, converted
// to TI, and inverts it in place (16 x 16 diagonal blocks, then two pairwise
// merges: 40 barriers instead of the 128 of a column-by-column inverse at
// KB = 64).  4 * KB threads with indices t = 0 .. 4 KB - 1 work on one block;
// every thread of the workgroup calls this (the same number of barriers), so
// a workgroup of 8 KB threads inverts two blocks at once, each half passing
// its own t and L.
template <class T, class TI, bool FWD, bool TRANS, bool UNIT, int KB = kB>
__device__ __forceinline__ void invert_diag(int64_t n, const T* A, int64_t lda, int64_t r0, TI (&L)[KB][KB + 1],
                                            int t) {
  TI v[per_thread<KB>()];
  load_diag<T, TI, FWD, TRANS, UNIT, KB>(n, A, lda, r0, t, v);
  invert_loaded<TI, TRANS, KB>(L, t, v);
}

// Writes the entries load_diag gave this thread into L and inverts L in place.
template <class TI, bool TRANS, int KB>
__device__ __forceinline__ void invert_loaded(TI (&L)[KB][KB + 1], int t, const TI (&v)[per_thread<KB>()]) {
  constexpr int NT = 4 * KB;
#pragma unroll
  for (int e = 0; e < per_thread<KB>(); ++e) {
    const int idx = t + NT * e;
    const int a = idx / KB, b = idx % KB;
    L[TRANS ? b : a][TRANS ? a : b] = v[e];
  }
  __syncthreads();
  // 1. The KB/16 diagonal 16 x 16 blocks by forward substitution, no
  //    barriers: thread 16 * blk + j computes column j of inv(D_blk) in
  //    registers, right-looking: once X[i][j] = s_i / D[i][i] is known it
  //    updates the later rows' sums s_m -= D[m][i] X[i][j], so the
  //    dependent path is one multiply-add and one multiply per row (the
  //    reciprocals of the diagonal are taken first, one per thread and
  //    passed round the 16 threads of the block; X[k][j] = 0 for k < j); D
  //    is read by broadcast from LDS, and the columns are written over D
  //    once every thread has read it.
  {
    TI xc[16];
    const int b = (t / 16) * 16, j = t % 16;
    if (t < KB) {
      const TI rj = TI(1) / L[b + j][b + j];
      TI rd[16];
#pragma unroll
      for (int i = 0; i < 16; ++i) rd[i] = __shfl(rj, i, 16);
      TI sr[16];
#pragma unroll
      for (int i = 0; i < 16; ++i) sr[i] = i == j ? TI(1) : TI(0);
#pragma unroll
      for (int i = 0; i < 16; ++i) {
        xc[i] = sr[i] * rd[i];
#pragma unroll
        for (int m = i + 1; m < 16; ++m) sr[m] -= L[b + m][b + i] * xc[i];
      }
    }
    __syncthreads();
    if (t < KB) {
#pragma unroll
      for (int i = 0; i < 16; ++i)
        if (i >= j) L[b + i][b + j] = xc[i];
    }
    __syncthreads();
  }
  // 2. Merge inverted diagonal blocks pairwise, 16 -> 32 -> ... -> KB:
  //    inv([A 0; B C]) = [inv(A) 0; -inv(C) B inv(A), inv(C)], written over
  //    B.  At KB = 64 and 128 the products run on the matrix cores (merge_mfma);
  //    otherwise they are formed in registers from LDS with compile-time
  //    bounds (the strictly upper parts of inv(A) and inv(C) hold zeros).
  if constexpr (KB == 64 || KB == 128)
    merge_mfma<TI, KB, 16>(L, t);
  else
    merge_level<TI, KB, 16>(L, t);
}

// G = inv(D) M on the matrix cores (v_mfma_*_16x16x4, FP32 or FP64), over
// M in place: with KB / 16 waves, wave w forms the 16-column tile w of G for
// all KB / 16 row tiles; inv(D) is lower triangular, so row tile I sums K
// tiles 0 .. I only.  Within a K tile, lane group lg takes k = 16 K + 4 lg + j
// in step j (A and B agree on it, which is all the product needs).  Called
// by all 4 KB threads; ends with a barrier.
template <class T, int KB>
__device__ __forceinline__ void fold_g(const T (&inv)[KB][KB + 1], T (*mb)[mb_stride<T, KB>()], int t) {
  static_assert(KB == 64 || KB == 128, "KB / 16 waves, one 16-column tile each");
  constexpr int RT = KB / 16;
  using MF = gemm::Mfma<T>;
  const int J = t / 64, lane = t % 64, lm = lane % 16, lg = lane / 16;
  typename MF::acc acc[RT];
#pragma unroll
  for (int I = 0; I < RT; ++I)
#pragma unroll
    for (int q = 0; q < 4; ++q) acc[I][q] = T(0);
#pragma unroll
  for (int K = 0; K < RT; ++K) {
    T b[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) b[j] = mb[16 * K + 4 * lg + j][16 * J + lm];
#pragma unroll
    for (int I = K; I < RT; ++I) {
      T a[4];
#pragma unroll
      for (int j = 0; j < 4; ++j) a[j] = inv[16 * I + lm][16 * K + 4 * lg + j];
#pragma unroll
      for (int j = 0; j < 4; ++j) acc[I] = MF::step(a[j], b[j], acc[I]);
    }
  }
  __syncthreads();  // every wave has read mb
#pragma unroll
  for (int I = 0; I < RT; ++I)
#pragma unroll
    for (int q = 0; q < 4; ++q) mb[16 * I + MF::out_row(lg, q)][16 * J + lm] = acc[I][q];
  __syncthreads();
}

// One processing step p of the solve; called by all 4 * KB threads of a
// workgroup.  With PF the last part of the step, the only part that waits on
// the previous block, runs on one wave (KB = 64 rows, one per lane).
// PF: load each panel block one iteration ahead (a kernel that keeps other
// work items in its registers passes false).
// BAND: M is a band of kb off-diagonals (see mat); the panel starts at the
// first block that reaches into the band.
template <class T, bool FWD, bool TRANS, bool UNIT, int KB = kB, bool PF = true, bool BAND = false>
__device__ __forceinline__ void step(int p, int64_t n, const T* __restrict__ A, int64_t lda, T* x, unsigned* pub,
                                     unsigned marker, unsigned* err, Smem<T, T, KB>& sm,
                                     T (*mb)[mb_stride<T, KB>()] = nullptr, T (*mb2)[mb_stride<T, KB>()] = nullptr,
                                     int64_t kb = 0) {
  constexpr int CW = KB / 4;  // columns per thread in the block mat-vecs
  const int nb = (int)((n + KB - 1) / KB);
  const int t = threadIdx.x;
  const int64_t bi = FWD ? p : nb - 1 - p;
  const int64_t r0 = bi * KB;
  // With PF, this block's right-hand side (to vs, local order; only this
  // workgroup writes these rows of x, at the end of the step) and
  // M(this block, block p-1) (to mb: mb[k][c] = M(row_of(k), column c of
  // block p-1)) are read before anything waits, together with the diagonal
  // block: every load is issued before the first is used, so the three
  // cost one memory round trip; the inversion's first barrier orders the
  // LDS writes before their use.
  constexpr int PT = per_thread<KB>();
  T rhs = T(0);
  if (PF && t < KB) {
    const int64_t g = row_of<FWD, KB>(r0, t);
    if (g < n) rhs = x[g];
  }
  // M(this block, block p - d) for d = 1, 2 (the blocks folded into G1, G2)
  auto load_prev = [&](int d, T (&m)[PT]) {
    const int64_t cb0 = (FWD ? (int64_t)(p - d) : (int64_t)(nb - 1 - p + d)) * KB;
#pragma unroll
    for (int e = 0; e < PT; ++e) {
      const int idx = t + 4 * KB * e;
      const int a = idx / KB, c2 = idx % KB;
      const int k = TRANS ? c2 : a, c = TRANS ? a : c2;  // lanes on consecutive addresses
      const int64_t gk = row_of<FWD, KB>(r0, k), gc = cb0 + c;
      m[e] = (gk < n && gc < n) ? mat<T, TRANS, BAND>(A, lda, gk, gc, kb) : T(0);
    }
  };
  auto stage_prev = [&](const T (&m)[PT], T (*dst)[mb_stride<T, KB>()]) {
#pragma unroll
    for (int e = 0; e < PT; ++e) {
      const int idx = t + 4 * KB * e;
      const int a = idx / KB, c2 = idx % KB;
      dst[TRANS ? c2 : a][TRANS ? a : c2] = m[e];
    }
  };
  T mv[PT], mv2[PT];
  if constexpr (PF) {
    if (p > 0) load_prev(1, mv);
    if (p > 1) load_prev(2, mv2);
  }
  T dv[PT];
  load_diag<T, T, FWD, TRANS, UNIT, KB, BAND>(n, A, lda, r0, t, dv, kb);
  if (PF && t < KB) sm.vs[t] = rhs;
  if constexpr (PF) {
    if (p > 0) stage_prev(mv, mb);
    if (p > 1) stage_prev(mv2, mb2);
  }
  invert_loaded<T, TRANS, KB>(sm.inv, t, dv);
  // wave 0's rows of inv(D) and G for the end of the step (PF)
  constexpr bool kInvInRegs = sizeof(T) == 4;
  T ir[kInvInRegs ? KB : 1];
  T gr[PF ? KB : 1];
  if constexpr (PF && kInvInRegs) {
    if (t < KB) {
#pragma unroll
      for (int c = 0; c < KB; ++c) ir[c] = sm.inv[t][c];
    }
  }
  // NoTrans: 4 threads per row, CW consecutive columns each (contiguous
  // loads); Trans: one column of KB rows per KB threads, CW columns per
  // group (consecutive lanes read consecutive addresses).
  const int r = TRANS ? t % KB : t / 4;
  const int cq = TRANS ? t / KB : t % 4;
  const int64_t gi = r0 + r;
  const int i4 = t / 4, q4 = t % 4;  // row (local order) and column quarter of the inv(D) products
  // PF: the two previous blocks' updates are folded into inv(D) ahead of
  // time, G1 = inv(D) M(this block, block p-1) and G2 = inv(D) M(this block,
  // block p-2), so the panel loop covers blocks 0 .. p-3 and
  // x_block = y0 - G2 x_{p-2} - G1 x_{p-1} (y0 = inv(D)(b - the other
  // blocks)): the panel's last block and y0 are a whole step ahead of the
  // chain, G2 x_{p-2} half a step, and once x_{p-1} appears the step is one
  // block mat-vec.  mb / mb2 hold M(row_of(k), column c of block p-1 / p-2);
  // the products run on the matrix cores (fold_g) and go over mb / mb2;
  // wave 0 keeps its rows of G1 in registers and reads G2 from LDS.
  const int pp = PF ? (p > 1 ? p - 2 : 0) : p;  // blocks the panel loop covers
  // a band's blocks p - q with (p - q - 1) KB >= kb hold no entry of it
  const int q0 = BAND ? (int)(p - ceil_div(kb, KB) > 0 ? p - ceil_div(kb, KB) : 0) : 0;
  if constexpr (PF) {
    if (p > 0) {
      fold_g<T, KB>(sm.inv, mb, t);
      if (t < KB) {
#pragma unroll
        for (int c = 0; c < KB; ++c) gr[c] = mb[t][c];
      }
    }
    if (p > 1) fold_g<T, KB>(sm.inv, mb2, t);
  }
  // This thread's CW matrix entries of block q, one bounds test for the
  // whole run when it is inside the matrix (so NoTrans loads merge into
  // 16-byte accesses).
  auto load_m = [&](int q, T (&m)[CW]) {
    const int64_t bj = FWD ? q : nb - 1 - q;
    const int64_t c0 = bj * KB + cq * CW;
    // (a band's run is whole when all CW entries lie inside the band)
    const bool whole = !BAND || (gi - c0 <= kb && c0 + CW - 1 - gi <= kb);
    if (whole && gi < n && c0 + CW <= n) {
#pragma unroll
      for (int e = 0; e < CW; ++e) m[e] = mat<T, TRANS>(A, lda, gi, c0 + e);
    } else {
#pragma unroll
      for (int e = 0; e < CW; ++e)
        m[e] = (gi < n && c0 + e < n) ? mat<T, TRANS, BAND>(A, lda, gi, c0 + e, kb) : T(0);
    }
  };
  // Block q's entries are loaded one iteration ahead, so an iteration costs
  // the wait for x_q and not also a matrix load's latency: the panel loop
  // then keeps ahead of the chain of published x blocks.
  T acc = T(0);
  if constexpr (!PF) {
    for (int q = q0; q < p; ++q) {
      T m[CW];
      load_m(q, m);
      T* xs = sm.xs[q & 1];
      fetch_block<T, KB>(n, FWD ? q : nb - 1 - q, pub, marker, err, xs);
      __syncthreads();
#pragma unroll
      for (int e = 0; e < CW; ++e) acc += m[e] * xs[cq * CW + e];
    }
  }
  T m[2][CW];
  if (PF && pp > q0) load_m(q0, m[0]);
  for (int q = q0; PF && q < pp; q += 2) {
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int qq = q + h;
      if (qq < pp) {
        if (qq + 1 < pp) load_m(qq + 1, m[h ^ 1]);
        const int64_t bj = FWD ? qq : nb - 1 - qq;
        T* xs = sm.xs[h];  // double-buffered: one barrier per block
        fetch_block<T, KB>(n, bj, pub, marker, err, xs);
        __syncthreads();
#pragma unroll
        for (int e = 0; e < CW; ++e) acc += m[h][e] * xs[cq * CW + e];
      }
    }
  }
  if constexpr (!PF) {
    // v = rhs - panel sums in local order, then x_local = inv(D) v with 4
    // threads per row
    sm.red[cq][r] = acc;
    __syncthreads();
    if (t < KB) {
      const int64_t g = row_of<FWD, KB>(r0, t);
      const int nr = (int)(g - r0);
      T v = T(0);
      if (g < n) {
        T s = T(0);
#pragma unroll
        for (int w = 0; w < 4; ++w) s += sm.red[w][nr];
        v = x[g] - s;
      }
      sm.vs[t] = v;
    }
    __syncthreads();
    T s = T(0);
#pragma unroll
    for (int e = 0; e < CW; ++e) s += sm.inv[i4][q4 * CW + e] * sm.vs[q4 * CW + e];
    s += __shfl_xor(s, 1, 64);
    s += __shfl_xor(s, 2, 64);
    const int64_t g = row_of<FWD, KB>(r0, i4);
    if (q4 == 0 && g < n) publish_value<T>(g, s, x, pub, marker);
  } else {
    static_assert(KB == 64, "the chain part of a step runs on one wave, one row per lane");
    // v = rhs - panel sums (vs holds the rhs in local order): NoTrans keeps
    // a row's four partial sums in adjacent lanes, so two shuffles reduce
    // them; Trans spreads them over the waves, so they meet in LDS.
    if constexpr (!TRANS) {
      T ps = acc + __shfl_xor(acc, 1, 64);
      ps += __shfl_xor(ps, 2, 64);
      if (cq == 0 && gi < n) {
        const int k = FWD ? r : KB - 1 - r;
        sm.vs[k] -= ps;
      }
    } else {
      sm.red[cq][r] = acc;
      __syncthreads();
      if (t < KB) {
        const int64_t g = row_of<FWD, KB>(r0, t);
        if (g < n) {
          const int nr = (int)(g - r0);
          sm.vs[t] -= (sm.red[0][nr] + sm.red[1][nr]) + (sm.red[2][nr] + sm.red[3][nr]);
        }
      }
    }
    __syncthreads();
    // The rest on wave 0, lane i = local row i: y0 = inv(D) v (FP32: its
    // rows of inv(D) in registers since the inversion), then
    // y0 - G2 x_{p-2} (G2 and x_{p-2} from LDS), then, once x_{p-1} is
    // published, x = y0 - G2 x_{p-2} - G1 x_{p-1} with the rows of G1 in
    // registers.  x passes through LDS with no workgroup barrier on the
    // chain (LDS operations of one wave complete in order).  (Measured: a
    // second wave forming y0, or y0 - G2 x_{p-2}, and handing it over
    // through LDS and a flag gave no shorter step: the step is the
    // publication's round trip plus the G1 product.)
    if (t < KB) {
      constexpr int VN = 16 / (int)sizeof(T);
      const VecN<T, VN>* vv = reinterpret_cast<const VecN<T, VN>*>(sm.vs);
      // every LDS read issued before the first product (one LDS latency,
      // not one per pair of reads)
      VecN<T, VN> vr[KB / VN];
#pragma unroll
      for (int c = 0; c < KB / VN; ++c) vr[c] = vv[c];
      T y[4] = {T(0), T(0), T(0), T(0)};
#pragma unroll
      for (int c = 0; c < KB / VN; ++c) {
        const VecN<T, VN> v4 = vr[c];
#pragma unroll
        for (int e = 0; e < VN; ++e) {
          const int k = c * VN + e;
          const T iv = kInvInRegs ? ir[k] : sm.inv[t][k];
          y[k % 4] += iv * v4.v[e];
        }
      }
      T xv = (y[0] + y[1]) + (y[2] + y[3]);
      if (p > 1) {
        const int64_t g2 = (FWD ? (int64_t)(p - 2) : (int64_t)(nb - 1 - p + 2)) * KB + t;
        T* xs = sm.xs[1];
        xs[t] = g2 < n ? tagged::read<T>(pub, g2, marker, err) : T(0);
        __builtin_amdgcn_wave_barrier();
        const VecN<T, VN>* x4 = reinterpret_cast<const VecN<T, VN>*>(xs);
        const VecN<T, VN>* g4 = reinterpret_cast<const VecN<T, VN>*>(mb2[t]);
        T u[4] = {T(0), T(0), T(0), T(0)};
#pragma unroll
        for (int c = 0; c < KB / VN; ++c) {
          const VecN<T, VN> xx = x4[c], gg = g4[c];
#pragma unroll
          for (int e = 0; e < VN; ++e) u[(c * VN + e) % 4] += gg.v[e] * xx.v[e];
        }
        xv -= (u[0] + u[1]) + (u[2] + u[3]);
      }
      if (p > 0) {
        const int64_t bj = FWD ? (int64_t)(p - 1) : (int64_t)(nb - p);
        const int64_t g1 = bj * KB + t;
        T* xs = sm.xs[0];
        xs[t] = g1 < n ? tagged::read<T>(pub, g1, marker, err) : T(0);
        __builtin_amdgcn_wave_barrier();
        const VecN<T, VN>* x4 = reinterpret_cast<const VecN<T, VN>*>(xs);
        // x read in chunks whose loads are all issued before the products
        // (FP64: four, so the rows of G1 and the chunk fit the VGPRs)
        constexpr int CH = sizeof(T) == 4 ? KB / VN : KB / VN / 4;
        T u[4] = {T(0), T(0), T(0), T(0)};
#pragma unroll
        for (int c0 = 0; c0 < KB / VN; c0 += CH) {
          VecN<T, VN> xr[CH];
#pragma unroll
          for (int c = 0; c < CH; ++c) xr[c] = x4[c0 + c];
#pragma unroll
          for (int c = 0; c < CH; ++c) {
            const VecN<T, VN> xx = xr[c];
#pragma unroll
            for (int e = 0; e < VN; ++e) u[((c0 + c) * VN + e) % 4] += gr[(c0 + c) * VN + e] * xx.v[e];
          }
        }
        xv -= (u[0] + u[1]) + (u[2] + u[3]);
      }
      const int64_t g = row_of<FWD, KB>(r0, t);
      if (g < n) publish_value<T>(g, xv, x, pub, marker);
    }
  }
  __syncthreads();  // the shared buffers are reused by the caller's next step
}

}  // namespace trsv
}  // namespace exoblas
