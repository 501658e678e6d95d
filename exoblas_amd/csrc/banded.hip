// Banded level-2 kernels for gfx950: gbmv, sbmv, tbmv (one band-gather
// kernel) and tbsv (a one-wave sequential solve).
//
// Band storage (row-major, the reference's layout): a band matrix row i keeps
// its in-band entries in A[i, c0 + d] for offsets d = j - i of the band,
// where c0 is the column of the diagonal (kl for gbmv, 0 for upper, k for
// lower storage).  A symmetric or transposed product also reads entries
// (i, j) from the other triangle, stored as (j, i): A[j, c0 - d].  Every
// product row is therefore a sum over at most two offset ranges ("segments"),
// each read row-wise (A[i, c0 + d]: contiguous per row) or column-wise
// (A[j, c0 - d]: one element per stored row).
//
// band_lds_kernel (bands whose rows fit in LDS): a workgroup owns 256, 128 or
// 64 consecutive output rows (the most whose staging fits in 64 KiB), copies the band rows they touch (their own and, for
// column-wise segments, the k rows either side) and the x window into LDS with
// coalesced loads, then each thread sums one row from LDS in offset order.
// band_mv_kernel (wider bands): a group of G lanes per output row (G = 1, 4,
// 16 or 64 by band width) sums its row's products lane-strided and reduces
// with xor shuffles in a fixed order.  Both are deterministic.
// tbsv_window_kernel (k <= 32): one wave solves blocks of 64 rows from LDS
// while it prefetches the next block, with the last k solutions in registers.
// tbsv_kernel (wider bands): one wave walks the rows in dependency order;
// lane l handles offsets d = 1 + l, 1 + l + 64, ...; the last k solutions live
// in an LDS ring, so each step is one wave reduction and one division.
#include "device.hpp"
#include "ops.hpp"

namespace exoblas {
namespace {

constexpr int kNT = 256;

// Offsets d in [lo, hi] of row i; col: element (i, i + d) at A[i, c0 + d]
// (row-wise, col == false) or A[i + d, c0 - d] (column-wise, col == true).
struct Seg {
  int64_t lo, hi, c0;
  int col;
};

enum Out : int { kGbmv = 0, kAccumulate = 1, kOverwrite = 2 };

// Row i of y: OUT kGbmv y = beta*y + alpha*s; kAccumulate y += alpha*s;
// kOverwrite y = s (tbmv, reading x from a copy).  unit: the diagonal entry
// (d == 0) is 1 and A's diagonal is not read.
template <class T, int G>
__global__ __launch_bounds__(kNT) void band_mv_kernel(int64_t rows, int64_t ncols, Seg s0, Seg s1, int nseg,
                                                      const T* __restrict__ A, int64_t lda,
                                                      const T* __restrict__ x, int64_t incx, T* y, int64_t incy,
                                                      T alpha, T beta, int out, int unit) {
  constexpr int RPB = kNT / G;  // rows per workgroup
  const int g = threadIdx.x % G;
  const int64_t i = (int64_t)blockIdx.x * RPB + threadIdx.x / G;
  T s = T(0);
  if (i < rows) {
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      if (q >= nseg) break;
      const Seg sg = q == 0 ? s0 : s1;
      // j = i + d within [0, ncols)
      const int64_t dlo = sg.lo > -i ? sg.lo : -i;
      const int64_t dhi = sg.hi < ncols - 1 - i ? sg.hi : ncols - 1 - i;
      for (int64_t d = dlo + g; d <= dhi; d += G) {
        const int64_t j = i + d;
        T a;
        if (unit && d == 0)
          a = T(1);
        else
          a = sg.col ? A[j * lda + (sg.c0 - d)] : A[i * lda + (sg.c0 + d)];
        s += a * x[j * incx];
      }
    }
  }
#pragma unroll
  for (int off = G / 2; off > 0; off >>= 1) s += __shfl_xor(s, off, kWave);
  if (g == 0 && i < rows) {
    T* yi = y + i * incy;
    if (out == kGbmv)
      *yi = beta * *yi + alpha * s;
    else if (out == kAccumulate)
      *yi += alpha * s;
    else
      *yi = s;
  }
}

// Rows [i0, i0 + NT) of the band product with the band rows [rs0, rs1) and
// x[xs0 .. xs1) staged in LDS; band row r occupies LDS [(r - rs0) * ws, +wc).
// When lda == wc the staged rows are one contiguous range copied as is (ws ==
// wc, no index arithmetic); otherwise ws = wc rounded up to odd, which keeps
// the row-per-thread reads free of bank conflicts.
template <class T, int NT, int U>
__global__ __launch_bounds__(NT) void band_lds_kernel(int64_t rows, int64_t ncols, int64_t srows, Seg s0, Seg s1,
                                                       int nseg, int64_t wc, int64_t ws, int64_t reach_lo,
                                                       int64_t reach_hi, int64_t xlo, int64_t xhi,
                                                       const T* __restrict__ A, int64_t lda,
                                                       const T* __restrict__ x, int64_t incx, T* y, int64_t incy,
                                                       T alpha, T beta, int out, int unit) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int64_t i0 = (int64_t)blockIdx.x * NT;
  const int64_t rs0 = i0 + reach_lo > 0 ? i0 + reach_lo : 0;
  const int64_t rs1 = i0 + NT + reach_hi < srows ? i0 + NT + reach_hi : srows;
  const int64_t nr = rs1 > rs0 ? rs1 - rs0 : 0;
  T* band = reinterpret_cast<T*>(smem);
  T* xs = band + (NT + reach_hi - reach_lo) * ws;
  const int64_t xs0 = i0 + xlo > 0 ? i0 + xlo : 0;
  const int64_t xs1 = i0 + NT + xhi < ncols ? i0 + NT + xhi : ncols;
  // global -> LDS, 8 loads in flight per thread (LDS offsets fit 32 bits)
  const uint32_t E = (uint32_t)(nr * wc), wcu = (uint32_t)wc, wsu = (uint32_t)ws;
  if (lda == wc) {
    // one contiguous range: its 16-byte-aligned interior as streaming 16-byte
    // loads (U in flight per thread), the ragged ends element by element; no
    // load leaves [src, src + E)
    using V = typename Vec16<T>::type;
    constexpr int VN = Vec16<T>::n;
    const T* src = A + rs0 * lda;
    const uint32_t mis = (uint32_t)((reinterpret_cast<uintptr_t>(src) & 15) / sizeof(T));
    const uint32_t h = std::min<uint32_t>(mis ? (uint32_t)VN - mis : 0u, E);
    const uint32_t nv = (E - h) / VN, t0 = h + nv * VN;
    for (uint32_t e = threadIdx.x; e < h; e += NT) band[e] = src[e];
    for (uint32_t e = t0 + threadIdx.x; e < E; e += NT) band[e] = src[e];
    const T* body = src + h;
    for (uint32_t v0 = threadIdx.x; v0 < nv; v0 += U * NT) {
      V r[U];
#pragma unroll
      for (int u = 0; u < U; ++u)
        if (v0 + u * NT < nv) r[u] = ld_stream<T>(body + (size_t)(v0 + u * NT) * VN);
#pragma unroll
      for (int u = 0; u < U; ++u)
        if (v0 + u * NT < nv)
#pragma unroll
          for (int c = 0; c < VN; ++c) band[h + (v0 + u * NT) * VN + c] = vget<T>(r[u], c);
    }
  } else {
    for (uint32_t e0 = threadIdx.x; e0 < E; e0 += 8 * NT) {
      T v[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const uint32_t e = e0 + u * NT, r = e / wcu;
        if (e < E) v[u] = A[(rs0 + r) * lda + (e - r * wcu)];
      }
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const uint32_t e = e0 + u * NT, r = e / wcu;
        if (e < E) band[r * wsu + (e - r * wcu)] = v[u];
      }
    }
  }
  for (int64_t j = xs0 + threadIdx.x; j < xs1; j += NT) xs[j - xs0] = x[j * incx];
  __syncthreads();
  const int64_t i = i0 + threadIdx.x;
  if (i >= rows) return;
  const int ii = (int)(i - rs0), jx = (int)(i - xs0);  // this row's LDS row and x offset
  T s = T(0);
#pragma unroll
  for (int q = 0; q < 2; ++q) {
    if (q >= nseg) break;
    const Seg sg = q == 0 ? s0 : s1;
    const int dlo = (int)(sg.lo > -i ? sg.lo : -i);
    const int dhi = (int)(sg.hi < ncols - 1 - i ? sg.hi : ncols - 1 - i);
    const int c0 = (int)sg.c0, w = (int)ws;
    for (int d = dlo; d <= dhi; ++d) {
      T a;
      if (unit && d == 0)
        a = T(1);
      else
        a = sg.col ? band[(ii + d) * w + (c0 - d)] : band[ii * w + (c0 + d)];
      s += a * xs[jx + d];
    }
  }
  T* yi = y + i * incy;
  if (out == kGbmv)
    *yi = beta * *yi + alpha * s;
  else if (out == kAccumulate)
    *yi += alpha * s;
  else
    *yi = s;
}

// srows: rows of the band storage; wc: its columns in use.
template <class T>
void launch_band_mv(Ctx& c, int64_t rows, int64_t ncols, int64_t srows, int64_t wc, Seg s0, Seg s1, int nseg,
                    const T* A, int64_t lda, const T* x, int64_t incx, T* y, int64_t incy, T alpha, T beta, int out,
                    bool unit) {
  if (rows <= 0) return;
  // rows the column-wise segments reach beyond the block, and the x window
  int64_t reach_lo = 0, reach_hi = 0, xlo = s0.lo, xhi = s0.hi;
  for (int q = 0; q < nseg; ++q) {
    const Seg& sg = q == 0 ? s0 : s1;
    if (sg.col) {
      reach_lo = std::min(reach_lo, sg.lo);
      reach_hi = std::max(reach_hi, sg.hi);
    }
    xlo = std::min(xlo, sg.lo);
    xhi = std::max(xhi, sg.hi);
  }
  const int64_t ws = lda == wc ? wc : (wc | 1);
  // the largest row block (256, 128 or 64 rows) whose staging fits in 64 KiB
  auto lds_of = [&](int64_t nt) { return sizeof(T) * (size_t)((nt + reach_hi - reach_lo) * ws + (nt + xhi - xlo)); };
  // 16 staging loads in flight per thread when a thread has more than 8
  // vectors to copy, else 8 (measured on MI355X at n = 2^22, k = 16: sgbmv
  // 4.39 -> 4.59 TB/s, dsbmv 4.54 -> 4.98 with 16; ssbmv 4.56 -> 4.08 with 16)
#define EXO_BAND_LDS(NTT)                                                                                        \
  if (lds_of(NTT) <= (64 << 10)) {                                                                               \
    const bool deep = (NTT + reach_hi - reach_lo) * wc * (int64_t)sizeof(T) > (int64_t)NTT * 8 * 16;           \
    if (deep)                                                                                                    \
      hipLaunchKernelGGL((band_lds_kernel<T, NTT, 16>), dim3((unsigned)ceil_div(rows, NTT)), dim3(NTT),          \
                         lds_of(NTT), c.stream, rows, ncols, srows, s0, s1, nseg, wc, ws, reach_lo, reach_hi, xlo, \
                         xhi, A, lda, x, incx, y, incy, alpha, beta, out, (int)unit);                             \
    else                                                                                                         \
      hipLaunchKernelGGL((band_lds_kernel<T, NTT, 8>), dim3((unsigned)ceil_div(rows, NTT)), dim3(NTT),           \
                         lds_of(NTT), c.stream, rows, ncols, srows, s0, s1, nseg, wc, ws, reach_lo, reach_hi, xlo, \
                         xhi, A, lda, x, incx, y, incy, alpha, beta, out, (int)unit);                             \
    EXO_LAUNCH_CHECK();                                                                                          \
    return;                                                                                                      \
  }
  EXO_BAND_LDS(256)
  EXO_BAND_LDS(128)
  EXO_BAND_LDS(64)
#undef EXO_BAND_LDS
  int64_t w = s0.hi - s0.lo + 1 + (nseg > 1 ? s1.hi - s1.lo + 1 : 0);
#define EXO_BAND(GG)                                                                                       \
  hipLaunchKernelGGL((band_mv_kernel<T, GG>), dim3((unsigned)ceil_div(rows, kNT / GG)), dim3(kNT), 0, c.stream, \
                     rows, ncols, s0, s1, nseg, A, lda, x, incx, y, incy, alpha, beta, out, (int)unit)
  if (w >= 128)
    EXO_BAND(64);
  else if (w >= 24)
    EXO_BAND(16);
  else if (w >= 6)
    EXO_BAND(4);
  else
    EXO_BAND(1);
#undef EXO_BAND
  EXO_LAUNCH_CHECK();
}

// y[i] = x[i * incx] for a scratch copy (tbmv reads the old x while writing).
template <class T>
__global__ __launch_bounds__(kNT) void gather_kernel(int64_t n, const T* x, int64_t incx, T* out) {
  for (int64_t i = (int64_t)blockIdx.x * kNT + threadIdx.x; i < n; i += (int64_t)gridDim.x * kNT)
    out[i] = x[i * incx];
}

// One wave solves op(A) x = b in place for a band of k off-diagonals.  fwd:
// rows 0, 1, ... (x_i needs x_{i-d}); else rows n-1, n-2, ... (x_{i+d}).
// Entry (i, j = i -/+ d) is A[i, c0 + (j - i)] (row-wise) or A[j, c0 - (j - i)]
// (column-wise); the diagonal is A[i, c0] (1 when unit).  ring holds
// R >= k + 1 solved values, indexed by row mod R.
template <class T>
__global__ __launch_bounds__(kWave) void tbsv_kernel(int64_t n, int64_t k, int fwd, int col, int64_t c0, int unit,
                                                     const T* __restrict__ A, int64_t lda, T* x, int64_t incx,
                                                     int64_t R) {
  extern __shared__ __align__(16) unsigned char smem[];
  T* ring = reinterpret_cast<T*>(smem);
  const int lane = threadIdx.x;
  const int64_t sgn = fwd ? -1 : 1;  // j = i + sgn * d
  for (int64_t s = 0; s < n; ++s) {
    const int64_t i = fwd ? s : n - 1 - s;
    const T b = x[i * incx];
    const T piv = unit ? T(1) : A[i * lda + c0];
    T part = T(0);
    for (int64_t d = 1 + lane; d <= k; d += kWave) {
      const int64_t j = i + sgn * d;
      if (j < 0 || j >= n) break;  // d only grows: the rest are out of range too
      const int64_t off = j - i;
      const T a = col ? A[j * lda + (c0 - off)] : A[i * lda + (c0 + off)];
      part += a * ring[j & (R - 1)];
    }
    const T dot = wave_sum(part);
    const T xi = (b - dot) / piv;
    if (lane == 0) {
      ring[i & (R - 1)] = xi;
      x[i * incx] = xi;
    }
    __syncthreads();  // the ring entry is visible to the next row's reads
  }
}

// Narrow bands (k <= KMAX <= 32): the rows are solved in blocks of 64 in
// dependency order.  While block b is solved, every lane loads row l of
// block b + 1 (its pivot, its KMAX offset entries, 0 past the band or the
// matrix, and its right-hand side) into registers; after block b they go to
// LDS.  The solve itself reads the block from LDS (uniform addresses) and
// keeps the last KMAX solutions in a register window, so a step is KMAX
// multiply-adds in four chains and one multiply by the pivot's reciprocal
// (taken when the row is fetched), with no memory round trip.
template <class T, int KMAX>
__global__ __launch_bounds__(kWave) void tbsv_window_kernel(int64_t n, int64_t k, int fwd, int col, int64_t c0,
                                                            int unit, const T* __restrict__ A, int64_t lda, T* x,
                                                            int64_t incx) {
  __shared__ T arow[kWave][KMAX + 1];  // [r][0] 1 / pivot, [r][d] entry of offset d
  __shared__ T bv[kWave];
  const int lane = threadIdx.x;
  const int64_t nb = ceil_div(n, kWave);
  T ra[KMAX + 1], rb = T(0);
  auto row_of = [&](int64_t s) { return fwd ? s : n - 1 - s; };
  // fetch only issues loads (every one unconditional, from the row's diagonal
  // where the entry lies outside the band or the matrix), so none waits for
  // another; stash, after the block's solve, masks them, takes the pivot's
  // reciprocal and writes LDS.
  auto fetch = [&](int64_t blk) {
    const int64_t s = blk * kWave + lane;
    const int64_t i = row_of(s < n ? s : n - 1);
    const int64_t dg = i * lda + c0;
    ra[0] = A[dg];
#pragma unroll
    for (int d = 1; d <= KMAX; ++d) {
      const int64_t j = fwd ? i - d : i + d;
      const bool in = d <= k && j >= 0 && j < n;
      ra[d] = A[in ? (col ? j * lda + (c0 - (j - i)) : dg + (j - i)) : dg];
    }
    rb = x[i * incx];
  };
  auto stash = [&](int64_t blk) {
    const int64_t s = blk * kWave + lane;
    const int64_t i = row_of(s < n ? s : n - 1);
    arow[lane][0] = unit ? T(1) : T(1) / ra[0];
#pragma unroll
    for (int d = 1; d <= KMAX; ++d) {
      const int64_t j = fwd ? i - d : i + d;
      arow[lane][d] = (d <= k && j >= 0 && j < n) ? ra[d] : T(0);
    }
    bv[lane] = rb;
  };
  T xw[KMAX];  // xw[d - 1] = solution of the row d steps back
#pragma unroll
  for (int d = 0; d < KMAX; ++d) xw[d] = T(0);
  fetch(0);
  stash(0);
  __syncthreads();
  for (int64_t blk = 0; blk < nb; ++blk) {
    if (blk + 1 < nb) fetch(blk + 1);
    const int cnt = (int)(n - blk * kWave < kWave ? n - blk * kWave : kWave);
    // FP32: branch-free and unrolled, so the LDS reads of the next rows issue
    // ahead of this row's chain; lane r keeps row r's solution and the block's
    // solutions are stored together afterwards (11.1 -> 10.2 ms at n = 2^16,
    // k = 16).  FP64 stores each row as it is solved (13.8 ms against 17.0
    // with the block store, whose registers it runs short of).
    constexpr bool kBlockStore = sizeof(T) == 4;
    T mine = T(0);
#pragma unroll 4
    for (int r = 0; r < cnt; ++r) {
      T acc[4] = {T(0), T(0), T(0), T(0)};
#pragma unroll
      for (int d = 1; d <= KMAX; ++d) acc[(d - 1) & 3] += arow[r][d] * xw[d - 1];
      const T xi = (bv[r] - ((acc[0] + acc[1]) + (acc[2] + acc[3]))) * arow[r][0];
#pragma unroll
      for (int d = KMAX - 1; d > 0; --d) xw[d] = xw[d - 1];
      xw[0] = xi;
      if constexpr (kBlockStore)
        mine = lane == r ? xi : mine;
      else if (lane == 0)
        x[row_of(blk * kWave + r) * incx] = xi;
    }
    if (kBlockStore && lane < cnt) x[row_of(blk * kWave + lane) * incx] = mine;
    __syncthreads();
    if (blk + 1 < nb) stash(blk + 1);
    __syncthreads();
  }
}

template <class T>
void launch_tbsv(Ctx& c, int64_t n, int64_t k, bool fwd, bool col, int64_t c0, bool unit, const T* A, int64_t lda,
                 T* x, int64_t incx) {
  if (n <= 0) return;
#define EXO_TBSV_W(KM)                                                                                          \
  if (k <= KM) {                                                                                                \
    hipLaunchKernelGGL((tbsv_window_kernel<T, KM>), dim3(1), dim3(kWave), 0, c.stream, n, k, (int)fwd, (int)col, \
                       c0, (int)unit, A, lda, x, incx);                                                         \
    EXO_LAUNCH_CHECK();                                                                                         \
    return;                                                                                                     \
  }
  EXO_TBSV_W(4)
  EXO_TBSV_W(8)
  EXO_TBSV_W(16)
  EXO_TBSV_W(32)
#undef EXO_TBSV_W
  int64_t R = 1;
  while (R < k + 1) R <<= 1;
  const size_t bytes = sizeof(T) * (size_t)R;
  constexpr size_t kMaxLds = 128 << 10;
  if (bytes > kMaxLds)
    throw Error(kInvalidSize, "tbsv: band of " + std::to_string(k) + " off-diagonals exceeds the solver's " +
                                  std::to_string(kMaxLds / sizeof(T) - 1) + " (use trsv)");
  static std::atomic<uint64_t> lds_set{0};
  if (bytes > (64 << 10))
    ensure_lds_limit(reinterpret_cast<const void*>(&tbsv_kernel<T>), kMaxLds, c.device, lds_set);
  hipLaunchKernelGGL((tbsv_kernel<T>), dim3(1), dim3(kWave), bytes, c.stream, n, k, (int)fwd, (int)col, c0,
                     (int)unit, A, lda, x, incx, R);
  EXO_LAUNCH_CHECK();
}

}  // namespace

namespace ops {

// src/level2/gbmv.py:13-52.  y = beta*y + alpha*A x, A m x n with kl sub- and
// ku super-diagonals, a[i, kl + (j - i)] = A(i, j).  Rows with no in-band
// column (i >= n + kl) get y = beta*y.
template <class T>
void gbmv_row_major_NonTrans(Ctx& c, int64_t m, int64_t n, int64_t kl, int64_t ku, T alpha, T beta, const T* a,
                             int64_t lda, const T* x, int64_t incx, T* y, int64_t incy) {
  const Seg s{-kl, ku, kl, 0};
  launch_band_mv<T>(c, m, n, m, kl + ku + 1, s, s, 1, a, lda, x, incx, y, incy, alpha, beta, kGbmv, false);
}

// src/level2/sbmv.py:12-16  y = beta*y (the wrapper calls it when beta != 1)
template <class T>
void sbmv_scal_y(Ctx& c, int64_t n, T beta, T* y, int64_t incy) {
  scal<T>(c, n, beta, y, incy);
}

// src/level2/sbmv.py:19-33.  y += alpha*S x, S symmetric with k
// super-diagonals stored: S(i, j>=i) = A[i, j - i].
template <class T>
void sbmv_row_major_Upper(Ctx& c, int64_t n, int64_t k, T alpha, const T* A, int64_t lda, const T* x, int64_t incx,
                          T* y, int64_t incy) {
  launch_band_mv<T>(c, n, n, n, k + 1, Seg{0, k, 0, 0}, Seg{-k, -1, 0, 1}, k > 0 ? 2 : 1, A, lda, x, incx, y,
                    incy, alpha, T(0), kAccumulate, false);
}

// src/level2/sbmv.py:36-50.  S(i, j<=i) = A[i, k + (j - i)].
template <class T>
void sbmv_row_major_Lower(Ctx& c, int64_t n, int64_t k, T alpha, const T* A, int64_t lda, const T* x, int64_t incx,
                          T* y, int64_t incy) {
  launch_band_mv<T>(c, n, n, n, k + 1, Seg{-k, 0, k, 0}, Seg{1, k, k, 1}, k > 0 ? 2 : 1, A, lda, x, incx, y,
                    incy, alpha, T(0), kAccumulate, false);
}

// x = op(T) x for a triangular band (tbmv): gather the old x, then overwrite.
template <class T>
void tbmv(Ctx& c, int64_t n, int64_t k, T* x, int64_t incx, const T* A, int64_t lda, int64_t Diag, Seg s) {
  if (n <= 0) return;
  T* xs = static_cast<T*>(c.scratch(sizeof(T) * (size_t)n, 1));
  hipLaunchKernelGGL((gather_kernel<T>), dim3((unsigned)std::min<int64_t>(ceil_div(n, kNT), 1024)), dim3(kNT), 0,
                     c.stream, n, x, incx, xs);
  EXO_LAUNCH_CHECK();
  launch_band_mv<T>(c, n, n, n, k + 1, s, s, 1, A, lda, xs, 1, x, incx, T(1), T(0), kOverwrite, Diag != 0);
}

// src/level2/tbmv.py:12-43  x = U x, U(i, j>=i) = A[i, j - i]
template <class T>
void tbmv_row_major_Upper_NonTrans(Ctx& c, int64_t n, int64_t k, T* x, int64_t incx, const T* A, int64_t lda,
                                   int64_t Diag) {
  tbmv<T>(c, n, k, x, incx, A, lda, Diag, Seg{0, k, 0, 0});
}
// src/level2/tbmv.py:46-77  x = L x, L(i, j<=i) = A[i, k + (j - i)]
template <class T>
void tbmv_row_major_Lower_NonTrans(Ctx& c, int64_t n, int64_t k, T* x, int64_t incx, const T* A, int64_t lda,
                                   int64_t Diag) {
  tbmv<T>(c, n, k, x, incx, A, lda, Diag, Seg{-k, 0, k, 0});
}
// src/level2/tbmv.py:80-108  x = U^T x: row i sums U(j, i) x_j, j in [i - k, i]
template <class T>
void tbmv_row_major_Upper_Trans(Ctx& c, int64_t n, int64_t k, T* x, int64_t incx, const T* A, int64_t lda,
                                int64_t Diag) {
  tbmv<T>(c, n, k, x, incx, A, lda, Diag, Seg{-k, 0, 0, 1});
}
// src/level2/tbmv.py:111-141  x = L^T x: row i sums L(j, i) x_j, j in [i, i + k]
template <class T>
void tbmv_row_major_Lower_Trans(Ctx& c, int64_t n, int64_t k, T* x, int64_t incx, const T* A, int64_t lda,
                                int64_t Diag) {
  tbmv<T>(c, n, k, x, incx, A, lda, Diag, Seg{0, k, k, 1});
}

// src/level2/tbsv.py:12-45  U x = b, rows n-1 .. 0
template <class T>
void tbsv_row_major_Upper_NonTrans(Ctx& c, int64_t n, int64_t k, T* x, int64_t incx, const T* A, int64_t lda,
                                   int64_t Diag) {
  launch_tbsv<T>(c, n, k, false, false, 0, Diag != 0, A, lda, x, incx);
}
// src/level2/tbsv.py:48-81  L x = b, rows 0 .. n-1
template <class T>
void tbsv_row_major_Lower_NonTrans(Ctx& c, int64_t n, int64_t k, T* x, int64_t incx, const T* A, int64_t lda,
                                   int64_t Diag) {
  launch_tbsv<T>(c, n, k, true, false, k, Diag != 0, A, lda, x, incx);
}
// src/level2/tbsv.py:84-105  U^T x = b, rows 0 .. n-1
template <class T>
void tbsv_row_major_Upper_Trans(Ctx& c, int64_t n, int64_t k, T* x, int64_t incx, const T* A, int64_t lda,
                                int64_t Diag) {
  launch_tbsv<T>(c, n, k, true, true, 0, Diag != 0, A, lda, x, incx);
}
// src/level2/tbsv.py:108-129  L^T x = b, rows n-1 .. 0
template <class T>
void tbsv_row_major_Lower_Trans(Ctx& c, int64_t n, int64_t k, T* x, int64_t incx, const T* A, int64_t lda,
                                int64_t Diag) {
  launch_tbsv<T>(c, n, k, false, true, k, Diag != 0, A, lda, x, incx);
}

#define EXO_BANDED_INST(T)                                                                                       \
  template void gbmv_row_major_NonTrans<T>(Ctx&, int64_t, int64_t, int64_t, int64_t, T, T, const T*, int64_t,  \
                                           const T*, int64_t, T*, int64_t);                                    \
  template void sbmv_scal_y<T>(Ctx&, int64_t, T, T*, int64_t);                                                   \
  template void sbmv_row_major_Upper<T>(Ctx&, int64_t, int64_t, T, const T*, int64_t, const T*, int64_t, T*,     \
                                        int64_t);                                                                \
  template void sbmv_row_major_Lower<T>(Ctx&, int64_t, int64_t, T, const T*, int64_t, const T*, int64_t, T*,     \
                                        int64_t);                                                                \
  template void tbmv_row_major_Upper_NonTrans<T>(Ctx&, int64_t, int64_t, T*, int64_t, const T*, int64_t, int64_t); \
  template void tbmv_row_major_Lower_NonTrans<T>(Ctx&, int64_t, int64_t, T*, int64_t, const T*, int64_t, int64_t); \
  template void tbmv_row_major_Upper_Trans<T>(Ctx&, int64_t, int64_t, T*, int64_t, const T*, int64_t, int64_t);    \
  template void tbmv_row_major_Lower_Trans<T>(Ctx&, int64_t, int64_t, T*, int64_t, const T*, int64_t, int64_t);    \
  template void tbsv_row_major_Upper_NonTrans<T>(Ctx&, int64_t, int64_t, T*, int64_t, const T*, int64_t, int64_t); \
  template void tbsv_row_major_Lower_NonTrans<T>(Ctx&, int64_t, int64_t, T*, int64_t, const T*, int64_t, int64_t); \
  template void tbsv_row_major_Upper_Trans<T>(Ctx&, int64_t, int64_t, T*, int64_t, const T*, int64_t, int64_t);    \
  template void tbsv_row_major_Lower_Trans<T>(Ctx&, int64_t, int64_t, T*, int64_t, const T*, int64_t, int64_t);
EXO_BANDED_INST(float)
EXO_BANDED_INST(double)

}  // namespace ops
}  // namespace exoblas
