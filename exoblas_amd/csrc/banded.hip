// Banded level-2 kernels for gfx950: gbmv, sbmv, tbmv (one band-gather
// kernel), tbsv (a blocked solve for narrow bands, trsv's block solver for
// wide ones).
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
// tbsv, k <= 32: a blocked solve (per-block responses to the incoming
// state, a k x k chain over the blocks, a per-block finish), or one wave over
// all rows for short systems.
// tbsv, k > 32 (and short systems with k <= 32, where it is faster): the
// block solver of trsv (64-row blocks chained between workgroups) over the
// band seen as a dense triangle, zero outside the band.
#include <type_traits>

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

}  // namespace

// The narrow-band solve's kernels have external linkage: inside the unnamed
// namespace the compiler allocates tbsv_rows_kernel<..., kResponses> twice
// the registers (256 against 118 for FP32, k <= 16) and spills.
namespace tbsv_blocked {

// Narrow bands (k <= KMAX <= 32).  Rows are taken in dependency order s
// (row i = s forward, n - 1 - s backward): x_s = (b_s - sum_{d=1..k} a_s(d)
// x_{s-d}) / pivot_s.  Blocked solve: the rows are cut into blocks of blen
// (a multiple of 64).  A block's solutions are affine in the last KMAX
// solutions of the block before it (its incoming "state", entry e = the
// solution e + 1 rows back), so
//   tbsv_rows_kernel<kResponses>: every block at once, lane c < KMAX solves
//     the block from state e_c with a zero right-hand side and lane KMAX from
//     a zero state with b; the final windows give the k x k map H and the
//     offset z of w_out = z + H w_in, stored as M[block][e][c] (c = KMAX: z);
//   tbsv_chain_kernel: one wave runs w_b = z_b + H_b w_{b-1} over the blocks
//     (a k x k product per block instead of blen rows);
//   tbsv_rows_kernel<kFinish>: every block at once, re-solved from its
//     incoming state W[b - 1].
// Short systems use tbsv_rows_kernel<kWhole>: one wave over all rows.
//
// Row step (all modes): a workgroup stages 64 rows at a time in LDS (row r:
// a(d) at [d - 1], the pivot's reciprocal at [KMAX], b at [KMAX + 1]), rows
// past the block padded with zeros so every 64-row pass runs whole; the
// next 64 rows are loaded into registers meanwhile.  The pass is unrolled in
// groups of KMAX rows with a circular register window (slot u holds the
// group's row u), so a row is KMAX multiply-adds in one chain (the previous
// row's solution last) and one multiply.
enum TbsvMode : int { kWhole = 0, kResponses = 1, kFinish = 2 };

// Value of v in lane l (uniform l), through scalar registers.
template <class T>
__device__ __forceinline__ T lane_value(T v, int l) {
  if constexpr (sizeof(T) == 4) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
  } else {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffff), l);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
  }
}

template <class T, int KMAX>
constexpr int tbsv_pitch() {  // M row pitch: H row, z, padding to 16 bytes
  return KMAX + 16 / (int)sizeof(T);
}

template <class T, int KMAX, int MODE>
__global__ __launch_bounds__(kWave) void tbsv_rows_kernel(int64_t n, int64_t k, int fwd, int col, int64_t c0,
                                                          int unit, const T* __restrict__ A, int64_t lda, T* x,
                                                          int64_t incx, int64_t blen, T* M, const T* W) {
  constexpr int P = tbsv_pitch<T, KMAX>();
  constexpr int VW = 16 / (int)sizeof(T);  // values per 16-byte LDS read
  constexpr int RP = KMAX + VW;            // LDS row pitch
  constexpr int NV = (KMAX + 2 + VW - 1) / VW;
  static_assert(NV * VW <= RP && KMAX < kWave && kWave % KMAX == 0, "tbsv row layout");
  using V = typename std::conditional<sizeof(T) == 4, float4, double2>::type;
  __shared__ __align__(16) T arow[kWave * RP];
  __shared__ T xs[kWave];
  const int lane = threadIdx.x;
  const int64_t blk = blockIdx.x;
  const int64_t sb = blk * blen;
  const int64_t se = sb + blen < n ? sb + blen : n;
  const int64_t passes = ceil_div(se - sb, kWave);
  T ra[KMAX + 1], rb = T(0);
  auto row_of = [&](int64_t s) { return fwd ? s : n - 1 - s; };
  // fetch only issues loads (all unconditional: from the row's diagonal
  // where the entry lies outside the band or the matrix); stash masks them
  // and writes LDS after the pass that used the previous rows.
  auto fetch = [&](int64_t pass) {
    const int64_t s = sb + pass * kWave + lane;
    const int64_t i = row_of(s < se ? s : se - 1);
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
  auto stash = [&](int64_t pass) {
    const int64_t s = sb + pass * kWave + lane;
    const bool real = s < se;
    const int64_t i = row_of(real ? s : se - 1);
    T* ar = arow + lane * RP;
#pragma unroll
    for (int d = 1; d <= KMAX; ++d) {
      const int64_t j = fwd ? i - d : i + d;
      ar[d - 1] = (real && d <= k && j >= 0 && j < n) ? ra[d] : T(0);
    }
    ar[KMAX] = real ? (unit ? T(1) : T(1) / ra[0]) : T(0);
    ar[KMAX + 1] = real ? rb : T(0);
  };
  T xw[KMAX];
#pragma unroll
  for (int u = 0; u < KMAX; ++u) {
    const int e = KMAX - 1 - u;  // before row 0, slot u holds state entry e
    if constexpr (MODE == kWhole) xw[u] = T(0);
    if constexpr (MODE == kResponses) xw[u] = lane == e ? T(1) : T(0);
    if constexpr (MODE == kFinish) xw[u] = blk > 0 ? W[(blk - 1) * kWave + e] : T(0);
  }
  fetch(0);
  stash(0);
  __syncthreads();
  // LDS rows are read LDD rows ahead into a ring of register rows; the
  // compiler barriers keep each read where it is written.
  constexpr int LDD = (KMAX + 2) * (int)sizeof(T) <= 160 ? 2 : 1;
  T rv[LDD + 1][NV * VW];
  auto lds_row = [&](int slot, int r) {
    const V* p = reinterpret_cast<const V*>(arow + r * RP);
#pragma unroll
    for (int q = 0; q < NV; ++q) {
      const V v = p[q];
      if constexpr (sizeof(T) == 4) {
        rv[slot][4 * q] = v.x;
        rv[slot][4 * q + 1] = v.y;
        rv[slot][4 * q + 2] = v.z;
        rv[slot][4 * q + 3] = v.w;
      } else {
        rv[slot][2 * q] = v.x;
        rv[slot][2 * q + 1] = v.y;
      }
    }
  };
  for (int64_t pass = 0; pass < passes; ++pass) {
    if (pass + 1 < passes) fetch(pass + 1);
#pragma unroll
    for (int r = 0; r < LDD; ++r) lds_row(r, r);
#pragma unroll
    for (int g = 0; g < kWave / KMAX; ++g) {
#pragma unroll
      for (int u = 0; u < KMAX; ++u) {
        const int r = g * KMAX + u;
        if (r + LDD < kWave) lds_row((r + LDD) % (LDD + 1), r + LDD);
        asm volatile("" ::: "memory");
        __builtin_amdgcn_sched_barrier(0);
        const T* ar = rv[r % (LDD + 1)];
        // one multiply-add chain, farthest term first (as the reference's
        // loop runs): a wave issues a VALU instruction about every 7 cycles
        // dependent or not, so split accumulators would only add the adds
        // that join them
        T t = ar[KMAX + 1];
        if constexpr (MODE == kResponses) t = lane == KMAX ? t : T(0);
#pragma unroll
        for (int d = KMAX; d >= 1; --d) t -= ar[d - 1] * xw[(u - d + KMAX) % KMAX];
        xw[u] = t * ar[KMAX];
      }
      // Every mode stores the group's solutions (kResponses does not use
      // them): without this store between groups the compiler keeps enough
      // values live to spill.
      if (lane == 0) {
#pragma unroll
        for (int u = 0; u < KMAX; ++u) xs[g * KMAX + u] = xw[u];
      }
    }
    __syncthreads();
    if constexpr (MODE != kResponses) {
      const int64_t s = sb + pass * kWave + lane;
      if (s < se) x[row_of(s) * incx] = xs[lane];
    }
    if (pass + 1 < passes) stash(pass + 1);
    __syncthreads();
  }
  if constexpr (MODE == kResponses) {
    // full blocks only, so the window ends on slot KMAX - 1 (the last row)
    if (lane < P) {
#pragma unroll
      for (int e = 0; e < KMAX; ++e) M[(blk * KMAX + e) * P + lane] = xw[KMAX - 1 - e];
    }
  }
}

// w_b = z_b + H_b w_{b-1} for b < steps (w_{-1} = 0), one wave: lane e
// keeps state entry e and takes row e of block b's map, prefetched D blocks
// ahead (D sized so the loads in flight stay within the vector memory
// counter).  The loop body has no branch: M holds D more blocks than are
// used and W room for the padded steps (their values are never read), and
// the sched barriers keep the loads in issue order so the waits count them.
// W holds kWave values per block.
template <class T, int KMAX, int D>
__global__ __launch_bounds__(kWave) void tbsv_chain_kernel(int64_t steps, const T* __restrict__ M, T* W) {
  constexpr int P = tbsv_pitch<T, KMAX>();
  const int lane = threadIdx.x;
  const int e = lane < KMAX ? lane : KMAX - 1;
  T buf[D][KMAX + 1];
  const T* row = M + e * P;
  auto load = [&](int q) {
#pragma unroll
    for (int c = 0; c <= KMAX; ++c) buf[q][c] = row[c];
    row += KMAX * P;
  };
#pragma unroll
  for (int q = 0; q < D; ++q) {
    load(q);
    __builtin_amdgcn_sched_barrier(0);
  }
  T w = T(0);
  T* wp = W + lane;
  for (int64_t b0 = 0; b0 < steps; b0 += D) {
#pragma unroll
    for (int q = 0; q < D; ++q) {
      T wv[KMAX];
#pragma unroll
      for (int c = 0; c < KMAX; ++c) wv[c] = lane_value(w, c);
      // four accumulators: the products start as the first lane values
      // arrive while later readlanes are still issuing (one chain measured
      // 26.8 against 25.7 us for 255 steps, FP32 k = 16)
      T acc[4] = {buf[q][KMAX], T(0), T(0), T(0)};
#pragma unroll
      for (int c = 0; c < KMAX; ++c) acc[c & 3] += buf[q][c] * wv[c];
      w = (acc[0] + acc[1]) + (acc[2] + acc[3]);
      *wp = w;
      wp += kWave;
      load(q);
      __builtin_amdgcn_sched_barrier(0);
    }
  }
}

template <class T, int KMAX>
constexpr int tbsv_chain_depth() {
  constexpr int loads = ((KMAX + 1) * (int)sizeof(T) + 15) / 16;  // per block row
  constexpr int d = 60 / (loads + 1);
  return d < 2 ? 2 : (d > 8 ? 8 : d);
}

// Systems of at least this many rows take the blocked solve.
constexpr int64_t kTbsvBlockedMin = 512;

// Block length: the blocked solve costs about 2 blen row steps plus n / blen
// chain steps of similar length, so blen near sqrt(n / 2), as a power of two
// (64 .. 8192).
inline int64_t tbsv_block_rows(int64_t n) {
  int64_t blen = kWave;
  while (blen < 8192 && 4 * blen * blen <= (n * 6) / 5) blen *= 2;
  return blen;
}

template <class T, int KMAX>
void launch_tbsv_narrow(Ctx& c, int64_t n, int64_t k, bool fwd, bool col, int64_t c0, bool unit, const T* A,
                        int64_t lda, T* x, int64_t incx) {
  if (n < kTbsvBlockedMin) {
    hipLaunchKernelGGL((tbsv_rows_kernel<T, KMAX, kWhole>), dim3(1), dim3(kWave), 0, c.stream, n, k, (int)fwd,
                       (int)col, c0, (int)unit, A, lda, x, incx, n, (T*)nullptr, (const T*)nullptr);
    EXO_LAUNCH_CHECK();
    return;
  }
  constexpr int D = tbsv_chain_depth<T, KMAX>();
  constexpr int P = tbsv_pitch<T, KMAX>();
  const int64_t blen = tbsv_block_rows(n);
  const int64_t nb = ceil_div(n, blen);
  const int64_t rows = nb + 2 * D;  // blocks of M and W, with the chain's padding
  const size_t m_elems = (size_t)rows * KMAX * P;
  T* M = static_cast<T*>(c.scratch(sizeof(T) * (m_elems + (size_t)rows * kWave), 3));
  T* W = M + m_elems;
  hipLaunchKernelGGL((tbsv_rows_kernel<T, KMAX, kResponses>), dim3((unsigned)(nb - 1)), dim3(kWave), 0, c.stream, n,
                     k, (int)fwd, (int)col, c0, (int)unit, A, lda, x, incx, blen, M, (const T*)nullptr);
  EXO_LAUNCH_CHECK();
  hipLaunchKernelGGL((tbsv_chain_kernel<T, KMAX, D>), dim3(1), dim3(kWave), 0, c.stream, nb - 1, (const T*)M, W);
  EXO_LAUNCH_CHECK();
  hipLaunchKernelGGL((tbsv_rows_kernel<T, KMAX, kFinish>), dim3((unsigned)nb), dim3(kWave), 0, c.stream, n, k,
                     (int)fwd, (int)col, c0, (int)unit, A, lda, x, incx, blen, (T*)nullptr, (const T*)W);
  EXO_LAUNCH_CHECK();
}

}  // namespace tbsv_blocked

namespace {

template <class T>
void launch_tbsv(Ctx& c, int64_t n, int64_t k, bool fwd, bool col, int64_t c0, bool unit, const T* A, int64_t lda,
                 T* x, int64_t incx) {
  if (n <= 0) return;
  // Measured on MI355X (s/d, n = 64 .. 16384): trsv's block solver takes
  // about 15 us + 0.65 us per 64 rows; the one-wave solve (n < 512) about
  // 6 us + n row steps of 0.005 + 0.0055 KMAX us; the blocked solve (n >= 512)
  // 13-40 us up to n = 2048, where the block solver overtakes it for
  // KMAX = 32 up to n ~ 1536.  Each system goes to the cheaper of the two.
  const int km = k <= 4 ? 4 : k <= 8 ? 8 : k <= 16 ? 16 : 32;
  const bool short_band = n < tbsv_blocked::kTbsvBlockedMin &&
                          6.0 + (double)n * (0.005 + 0.0055 * km) > 15.0 + (double)n * (0.65 / 64);
  if (k > 32 || short_band || (km == 32 && n <= 1536)) {
    // the block solver over the band seen as a dense triangle (zero outside it)
    ops::trsv_band_rm<T>(c, fwd, col, unit, n, k, A, lda, c0, x, incx);
    return;
  }
  using tbsv_blocked::launch_tbsv_narrow;
  if (k <= 4) return launch_tbsv_narrow<T, 4>(c, n, k, fwd, col, c0, unit, A, lda, x, incx);
  if (k <= 8) return launch_tbsv_narrow<T, 8>(c, n, k, fwd, col, c0, unit, A, lda, x, incx);
  if (k <= 16) return launch_tbsv_narrow<T, 16>(c, n, k, fwd, col, c0, unit, A, lda, x, incx);
  launch_tbsv_narrow<T, 32>(c, n, k, fwd, col, c0, unit, A, lda, x, incx);
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
