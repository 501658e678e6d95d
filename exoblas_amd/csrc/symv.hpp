// symv for gfx950: y = beta y + alpha A x with A symmetric, one triangle
// stored (src/level2/symv.py:10-35), as two kernels over the stored tiles
// of tiles.hpp: tile_kernel reads every stored element once and produces
// per-tile row and column partial sums, finish_kernel adds them in a fixed
// order.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>

#include "device.hpp"
#include "tiles.hpp"

namespace exoblas {
namespace symv {

constexpr int kNT = 256;
constexpr int kRB = 64;  // rows per tile

constexpr int kQ = 2;   // 16-byte vectors per lane and tile row: W = 64 x VN x kQ columns
constexpr int kFW = 16;  // waves of finish_kernel

// One workgroup per RB x W tile of the stored triangle (tiles.hpp), RB rows
// and W = 64 x VN x Q columns, each element read once.  Wave w takes rows
// r0 + w RB / 4 .. (RW = RB / 4 rows) across the whole tile: lane l loads the
// 16-byte column vectors l, l + 64, .., l + 64 (Q - 1) of each row (4 rows
// in flight), keeping RW row sums (A_ij x_j) and Q x VN column sums (A_ij
// x_i, off the diagonal) in registers.  (Measured on MI355X, n = 8192 FP32:
// Q = 2 takes 9 % less time than Q = 4, whose 576 tiles of 256 KiB leave
// the CUs that get a third tile running alone at the end; Q = 1 and 32-row
// tiles gain nothing more.)  Row sums are reduced over the wave's lanes
// (a butterfly: lane l ends with row l % RW) into rowpart[c][row]; column sums
// are added over the 4 waves through LDS into colpart[b][column].
// finish_kernel adds both in a fixed order, so results are
// deterministic.
// VEC: 16-byte loads (aligned rows and n a multiple of VN).
template <class T, bool UPPER, bool VEC, int RB, int Q = kQ>
__global__ __launch_bounds__(kNT) void tile_kernel(int64_t n, const T* __restrict__ A, int64_t lda,
                                                        const T* __restrict__ x, int64_t incx, T* rowpart,
                                                        T* colpart, tiles::TileList tl) {
  using V = typename Vec16<T>::type;
  constexpr int VN = Vec16<T>::n, W = kWave * VN * Q;
  constexpr int NW = kNT / kWave, RW = RB / NW, U = 2;  // RW rows per wave
  static_assert(RW >= 2 * U && RW % (2 * U) == 0 && RW <= 16, "rows per wave");
  __shared__ T red[NW][W];
  int64_t b, c;
  tl.tile(blockIdx.x, b, c);
  const int lane = threadIdx.x % kWave, wave = threadIdx.x / kWave;
  const int64_t r0 = b * RB, cw = c * W;
  // tile-local coordinates: rows il < nrow, columns jl < ncol are inside the
  // matrix; element (il, jl) is stored when jl + dg <= il (lower) or
  // jl + dg >= il (upper), which only the tiles crossing the diagonal check
  const int nrow = (int)(n - r0 < RB ? n - r0 : RB);
  const int ncol = (int)(n - cw < W ? n - cw : W);
  const bool diag = cw < r0 + RB && cw + W > r0;
  const int dg = diag ? (int)(cw - r0) : 0;
  const int wr = wave * RW;  // this wave's first tile row
  T xc[Q][VN], cacc[Q][VN];
#pragma unroll
  for (int q = 0; q < Q; ++q)
#pragma unroll
    for (int k = 0; k < VN; ++k) {
      const int jl = (lane + q * kWave) * VN + k;
      xc[q][k] = jl < ncol ? x[(cw + jl) * incx] : T(0);
      cacc[q][k] = T(0);
    }
  T racc[RW];
#pragma unroll
  for (int r = 0; r < RW; ++r) racc[r] = T(0);
  if (VEC && nrow == RB && ncol == W) {
    // a tile inside the matrix: rows in pairs, the next pair's loads issued
    // before this pair is used (a scheduling barrier keeps the compiler from
    // hoisting later rounds' loads as well).  Every vector is loaded without
    // a condition, so the compiler can count outstanding loads and wait for
    // only the ones it needs; tiles on the diagonal zero the other
    // triangle's elements after the load.
    auto load = [&](int il0, V (&a)[U][Q], T (&xr)[U]) {
#pragma unroll
      for (int u = 0; u < U; ++u) {
        xr[u] = x[(r0 + il0 + u) * incx];
        const T* row = A + (r0 + il0 + u) * lda + cw + lane * VN;
#pragma unroll
        for (int q = 0; q < Q; ++q) a[u][q] = ld_stream<T>(row + q * kWave * VN);
      }
    };
    // DIAG: the tile crosses the diagonal (mask the other triangle)
    auto run = [&](auto diag_tag) {
      constexpr bool DIAG = decltype(diag_tag)::value;
      auto accumulate = [&](int il0, int rr, const V (&a)[U][Q], const T (&xr)[U]) {
#pragma unroll
        for (int u = 0; u < U; ++u)
#pragma unroll
          for (int q = 0; q < Q; ++q)
#pragma unroll
            for (int k = 0; k < VN; ++k) {
              T v = vget<T>(a[u][q], k);
              if (DIAG) {
                const int il = il0 + u;
                const int jd = (lane + q * kWave) * VN + k + dg;  // j - i = jd - il
                if (UPPER ? jd < il : jd > il) v = T(0);          // the other triangle
                racc[rr + u] += v * xc[q][k];
                cacc[q][k] += jd == il ? T(0) : v * xr[u];
              } else {
                racc[rr + u] += v * xc[q][k];
                cacc[q][k] += v * xr[u];
              }
            }
      };
      V a0[U][Q], a1[U][Q];
      T x0[U], x1[U];
      load(wr, a0, x0);
#pragma unroll
      for (int rr = 0; rr < RW; rr += 2 * U) {
        load(wr + rr + U, a1, x1);
        __builtin_amdgcn_sched_barrier(0);
        accumulate(wr + rr, rr, a0, x0);
        if (rr + 2 * U < RW) load(wr + rr + 2 * U, a0, x0);
        __builtin_amdgcn_sched_barrier(0);
        accumulate(wr + rr + U, rr + U, a1, x1);
      }
    };
    if (diag)
      run(std::true_type{});
    else
      run(std::false_type{});
  } else {
    // a tile on the matrix edge, misaligned rows or n not a multiple of VN:
    // one row at a time, element by element, the other triangle masked
#pragma unroll
    for (int rr = 0; rr < RW; ++rr) {
      const int il = wr + rr;
      if (il >= nrow) break;
      const T xr = x[(r0 + il) * incx];
      const T* row = A + (r0 + il) * lda + cw + lane * VN;
#pragma unroll
      for (int q = 0; q < Q; ++q) {
        const int jl0 = (lane + q * kWave) * VN;
        bool m = jl0 < ncol;
        if (diag) m = m && (UPPER ? jl0 + VN - 1 + dg >= il : jl0 + dg <= il);
        if (!m) continue;
#pragma unroll
        for (int k = 0; k < VN; ++k) {
          const int jd = jl0 + k + dg;  // j - i = jd - il
          T v = jl0 + k < ncol ? row[q * kWave * VN + k] : T(0);
          if (diag && (UPPER ? jd < il : jd > il)) v = T(0);  // the other triangle
          racc[rr] += v * xc[q][k];
          cacc[q][k] += (diag && jd == il) ? T(0) : v * xr;
        }
      }
    }
  }
  // butterfly over lanes: after the stage with offset o (RW / 2, ..., 1) a
  // lane keeps the rows whose bit o matches its own, adding its partner's
  // sums for them; then the lanes holding the same row (l % RW) are added
#pragma unroll
  for (int o = RW / 2; o >= 1; o /= 2) {
    const bool hi = lane & o;
#pragma unroll
    for (int r = 0; r < o; ++r) {
      const T keep = hi ? racc[r + o] : racc[r];
      const T send = hi ? racc[r] : racc[r + o];
      racc[r] = keep + __shfl_xor(send, o, kWave);
    }
  }
  T rs = racc[0];
#pragma unroll
  for (int o = RW; o < kWave; o *= 2) rs += __shfl_xor(rs, o, kWave);
  const int64_t npad = tl.nb * RB;
  if (lane < RW && wr + lane < nrow) rowpart[c * npad + r0 + wr + lane] = rs;
#pragma unroll
  for (int q = 0; q < Q; ++q)
#pragma unroll
    for (int k = 0; k < VN; ++k) red[wave][(lane + q * kWave) * VN + k] = cacc[q][k];
  __syncthreads();
  for (int e = threadIdx.x; e < ncol; e += kNT) {
    T s = red[0][e];
#pragma unroll
    for (int w = 1; w < NW; ++w) s += red[w][e];
    colpart[b * npad + cw + e] = s;
  }
}

// y_i = beta y_i + alpha (sum of row i's tile row sums, in column-block
// order, + sum of the column sums for column i, in row-block order).  A
// workgroup finishes 64 consecutive i; its NW waves take every NW-th partial
// (8 loads in flight per lane) and are combined in wave order through LDS,
// so the sum's association is fixed (deterministic).
template <class T, bool UPPER, int RB, int Q = kQ, int NW = kFW>
__global__ __launch_bounds__(NW * kWave) void finish_kernel(int64_t n, T alpha, T beta, const T* rowpart,
                                                          const T* colpart, tiles::TileList tl, T* y, int64_t incy) {
  constexpr int VN = Vec16<T>::n, W = kWave * VN * Q;
  __shared__ T red[2][NW][kWave];
  const int lane = threadIdx.x % kWave, wave = threadIdx.x / kWave;
  const int64_t i = (int64_t)blockIdx.x * kWave + lane;
  const int64_t npad = tl.nb * RB;
  T s = T(0), t = T(0);
  if (i < n) {
    const int64_t b = i / RB, c = i / W;
    // row block b's tiles: column blocks [c_lo, c_hi]
    const int64_t c_lo = tl.first(b), c_hi = UPPER ? tl.nc - 1 : b / tl.K;
    // tiles over column block c: row blocks [b_lo, b_hi]
    const int64_t b_lo = UPPER ? 0 : c * tl.K;
    const int64_t b_hi = UPPER ? ((c + 1) * tl.K - 1 < tl.nb - 1 ? (c + 1) * tl.K - 1 : tl.nb - 1) : tl.nb - 1;
    for (int64_t cc = c_lo + wave; cc <= c_hi; cc += NW) s += rowpart[cc * npad + i];
    int64_t bb = b_lo + wave;
    for (; bb + 7 * NW <= b_hi; bb += 8 * NW) {
      T v[8];
#pragma unroll
      for (int k = 0; k < 8; ++k) v[k] = colpart[(bb + k * NW) * npad + i];
#pragma unroll
      for (int k = 0; k < 8; ++k) t += v[k];
    }
    for (; bb <= b_hi; bb += NW) t += colpart[bb * npad + i];
  }
  red[0][wave][lane] = s;
  red[1][wave][lane] = t;
  __syncthreads();
  if (wave == 0 && i < n) {
    T rs = red[0][0][lane], cs = red[1][0][lane];
#pragma unroll
    for (int w = 1; w < NW; ++w) {
      rs += red[0][w][lane];
      cs += red[1][w][lane];
    }
    y[i * incy] = beta * y[i * incy] + alpha * (rs + cs);
  }
}

}  // namespace symv
}  // namespace exoblas
