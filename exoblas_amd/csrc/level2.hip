// Level-2 kernels for gfx950 (all HBM-bound on the matrix stream).
//
// gemv-N  long rows with a contiguous x: whole-workgroup rows, x held in
//         registers, two rows at a time; other tall matrices: one row per
//         wave; short ones: R rows per wave, x staged through LDS in panels,
//         wide matrices split their columns over workgroups (partials + a
//         finishing pass).
// gemv-T  a workgroup owns 64*VN consecutive columns and a chunk of rows;
//         lanes hold column sums, waves split the rows and reduce through LDS;
//         row splits are summed by the last workgroup of each column block.
// ger, syr, syr2   streaming rank-1/2 updates over the tiles that meet the
//         stored part (tiles.hpp), several rows of loads in flight per lane.
// symv    64-row tiles of the stored triangle, each element read once for
//         both its row sum (A_ij x_j -> y_i) and its column sum (A_ij x_i ->
//         y_j); per-tile partials are summed in a fixed order, so results are
//         deterministic.
// trmv    the triangular gemv-N / gemv-T kernels over a copy of x.
#include "device.hpp"
#include "ops.hpp"
#include "tagged.hpp"
#include "symv.hpp"
#include "tiles.hpp"

namespace exoblas {
namespace {

constexpr int kNT = 256;
constexpr int kNoTrans = 111;
constexpr int kUpper = 121;
constexpr int kUnit = 132;

enum Tri { kFull = 0, kLowerTri = 1, kUpperTri = 2 };

// ------------------------------------------------------------ gemv-N

// y[r] = beta*y[r] + alpha * sum_{c in row r's range} A[r,c] x[c]   (TRI == kFull;
// with ow set, y is not read: beta*0 in its place, for callers with beta == 0)
// For TRI != kFull the row range is the triangle; UNIT skips the diagonal and
// adds x[r]; the result overwrites y (trmv).
// With gridDim.y > 1 each workgroup row covers `cols_per` columns and writes
// its raw sums to part[blockIdx.y * M + r].
template <class T, int R, int P, bool VEC, int TRI, bool UNIT>
__global__ __launch_bounds__(kNT) void gemv_n_kernel(int64_t M, int64_t N, T alpha, T beta,
                                                     const T* __restrict__ A, int64_t lda,
                                                     const T* __restrict__ x, int64_t incx, T* y,
                                                     int64_t incy, int64_t cols_per, T* part, int ow) {
  using V = typename Vec16<T>::type;
  constexpr int VN = Vec16<T>::n;
  __shared__ __align__(16) T xs[P];
  const int wave = threadIdx.x / kWave;
  const int lane = threadIdx.x % kWave;
  const int64_t row0 = ((int64_t)blockIdx.x * (kNT / kWave) + wave) * R;
  int64_t c_beg = (int64_t)blockIdx.y * cols_per;
  int64_t c_end = c_beg + cols_per < N ? c_beg + cols_per : N;
  if (TRI != kFull) {  // columns any of this workgroup's rows can touch
    const int64_t rb0 = (int64_t)blockIdx.x * (kNT / kWave) * R;
    const int64_t rb1 = rb0 + (kNT / kWave) * R;
    if (TRI == kLowerTri && c_end > rb1) c_end = rb1;
    if (TRI == kUpperTri && c_beg < rb0) c_beg = rb0 - rb0 % VN;
  }
  T acc[R];
#pragma unroll
  for (int r = 0; r < R; ++r) acc[r] = T(0);

  for (int64_t p0 = c_beg; p0 < c_end; p0 += P) {
    const int pn = (int)(c_end - p0 < P ? c_end - p0 : P);
    __syncthreads();
    for (int i = threadIdx.x; i < pn; i += kNT) xs[i] = x[(p0 + i) * incx];
    __syncthreads();
    if (VEC) {
      const int nv = pn / VN;
      auto accum = [&](const V (&av)[R], const V& xv, int64_t cb) {
#pragma unroll
        for (int r = 0; r < R; ++r) {
          const int64_t row = row0 + r;
          if (row < M) {
#pragma unroll
            for (int k = 0; k < VN; ++k) {
              const int64_t col = cb + k;
              bool in = true;
              if (TRI == kLowerTri) in = UNIT ? col < row : col <= row;
              if (TRI == kUpperTri) in = UNIT ? col > row : col >= row;
              if (in) acc[r] += vget<T>(av[r], k) * vget<T>(xv, k);
            }
          }
        }
      };
      // UV vectors per lane per round: R * UV = 8 streaming loads in flight
      constexpr int UV = R >= 8 ? 1 : 8 / R;
      int v = lane;
      for (; v + (UV - 1) * kWave < nv; v += UV * kWave) {
        V av[UV][R];
#pragma unroll
        for (int u = 0; u < UV; ++u)
#pragma unroll
          for (int r = 0; r < R; ++r) {
            const int64_t row = row0 + r;
            if (row < M) av[u][r] = ld_stream<T>(A + row * lda + p0 + (int64_t)(v + u * kWave) * VN);
          }
#pragma unroll
        for (int u = 0; u < UV; ++u)
          accum(av[u], reinterpret_cast<const V*>(xs)[v + u * kWave], p0 + (int64_t)(v + u * kWave) * VN);
      }
      for (; v < nv; v += kWave) {
        V av[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
          const int64_t row = row0 + r;
          if (row < M) av[r] = ld_stream<T>(A + row * lda + p0 + (int64_t)v * VN);
        }
        accum(av, reinterpret_cast<const V*>(xs)[v], p0 + (int64_t)v * VN);
      }
      for (int i = nv * VN + lane; i < pn; i += kWave) {
#pragma unroll
        for (int r = 0; r < R; ++r) {
          const int64_t row = row0 + r, col = p0 + i;
          bool in = row < M;
          if (TRI == kLowerTri) in = in && (UNIT ? col < row : col <= row);
          if (TRI == kUpperTri) in = in && (UNIT ? col > row : col >= row);
          if (in) acc[r] += A[row * lda + col] * xs[i];
        }
      }
    } else {
      for (int i = lane; i < pn; i += kWave) {
#pragma unroll
        for (int r = 0; r < R; ++r) {
          const int64_t row = row0 + r, col = p0 + i;
          bool in = row < M;
          if (TRI == kLowerTri) in = in && (UNIT ? col < row : col <= row);
          if (TRI == kUpperTri) in = in && (UNIT ? col > row : col >= row);
          if (in) acc[r] += A[row * lda + col] * xs[i];
        }
      }
    }
  }
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const T s = wave_sum(acc[r]);
    const int64_t row = row0 + r;
    if (lane == 0 && row < M) {
      if (gridDim.y > 1) {
        part[(int64_t)blockIdx.y * M + row] = s;
      } else if (TRI != kFull) {
        y[row * incy] = UNIT ? s + x[row * incx] : s;
      } else {
        y[row * incy] = beta * (ow ? T(0) : y[row * incy]) + alpha * s;
      }
    }
  }
}

// Tall matrices with a contiguous, aligned x: one row per wave, x read
// straight from global memory (it stays in L2) and 8 streaming A vectors per
// lane in flight; no LDS and no barriers.  (Best of a sweep on MI355X over
// rows per wave, vectors in flight, block size and LDS-staged x: 6.4 TB/s
// for 8192 x 8192 FP32.)  Same TRI / UNIT meaning as gemv_n_kernel.
template <class T, int TRI, bool UNIT>
__global__ __launch_bounds__(kNT) void gemv_n_row_kernel(int64_t M, int64_t N, T alpha, T beta,
                                                         const T* __restrict__ A, int64_t lda,
                                                         const T* __restrict__ x, T* y, int64_t incy, int ow) {
  using V = typename Vec16<T>::type;
  constexpr int VN = Vec16<T>::n;
  constexpr int UV = 8;
  const int wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
  int64_t row = (int64_t)blockIdx.x * (kNT / kWave) + wave;
  if (TRI != kFull) {
    // a triangle's rows pair up to equal length: waves 2j and 2j + 1 of
    // workgroup w take rows i = 2w + j and M - 1 - i, so every workgroup
    // reads about M + 1 elements (dealt in order, 8 workgroups per CU at
    // M = 8192 gave the CUs with the long rows 12 % more than the mean)
    const int64_t h = (M + 1) / 2;  // rows i < h from the top, M - h from the bottom
    const int64_t i = 2 * (int64_t)blockIdx.x + wave / 2;
    if (wave % 2 == 0) {
      if (i >= h) return;
      row = i;
    } else {
      if (i >= M - h) return;
      row = M - 1 - i;
    }
  }
  if (row >= M) return;
  int64_t jl = 0, jh = N;  // columns [jl, jh) of this row
  if (TRI == kLowerTri) jh = UNIT ? row : row + 1;
  if (TRI == kUpperTri) jl = UNIT ? row + 1 : row;
  const int64_t vl = ceil_div(jl, VN);
  const int64_t vh = std::max<int64_t>(jh / VN, vl);  // whole vectors [vl, vh)
  const T* Ar = A + row * lda;
  T acc = T(0);
  int64_t v = vl + lane;
  for (; v + (UV - 1) * kWave < vh; v += UV * kWave) {
    V av[UV], xv[UV];
#pragma unroll
    for (int u = 0; u < UV; ++u) {
      av[u] = ld_stream<T>(Ar + (v + u * kWave) * VN);
      xv[u] = *reinterpret_cast<const V*>(x + (v + u * kWave) * VN);
    }
#pragma unroll
    for (int u = 0; u < UV; ++u)
#pragma unroll
      for (int k = 0; k < VN; ++k) acc += vget<T>(av[u], k) * vget<T>(xv[u], k);
  }
  for (; v < vh; v += kWave) {
    const V av = ld_stream<T>(Ar + v * VN);
    const V xv = *reinterpret_cast<const V*>(x + v * VN);
#pragma unroll
    for (int k = 0; k < VN; ++k) acc += vget<T>(av, k) * vget<T>(xv, k);
  }
  // the scalar columns before and after the whole vectors
  const int64_t hend = std::min<int64_t>(vl * VN, jh);
  for (int64_t j = jl + lane; j < hend; j += kWave) acc += Ar[j] * x[j];
  for (int64_t j = std::max<int64_t>(vh * VN, hend) + lane; j < jh; j += kWave) acc += Ar[j] * x[j];
  const T s = wave_sum(acc);
  if (lane == 0) {
    if (TRI != kFull)
      y[row * incy] = UNIT ? s + x[row] : s;
    else
      y[row * incy] = beta * (ow ? T(0) : y[row * incy]) + alpha * s;
  }
}

// gemv-N with a contiguous, aligned x of at most kNT * VN * UV elements: the
// workgroup keeps its slice of x in registers (thread t holds vectors t,
// t + kNT, ...) for all its rows, and reads R whole rows at a time with every
// thread (R * UV streaming loads in flight per lane), so each row is fetched
// as one contiguous burst and x is never re-read.  Rows blockIdx.x * R +
// k * gridDim.x * R ...: dealt round-robin, so a triangle's rows of uneven
// length balance over the workgroups.  TRI / UNIT as in gemv_n_kernel (a
// lane loads only the vectors that meet its row's part of the triangle).
template <class T, int UV, int R, int TRI = kFull, bool UNIT = false>
__global__ __launch_bounds__(kNT) void gemv_n_wg_kernel(int64_t M, int64_t N, T alpha, T beta,
                                                        const T* __restrict__ A, int64_t lda,
                                                        const T* __restrict__ x, T* y, int64_t incy, int ow) {
  using V = typename Vec16<T>::type;
  constexpr int VN = Vec16<T>::n;
  constexpr int NW = kNT / kWave;
  __shared__ T red[R][NW];
  const int t = threadIdx.x, wave = t / kWave, lane = t % kWave;
  const int64_t nv = N / VN;
  const int tail = (int)(N - nv * VN);
  V xr[UV];
#pragma unroll
  for (int u = 0; u < UV; ++u) {
    const int64_t v = t + (int64_t)u * kNT;
    if (v < nv) {
      xr[u] = *reinterpret_cast<const V*>(x + v * VN);
    } else {
#pragma unroll
      for (int k = 0; k < VN; ++k) vget<T>(xr[u], k) = T(0);
    }
  }
  const T xt = t < tail ? x[nv * VN + t] : T(0);
  // element (row, col) inside the row's part of the triangle
  auto in_tri = [](int64_t row, int64_t col) {
    if (TRI == kLowerTri) return UNIT ? col < row : col <= row;
    if (TRI == kUpperTri) return UNIT ? col > row : col >= row;
    return true;
  };
  // vector v meets the row's part
  auto meets = [](int64_t row, int64_t v) {
    if (TRI == kLowerTri) return v * VN <= (UNIT ? row - 1 : row);
    if (TRI == kUpperTri) return v * VN + VN - 1 >= (UNIT ? row + 1 : row);
    return true;
  };
  for (int64_t r0 = (int64_t)blockIdx.x * R; r0 < M; r0 += (int64_t)gridDim.x * R) {
    V av[R][UV];
#pragma unroll
    for (int rr = 0; rr < R; ++rr) {
      const int64_t row = r0 + rr < M ? r0 + rr : M - 1;
      const T* Ar = A + row * lda;
#pragma unroll
      for (int u = 0; u < UV; ++u) {
        const int64_t v = t + (int64_t)u * kNT;
        if (v < nv && meets(row, v)) av[rr][u] = ld_stream<T>(Ar + v * VN);
      }
    }
    T acc[R];
#pragma unroll
    for (int rr = 0; rr < R; ++rr) {
      const int64_t row = r0 + rr < M ? r0 + rr : M - 1;
      acc[rr] = T(0);
#pragma unroll
      for (int u = 0; u < UV; ++u) {
        const int64_t v = t + (int64_t)u * kNT;
        if (v < nv && meets(row, v)) {
          if (TRI == kFull) {
#pragma unroll
            for (int k = 0; k < VN; ++k) acc[rr] += vget<T>(av[rr][u], k) * vget<T>(xr[u], k);
          } else {
#pragma unroll
            for (int k = 0; k < VN; ++k)
              acc[rr] += in_tri(row, v * VN + k) ? vget<T>(av[rr][u], k) * vget<T>(xr[u], k) : T(0);
          }
        }
      }
      if (t < tail && in_tri(row, nv * VN + t)) acc[rr] += A[row * lda + nv * VN + t] * xt;
      acc[rr] = wave_sum(acc[rr]);
      if (lane == 0) red[rr][wave] = acc[rr];
    }
    __syncthreads();
    if (t < R && r0 + t < M) {
      T s = red[t][0];
#pragma unroll
      for (int w = 1; w < NW; ++w) s += red[t][w];
      if (TRI != kFull)
        y[(r0 + t) * incy] = UNIT ? s + x[r0 + t] : s;
      else
        y[(r0 + t) * incy] = beta * (ow ? T(0) : y[(r0 + t) * incy]) + alpha * s;
    }
    __syncthreads();
  }
}

// y[r] = beta*y[r] + alpha * sum_s part[s*M + r]  (or the triangular overwrite).
// A workgroup finishes 64 rows; its 4 waves take every 4th split each and
// combine through LDS in a fixed order (deterministic).
template <class T, bool TRI, bool UNIT>
__global__ __launch_bounds__(kNT) void gemv_finish_kernel(int64_t M, int splits, T alpha, T beta,
                                                          const T* part, const T* x, int64_t incx, T* y,
                                                          int64_t incy, int ow) {
  constexpr int NW = kNT / kWave;
  __shared__ T red[NW][kWave];
  const int rl = threadIdx.x % kWave, g = threadIdx.x / kWave;
  const int64_t r = (int64_t)blockIdx.x * kWave + rl;
  T s = T(0);
  if (r < M) {
    int k = g;
    for (; k + 3 * NW < splits; k += 4 * NW)
      s += (part[(int64_t)k * M + r] + part[(int64_t)(k + NW) * M + r]) +
           (part[(int64_t)(k + 2 * NW) * M + r] + part[(int64_t)(k + 3 * NW) * M + r]);
    for (; k < splits; k += NW) s += part[(int64_t)k * M + r];
  }
  red[g][rl] = s;
  __syncthreads();
  if (g == 0 && r < M) {
    T t = red[0][rl];
#pragma unroll
    for (int w = 1; w < NW; ++w) t += red[w][rl];
    if (TRI)
      y[r * incy] = UNIT ? t + x[r * incx] : t;
    else
      y[r * incy] = beta * (ow ? T(0) : y[r * incy]) + alpha * t;
  }
}

// ------------------------------------------------------------ gemv-T

// y[c] = beta*y[c] + alpha * sum_r A[r,c] x[r], A is Nr x Mc (rows x cols).
// Workgroup (bx, by): columns [bx*64*VN, +64*VN), rows [by*rows_per, +rows_per).
// With row splits (gridDim.y > 1) each workgroup publishes its column sums
// as tagged words in `pub` (tagged.hpp) and the last of a column block's
// workgroups (ticket tickets[blockIdx.x]) adds the splits in order 0, 1, ...
// and writes y; *err is the timeout word.  The marker comes from the
// stream's launch generation *gen, advanced by the last column block to
// complete (ticket *lt), so the launch can be replayed from a hipGraph.
template <class T, int RU, bool VEC, int TRI, bool UNIT, int NTT = kNT>
__global__ __launch_bounds__(NTT) void gemv_t_kernel(int64_t Mc, int64_t Nr, T alpha, T beta,
                                                     const T* __restrict__ A, int64_t lda,
                                                     const T* __restrict__ x, int64_t incx, T* y,
                                                     int64_t incy, int64_t rows_per, unsigned* tickets,
                                                     unsigned* gen, unsigned* lt, unsigned* err, unsigned* pub,
                                                     int ow) {
  using V = typename Vec16<T>::type;
  constexpr int VN = Vec16<T>::n;
  constexpr int NW = NTT / kWave;
  __shared__ T red[NW][kWave * VN];
  // The generation is loaded now and stored for the workgroup only at the
  // reduction below, so no wave waits for it before its streaming loads.
  __shared__ unsigned s_marker;
  const unsigned mk = (gridDim.y > 1 && threadIdx.x == 0) ? tagged::read_generation(gen) : 0u;
  const int wave = threadIdx.x / kWave;
  const int lane = threadIdx.x % kWave;
  const int64_t c0 = (int64_t)blockIdx.x * kWave * VN + (int64_t)lane * VN;  // this lane's first column
  int64_t r_beg = (int64_t)blockIdx.y * rows_per;
  int64_t r_end = r_beg + rows_per < Nr ? r_beg + rows_per : Nr;
  const int64_t cb0 = (int64_t)blockIdx.x * kWave * VN, cb1 = cb0 + kWave * VN;
  if (TRI == kLowerTri && r_beg < cb0) r_beg = cb0;  // rows >= column
  if (TRI == kUpperTri && r_end > cb1) r_end = cb1;  // rows <= column
  T acc[VN];
#pragma unroll
  for (int k = 0; k < VN; ++k) acc[k] = T(0);

  auto body = [&](int64_t r) {
    const T xr = x[r * incx];
    if (VEC && c0 + VN <= Mc) {
      const V av = ld_stream<T>(A + r * lda + c0);
#pragma unroll
      for (int k = 0; k < VN; ++k) {
        const int64_t c = c0 + k;
        bool in = true;
        if (TRI == kLowerTri) in = UNIT ? r > c : r >= c;
        if (TRI == kUpperTri) in = UNIT ? r < c : r <= c;
        if (in) acc[k] += vget<T>(av, k) * xr;
      }
    } else {
#pragma unroll
      for (int k = 0; k < VN; ++k) {
        const int64_t c = c0 + k;
        bool in = c < Mc;
        if (TRI == kLowerTri) in = in && (UNIT ? r > c : r >= c);
        if (TRI == kUpperTri) in = in && (UNIT ? r < c : r <= c);
        if (in) acc[k] += A[r * lda + c] * xr;
      }
    }
  };
  int64_t r = r_beg + wave;
  if (TRI == kFull && VEC && cb1 <= Mc) {
    // Interior column block of a full matrix: the same loop without the
    // per-lane edge and triangle tests (which the compiler keeps per row).
    const T* Ac = A + c0;
    for (; r + (RU - 1) * NW < r_end; r += RU * NW) {
      V av[RU];
      T xr[RU];
#pragma unroll
      for (int u = 0; u < RU; ++u) {
        av[u] = ld_stream<T>(Ac + (r + u * NW) * lda);
        xr[u] = x[(r + u * NW) * incx];
      }
#pragma unroll
      for (int u = 0; u < RU; ++u)
#pragma unroll
        for (int k = 0; k < VN; ++k) acc[k] += vget<T>(av[u], k) * xr[u];
    }
    for (; r < r_end; r += NW) {
      const V av = ld_stream<T>(Ac + r * lda);
      const T xr = x[r * incx];
#pragma unroll
      for (int k = 0; k < VN; ++k) acc[k] += vget<T>(av, k) * xr;
    }
  } else {
    for (; r + (RU - 1) * NW < r_end; r += RU * NW) {
#pragma unroll
      for (int u = 0; u < RU; ++u) body(r + u * NW);
    }
    for (; r < r_end; r += NW) body(r);
  }

#pragma unroll
  for (int k = 0; k < VN; ++k) red[wave][lane * VN + k] = acc[k];
  if (threadIdx.x == 0) s_marker = mk;
  __syncthreads();
  auto write_y = [&](int64_t c, T s) {
    if (TRI != kFull)
      y[c * incy] = UNIT ? s + x[c * incx] : s;
    else
      y[c * incy] = beta * (ow ? T(0) : y[c * incy]) + alpha * s;
  };
  for (int i = threadIdx.x; i < kWave * VN; i += NTT) {
    T s = T(0);
#pragma unroll
    for (int w = 0; w < NW; ++w) s += red[w][i];
    const int64_t c = cb0 + i;
    if (c < Mc) {
      if (gridDim.y > 1)
        tagged::publish<T>(pub, (int64_t)blockIdx.y * Mc + c, s, s_marker);
      else
        write_y(c, s);
    }
  }
  if (gridDim.y == 1) return;
  __shared__ int s_last;
  __syncthreads();
  if (threadIdx.x == 0) {
    // thread 0 published column cb0 with the marker it read: its ticket
    // follows that read.  The last column block to complete advances the
    // generation (every workgroup has read it by then).
    s_last = tagged::last_ticket(tickets + blockIdx.x, gridDim.y);
    if (s_last && tagged::last_ticket(lt, gridDim.x)) tagged::advance_generation(gen);
  }
  __syncthreads();
  if (!s_last) return;
  const int splits = (int)gridDim.y;
  for (int i = threadIdx.x; i < kWave * VN; i += NTT) {
    const int64_t c = cb0 + i;
    if (c >= Mc) break;
    T t = T(0);
    for (int s0 = 0; s0 < splits; s0 += 8) {
      T v[8];
      tagged::read_n<T, 8>(pub, (int64_t)s0 * Mc + c, Mc, splits - s0, v, s_marker, err);
#pragma unroll
      for (int k = 0; k < 8; ++k)
        if (s0 + k < splits) t += v[k];
    }
    write_y(c, t);
  }
}


template <class T, int TRI, bool UNIT>
void launch_gemv_n(Ctx& c, int64_t M, int64_t N, T alpha, T beta, const T* A, int64_t lda, const T* x,
                   int64_t incx, T* y, int64_t incy, bool overwrite = false) {
  const int ow = overwrite ? 1 : 0;
  constexpr int VN = Vec16<T>::n;
  constexpr int P = 2048;               // x panel staged in LDS
  constexpr int64_t SG = 2048;          // column-split granularity
  const bool vec = (lda % VN) == 0 && aligned16(A);
  const int64_t want = (int64_t)c.num_cus * 4;
  // Tall matrices: 2 rows per wave over the whole row, one launch.  Shorter
  // ones: 8 rows per wave and the columns split across workgroups, summed by
  // a second launch.
  // Rows of up to 16 vectors per thread and enough of them for 8 workgroups
  // per CU: whole-workgroup rows with x in registers, two rows at a time, 16
  // workgroups per CU for FP32 and 12 for FP64 (best of sweeps over 1-4 rows
  // and 4-24 workgroups per CU on MI355X at 8192 x 8192: FP32 80.5 % of HBM
  // peak at 16 against 77.7 % at 12 and 73.0 % at 4; FP64 82.8 % at 12,
  // within 0.3 % of the best).
  if (TRI == kFull && vec && incx == 1 && aligned16(x) && ceil_div(N, VN) <= (int64_t)kNT * 16 &&
      M >= 2 * 8 * (int64_t)c.num_cus) {
    constexpr int R = 2;
    constexpr int64_t kPerCu = sizeof(T) == 4 ? 16 : 12;
    const unsigned grid = (unsigned)std::min<int64_t>(ceil_div(M, R), kPerCu * c.num_cus);
    if (ceil_div(N, VN) <= (int64_t)kNT * 8)
      hipLaunchKernelGGL((gemv_n_wg_kernel<T, 8, R, TRI, UNIT>), dim3(grid), dim3(kNT), 0, c.stream, M, N, alpha,
                         beta, A, lda, x, y, incy, ow);
    else
      hipLaunchKernelGGL((gemv_n_wg_kernel<T, 16, R, TRI, UNIT>), dim3(grid), dim3(kNT), 0, c.stream, M, N, alpha,
                         beta, A, lda, x, y, incy, ow);
    EXO_LAUNCH_CHECK();
    return;
  }
  if (vec && incx == 1 && aligned16(x) && ceil_div(M, (int64_t)(kNT / kWave)) >= want) {
    hipLaunchKernelGGL((gemv_n_row_kernel<T, TRI, UNIT>), dim3((unsigned)ceil_div(M, kNT / kWave)), dim3(kNT), 0,
                       c.stream, M, N, alpha, beta, A, lda, x, y, incy, ow);
    EXO_LAUNCH_CHECK();
    return;
  }
  const bool tall = ceil_div(M, (int64_t)(kNT / kWave) * 2) >= want;
  const int R = tall ? 2 : 8;
  const int64_t row_blocks = ceil_div(M, (int64_t)(kNT / kWave) * R);
  int64_t splits = 1;
  if (!tall && row_blocks < want && N > 2 * SG)
    splits = std::min<int64_t>(ceil_div(want, row_blocks), ceil_div(N, SG));
  int64_t cols_per = N > 0 ? ceil_div(ceil_div(N, splits), SG) * SG : SG;
  splits = N > 0 ? ceil_div(N, cols_per) : 1;
  T* part = splits > 1 ? static_cast<T*>(c.scratch(sizeof(T) * splits * M)) : nullptr;
  dim3 grid((unsigned)row_blocks, (unsigned)splits);
#define EXO_GEMV_N(RR, V)                                                                                   \
  hipLaunchKernelGGL((gemv_n_kernel<T, RR, P, V, TRI, UNIT>), grid, dim3(kNT), 0, c.stream, M, N, alpha, beta, \
                     A, lda, x, incx, y, incy, cols_per, part, ow)
  if (tall) {
    if (vec) EXO_GEMV_N(2, true); else EXO_GEMV_N(2, false);
  } else {
    if (vec) EXO_GEMV_N(8, true); else EXO_GEMV_N(8, false);
  }
#undef EXO_GEMV_N
  EXO_LAUNCH_CHECK();
  if (splits > 1) {
    hipLaunchKernelGGL((gemv_finish_kernel<T, TRI != kFull, UNIT>), dim3((unsigned)ceil_div(M, kWave)), dim3(kNT), 0,
                       c.stream, M, (int)splits, alpha, beta, part, x, incx, y, incy, ow);
    EXO_LAUNCH_CHECK();
  }
}

template <class T, int TRI, bool UNIT>
void launch_gemv_t(Ctx& c, int64_t Mc, int64_t Nr, T alpha, T beta, const T* A, int64_t lda, const T* x,
                   int64_t incx, T* y, int64_t incy, bool overwrite = false) {
  const int ow = overwrite ? 1 : 0;
  constexpr int VN = Vec16<T>::n;
  constexpr int RU = 16;
  const bool vec = (lda % VN) == 0 && aligned16(A);
  const int64_t col_blocks = ceil_div(Mc, (int64_t)kWave * VN);
  // About one workgroup per CU, 16 rows of streaming loads in flight per
  // wave: long row ranges per workgroup and few partial rows for the second
  // launch (the best of a sweep over 0.5-2 workgroups per CU and 4-32 rows
  // in flight on MI355X: 72 % of HBM peak for 8192 x 8192 FP32, 81 % FP64).
  const int64_t want = (int64_t)c.num_cus;
  int64_t splits = std::max<int64_t>(1, std::min<int64_t>(ceil_div(want, col_blocks), ceil_div(Nr, 64)));
  const int64_t rows_per = Nr > 0 ? ceil_div(ceil_div(Nr, splits), 32) * 32 : 32;
  splits = std::max<int64_t>(1, ceil_div(Nr, rows_per));
  // splits > 1 only when col_blocks < want <= Ctx::kLaunchTicket: one ticket per
  // column block from the stream's counters
  unsigned epoch = 0;  // (not used: the marker comes from the launch generation)
  unsigned *tickets = nullptr, *gen = nullptr, *lt = nullptr, *err = nullptr, *pub = nullptr;
  if (splits > 1) {
    if (col_blocks > Ctx::kLaunchTicket) throw Error(kInvalidSize, "gemv: too many column blocks for split rows");
    tickets = c.counters();
    gen = c.counters() + Ctx::kGeneration;
    lt = c.counters() + Ctx::kLaunchTicket;
    unsigned* flags = c.epoch_flags((size_t)(2 + tagged::words<T>(splits * Mc)), &epoch);
    err = c.timeout_word();
    pub = flags + 2;  // 8-byte aligned
  }
  dim3 grid((unsigned)col_blocks, (unsigned)splits);
  // Full FP32 matrices: 16 waves along the rows with 8 rows in flight each
  // (73 % of HBM peak at 8192 x 8192 against 70 % with 4 waves of 16 rows);
  // FP64 keeps 4 waves of 16 rows (81 % against 75 %).
  if (sizeof(T) == 4 && vec && TRI == kFull)
    hipLaunchKernelGGL((gemv_t_kernel<T, 8, true, TRI, UNIT, 1024>), grid, dim3(1024), 0, c.stream, Mc, Nr, alpha,
                       beta, A, lda, x, incx, y, incy, rows_per, tickets, gen, lt, err, pub, ow);
  else if (vec)
    hipLaunchKernelGGL((gemv_t_kernel<T, RU, true, TRI, UNIT>), grid, dim3(kNT), 0, c.stream, Mc, Nr, alpha, beta,
                       A, lda, x, incx, y, incy, rows_per, tickets, gen, lt, err, pub, ow);
  else
    hipLaunchKernelGGL((gemv_t_kernel<T, RU, false, TRI, UNIT>), grid, dim3(kNT), 0, c.stream, Mc, Nr, alpha,
                       beta, A, lda, x, incx, y, incy, rows_per, tickets, gen, lt, err, pub, ow);
  EXO_LAUNCH_CHECK();
}

// out[i] = x[i * incx]: a strided or misaligned x made contiguous.
template <class T>
__global__ __launch_bounds__(kNT) void gather_kernel(int64_t n, const T* __restrict__ x, int64_t incx,
                                                     T* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * kNT + threadIdx.x;
  if (i < n) out[i] = x[i * incx];
}

// gemv-N over a row-major M x N matrix: a strided or misaligned x (N
// elements) is first gathered into scratch slot 1, so the row kernels that
// read x as 16-byte vectors apply (sgemv 8192^2 with incx = 2, incy = -3:
// 60.0 -> 49.4 us on MI355X, against the LDS-staged kernel that reads x
// element by element).
template <class T>
void launch_gemv_n_any_x(Ctx& c, int64_t M, int64_t N, T alpha, T beta, const T* A, int64_t lda, const T* x,
                         int64_t incx, T* y, int64_t incy, bool overwrite = false) {
  constexpr int VN = Vec16<T>::n;
  const bool fast_a = (lda % VN) == 0 && aligned16(A);
  if (fast_a && N > 0 && M > 0 && (incx != 1 || !aligned16(x))) {
    T* xs = static_cast<T*>(c.scratch(sizeof(T) * (size_t)N, 1));
    hipLaunchKernelGGL((gather_kernel<T>), dim3((unsigned)ceil_div(N, kNT)), dim3(kNT), 0, c.stream, N, x, incx, xs);
    EXO_LAUNCH_CHECK();
    x = xs;
    incx = 1;
  }
  launch_gemv_n<T, kFull, false>(c, M, N, alpha, beta, A, lda, x, incx, y, incy, overwrite);
}

// ------------------------------------------------------------ rank updates

// A[i,j] += (alpha u1[i]) c1[j] (+ (alpha u2[i]) c2[j] for rank 2) over the
// TRI part: the streaming tile update of tiles.hpp.
template <class T, int TRI, bool TWO>
void launch_rank_update(Ctx& c, int64_t m, int64_t n, T alpha, const T* u1, int64_t iu1, const T* c1,
                        int64_t ic1, const T* u2, int64_t iu2, const T* c2, int64_t ic2, T* A, int64_t lda) {
  const tiles::UpdateArgs<T> g{alpha, u1, iu1, c1, ic1, u2, iu2, c2, ic2};
  tiles::launch_update<T, TWO ? 2 : 1, TRI>(c.stream, m, n, A, lda, g);
}

// ------------------------------------------------------------ symv

}  // namespace

namespace ops {

// src/level2/gemv.py:10-22
template <class T>
void gemv_rm(Ctx& c, int64_t TransA, int64_t M, int64_t N, T alpha, T beta, const T* A, int64_t lda,
             const T* AT, int64_t ldat, const T* x, int64_t incx, T* y, int64_t incy) {
  if (M <= 0) return;
  if (TransA == kNoTrans)
    launch_gemv_n_any_x<T>(c, M, N, alpha, beta, A, lda, x, incx, y, incy);
  else
    launch_gemv_t<T, kFull, false>(c, M, N, alpha, beta, AT, ldat, x, incx, y, incy);
}

// CBLAS-shaped gemv: y = alpha op(A) x + beta y for an M x N matrix A in
// row-major (Order 101) or column-major (102) storage, y not read when beta
// == 0.  A column-major A is the row-major N x M array S = A^T, so A x is the
// transposed product of S and A^T x the plain one.
template <class T>
void gemv_ex_impl(Ctx& c, int64_t Order, int64_t TransA, int64_t M, int64_t N, T alpha, const T* A, int64_t lda,
                  const T* x, int64_t incx, T beta, T* y, int64_t incy) {
  const bool trans = TransA != kNoTrans, row = Order == 101;
  const int64_t ny = trans ? N : M;
  if (ny <= 0) return;
  const bool ow = beta == T(0);
  // y runs over the stored rows (row-major NoTrans, column-major Trans) or
  // over the stored columns (the other two)
  const int64_t srows = row ? M : N, scols = row ? N : M;
  if (row == trans)
    launch_gemv_t<T, kFull, false>(c, scols, srows, alpha, beta, A, lda, x, incx, y, incy, ow);
  else
    launch_gemv_n_any_x<T>(c, srows, scols, alpha, beta, A, lda, x, incx, y, incy, ow);
}

// src/level2/ger.py:9-15: A[i,j] += (alpha*x[i]) * y[j]
template <class T>
void ger_rm(Ctx& c, int64_t m, int64_t n, T alpha, const T* x, int64_t incx, const T* y, int64_t incy, T* A,
            int64_t lda) {
  launch_rank_update<T, kFull, false>(c, m, n, alpha, x, incx, y, incy, nullptr, 0, nullptr, 0, A, lda);
}

// src/level2/syr.py:10-19
template <class T>
void syr_rm(Ctx& c, int64_t Uplo, int64_t n, T alpha, const T* x, int64_t incx, const T* x_copy,
            int64_t incxc, T* A, int64_t lda) {
  if (Uplo == kUpper)
    launch_rank_update<T, kUpperTri, false>(c, n, n, alpha, x, incx, x_copy, incxc, nullptr, 0, nullptr, 0, A,
                                            lda);
  else
    launch_rank_update<T, kLowerTri, false>(c, n, n, alpha, x, incx, x_copy, incxc, nullptr, 0, nullptr, 0, A,
                                            lda);
}

// src/level2/syr2.py:9-27: A[i,j] += (alpha*x[i])*y[j] + (alpha*y[i])*x[j]
template <class T>
void syr2_rm(Ctx& c, int64_t Uplo, int64_t n, T alpha, const T* x, int64_t incx, const T* x_copy,
             int64_t incxc, const T* y, int64_t incy, const T* y_copy, int64_t incyc, T* A, int64_t lda) {
  if (Uplo == kUpper)
    launch_rank_update<T, kUpperTri, true>(c, n, n, alpha, x, incx, y, incy, y_copy, incyc, x_copy, incxc, A,
                                           lda);
  else
    launch_rank_update<T, kLowerTri, true>(c, n, n, alpha, x, incx, y, incy, y_copy, incyc, x_copy, incxc, A,
                                           lda);
}

// src/level2/symv.py:10-35
template <class T>
void symv_rm(Ctx& c, int64_t Uplo, int64_t n, T alpha, const T* A, int64_t lda, const T* x, int64_t incx,
             T beta, T* y, int64_t incy) {
  if (n <= 0) return;
  constexpr int VN = Vec16<T>::n;
  const bool upper = Uplo == kUpper;
  const tiles::TileList tl =
      tiles::make_list<symv::kRB, kWave * VN * symv::kQ>(upper ? tiles::kUpper : tiles::kLower, n, n);
  const int64_t npad = tl.nb * symv::kRB;
  T* part = static_cast<T*>(c.scratch(sizeof(T) * (size_t)((tl.nc + tl.nb) * npad)));
  T* rowpart = part;
  T* colpart = part + tl.nc * npad;
  const int64_t ntiles = tl.total();
  if (ntiles > 0x7fffffff) throw Error(kInvalidSize, "symv: n too large");
  const bool vec = (lda % VN) == 0 && (n % VN) == 0 && aligned16(A);
#define EXO_SYMV(U, V)                                                                                         \
  hipLaunchKernelGGL((symv::tile_kernel<T, U, V, symv::kRB>), dim3((unsigned)ntiles), dim3(kNT), 0, c.stream, n, A, lda, x, \
                     incx, rowpart, colpart, tl)
  if (upper) {
    if (vec) EXO_SYMV(true, true); else EXO_SYMV(true, false);
  } else {
    if (vec) EXO_SYMV(false, true); else EXO_SYMV(false, false);
  }
#undef EXO_SYMV
  EXO_LAUNCH_CHECK();
  const dim3 fgrid((unsigned)ceil_div(n, kWave)), fblock(symv::kFW * kWave);
  if (upper)
    hipLaunchKernelGGL((symv::finish_kernel<T, true, symv::kRB>), fgrid, fblock, 0, c.stream, n, alpha, beta, rowpart,
                       colpart, tl, y, incy);
  else
    hipLaunchKernelGGL((symv::finish_kernel<T, false, symv::kRB>), fgrid, fblock, 0, c.stream, n, alpha, beta, rowpart,
                       colpart, tl, y, incy);
  EXO_LAUNCH_CHECK();
}

// src/level2/trmv.py:9-43.  The product reads a copy of x (the spec's xCopy
// buffer plays the same role) and overwrites x.  For a unit diagonal the
// diagonal element is not read; x[i] is added instead.
template <class T>
void trmv_rm(Ctx& c, int64_t Uplo, int64_t TransA, int64_t Diag, int64_t n, T* x, int64_t incx, const T* A,
             int64_t lda) {
  if (n <= 0) return;
  // Contiguous copy of x in scratch slot 1 (gemv partials use slot 0).
  T* xs = static_cast<T*>(c.scratch(sizeof(T) * n, 1));
  ops::copy<T>(c, n, x, incx, xs, 1);
  const int64_t xinc = 1;
  const bool up = Uplo == kUpper, tr = TransA != kNoTrans, unit = Diag == kUnit;
  // Each variant: x = op(tri(A)) * xs.
  auto run = [&](auto tri_tag, auto unit_tag) {
    constexpr int TRI = decltype(tri_tag)::value;
    constexpr bool UNIT = decltype(unit_tag)::value;
    if (!tr)
      launch_gemv_n<T, TRI, UNIT>(c, n, n, T(1), T(0), A, lda, xs, xinc, x, incx);
    else
      launch_gemv_t<T, TRI, UNIT>(c, n, n, T(1), T(0), A, lda, xs, xinc, x, incx);
  };
  using L = std::integral_constant<int, kLowerTri>;
  using U = std::integral_constant<int, kUpperTri>;
  using Y = std::integral_constant<bool, true>;
  using N = std::integral_constant<bool, false>;
  if (up) {
    if (unit) run(U{}, Y{}); else run(U{}, N{});
  } else {
    if (unit) run(L{}, Y{}); else run(L{}, N{});
  }
}

#define EXO_L2_INST(T)                                                                                       \
  template void gemv_rm<T>(Ctx&, int64_t, int64_t, int64_t, T, T, const T*, int64_t, const T*, int64_t,     \
                           const T*, int64_t, T*, int64_t);                                                 \
  template void ger_rm<T>(Ctx&, int64_t, int64_t, T, const T*, int64_t, const T*, int64_t, T*, int64_t);     \
  template void symv_rm<T>(Ctx&, int64_t, int64_t, T, const T*, int64_t, const T*, int64_t, T, T*, int64_t); \
  template void syr_rm<T>(Ctx&, int64_t, int64_t, T, const T*, int64_t, const T*, int64_t, T*, int64_t);     \
  template void syr2_rm<T>(Ctx&, int64_t, int64_t, T, const T*, int64_t, const T*, int64_t, const T*,        \
                           int64_t, const T*, int64_t, T*, int64_t);                                        \
  template void trmv_rm<T>(Ctx&, int64_t, int64_t, int64_t, int64_t, T*, int64_t, const T*, int64_t);

EXO_L2_INST(float)
EXO_L2_INST(double)

void gemv_ex(Ctx& c, int64_t Order, int64_t TransA, int64_t M, int64_t N, float alpha, const float* A, int64_t lda,
             const float* x, int64_t incx, float beta, float* y, int64_t incy) {
  gemv_ex_impl<float>(c, Order, TransA, M, N, alpha, A, lda, x, incx, beta, y, incy);
}
void gemv_ex(Ctx& c, int64_t Order, int64_t TransA, int64_t M, int64_t N, double alpha, const double* A,
             int64_t lda, const double* x, int64_t incx, double beta, double* y, int64_t incy) {
  gemv_ex_impl<double>(c, Order, TransA, M, N, alpha, A, lda, x, incx, beta, y, incy);
}

}  // namespace ops
}  // namespace exoblas
