// Mixed-precision strsv with iterative refinement.
//
// update = 32 (FP32-stored updates):
//   refine = 1: one launch (trsv_refined_kernel): the FP32 look-back solve
//   and, trailing it by a few microseconds, the correction solve
//   d = solve(b - op(A) x) with an FP64 residual, sharing each block's
//   inverse and folds; x + d added in FP64 and rounded once (8192: 165 us
//   against 115 us for the plain solve and 275 us for the three-launch
//   form below).
//   refine > 1: classical refinement around the plain FP32 solve,
//     x_hi = solve(b)                 the FP32 look-back solve (trsv.hip)
//     repeat `refine` times:
//       r = b - op(A) (x_hi + x_lo)   FP64 accumulation over the FP32 matrix,
//                                     b read from the caller's vector, r
//                                     stored in FP32 as the next right-hand side
//       d = solve(r);  x_lo += d      (x_lo kept in FP32: it is ~eps |x|)
//     x = x_hi + x_lo                 added in FP64, rounded once to FP32
//   The solution is carried as the FP32 pair (x_hi, x_lo), so no FP64
//   vector is stored.
//
// update = 16 (BF16-stored updates):
//   solve   op(A) x0 = b   off-diagonal block updates in FP32 arithmetic on a
//                          BF16 copy of A, 64 x 64 diagonal blocks solved in
//                          FP64
//   repeat `refine` times:
//           r = b - op(A) x   FP64 accumulation over the FP32 matrix
//           solve op(A) d = r (as above),  x += d  (FP64)
//   x (FP32 out) = x
// That solve is the look-back block solve of trsv_core.hpp with a different
// element pipeline: right-hand side and solution in FP64, an FP32 shadow of
// the solution feeding the off-diagonal updates.
#include "device.hpp"
#include "ops.hpp"
#include "trsv_core.hpp"

namespace exoblas {
namespace {

using trsv::kB;
using trsv::kNT;
constexpr int kFlagBase = 2;

template <class TO>
__device__ __forceinline__ float to_f(TO v) { return (float)v; }

// One step of the mixed solve: the trsv_core step with FP32 off-diagonal
// updates (on TO-typed copies of A), the diagonal block inverted and applied
// in FP64, an FP64 right-hand side and solution, and an FP32 shadow of the
// solution feeding later off-diagonal updates.
template <class TO, bool FWD, bool TRANS, bool UNIT>
__global__ __launch_bounds__(kNT) void trsv_mixed_kernel(int64_t n, const float* __restrict__ A, int64_t lda,
                                                         const TO* __restrict__ Aoff, int64_t ldo,
                                                         const double* __restrict__ rhs, double* x, float* xf,
                                                         unsigned* flags, unsigned* gen, unsigned* err) {
  __shared__ trsv::Smem<float, double> sm;
  __shared__ int s_p;
  const int nb = (int)((n + kB - 1) / kB);
  const int t = threadIdx.x;
  __shared__ unsigned s_marker;
  if (t == 0) {  // ticket and marker from the stream's launch generation (graph-replayable)
    unsigned marker;
    s_p = (int)tagged::generation_ticket(flags, (unsigned)nb, gen, &marker);
    s_marker = marker;
  }
  __syncthreads();
  if (s_p >= nb) return;  // only if launches sharing the stream's ticket overlapped
  const int p = s_p;
  const unsigned epoch = s_marker;
  unsigned* done = flags + kFlagBase;
  const int64_t bi = FWD ? p : nb - 1 - p;
  const int64_t r0 = bi * kB;
  trsv::invert_diag<float, double, FWD, TRANS, UNIT>(n, A, lda, r0, sm.inv, (int)threadIdx.x);
  const int r = TRANS ? t % kB : t / 4;
  const int cq = TRANS ? t / kB : t % 4;
  const int64_t gi = r0 + r;
  // Off-diagonal entries of block q for this thread, one bounds test per run;
  // each block is loaded one iteration ahead of its x block's wait.
  auto load_m = [&](int q, float (&m)[16]) {
    const int64_t bj = FWD ? q : nb - 1 - q;
    const int64_t c0 = bj * kB + cq * 16;
    const bool in = gi < n && c0 + 16 <= n;
#pragma unroll
    for (int e = 0; e < 16; ++e) {
      const int64_t gj = c0 + e;
      m[e] = (in || (gi < n && gj < n)) ? to_f(TRANS ? Aoff[gj * ldo + gi] : Aoff[gi * ldo + gj]) : 0.f;
    }
  };
  float acc = 0.f;
  float m[2][16];
  if (p > 0) load_m(0, m[0]);
  for (int q = 0; q < p; q += 2) {
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int qq = q + h;
      if (qq < p) {
        if (qq + 1 < p) load_m(qq + 1, m[h ^ 1]);
        float* xs = sm.xs[h];
        trsv::fetch_block<float>(n, FWD ? qq : nb - 1 - qq, done, epoch, err, xs);
        __syncthreads();
#pragma unroll
        for (int e = 0; e < 16; ++e) acc += m[h][e] * xs[cq * 16 + e];
      }
    }
  }
  sm.red[cq][r] = acc;
  __syncthreads();
  if (t < kB) {
    const int64_t g = trsv::row_of<FWD>(r0, t);
    double v = 0.0;
    if (g < n) {
      float s = 0.f;
#pragma unroll
      for (int w = 0; w < kNT / kB; ++w) s += sm.red[w][g - r0];
      v = rhs[g] - (double)s;
    }
    sm.vs[t] = v;
  }
  __syncthreads();
  {
    const int i = t / 4, q4 = t % 4;
    double s = 0.0;
#pragma unroll
    for (int e = 0; e < 16; ++e) s += sm.inv[i][q4 * 16 + e] * sm.vs[q4 * 16 + e];
    s += __shfl_xor(s, 1, 64);
    s += __shfl_xor(s, 2, 64);
    const int64_t g = trsv::row_of<FWD>(r0, i);
    if (q4 == 0 && g < n) {
      x[g] = s;
      trsv::publish_value<float>(g, (float)s, xf, done, epoch);
    }
  }
}

// r[i] = b[i] - sum_j M(i, j) x[j] over the triangle (FP64 accumulation), one
// wave per row; NoTrans rows are contiguous, Trans rows are columns of A.
template <bool LOWER, bool TRANS, bool UNIT>
__global__ __launch_bounds__(256) void residual_kernel(int64_t n, const float* __restrict__ A, int64_t lda,
                                                       const double* __restrict__ b, const double* __restrict__ x,
                                                       double* r) {
  const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
  const int64_t i = (int64_t)blockIdx.x * 4 + wave;
  if (i >= n) return;
  // M = op(tri(A)): lower-notrans / upper-trans touch j <= i, the others j >= i
  const bool left = LOWER != TRANS;
  const int64_t jl = left ? 0 : i, jh = left ? i + 1 : n;
  double s = 0.0;
  for (int64_t j = jl + lane; j < jh; j += 64) {
    double m = (double)(TRANS ? A[j * lda + i] : A[i * lda + j]);
    if (UNIT && j == i) m = 1.0;
    s += m * x[j];
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
  if (lane == 0) r[i] = b[i] - s;
}

// r[i] = (float)(b[i] - sum_j M(i, j) (xh[j] + xl[j])), FP64 accumulation,
// b read at increment incb; xl may be null (zero).  NoTrans: one wave per
// row (rows contiguous, 16-byte loads when aligned).
template <bool LOWER, bool UNIT, bool VEC>
__global__ __launch_bounds__(256) void residual_rows_kernel(int64_t n, const float* __restrict__ A, int64_t lda,
                                                            const float* __restrict__ b, int64_t incb,
                                                            const float* __restrict__ xh, const float* __restrict__ xl,
                                                            float* r) {
  const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
  const int64_t i = (int64_t)blockIdx.x * 4 + wave;
  if (i >= n) return;
  const int64_t jl = LOWER ? 0 : i, jh = LOWER ? i + 1 : n;
  const float* Ar = A + i * lda;
  auto xv = [&](int64_t j) { return (double)xh[j] + (xl ? (double)xl[j] : 0.0); };
  double s = 0.0;
  int64_t j0 = jl;
  if (VEC) {  // the 16-byte aligned interior, 4 columns per lane per load
    const int64_t vl = (jl + 3) / 4, vh = jh / 4;
    for (int64_t j = jl + lane; j < std::min<int64_t>(vl * 4, jh); j += 64) s += (double)Ar[j] * xv(j);
    for (int64_t v = vl + lane; v < vh; v += 64) {
      const float4 a = reinterpret_cast<const float4*>(Ar)[v];
      const int64_t j = v * 4;
      s += (double)a.x * xv(j) + (double)a.y * xv(j + 1) + (double)a.z * xv(j + 2) + (double)a.w * xv(j + 3);
    }
    j0 = std::max<int64_t>(vh * 4, std::min<int64_t>(vl * 4, jh));
  }
  for (int64_t j = j0 + lane; j < jh; j += 64) s += (double)Ar[j] * xv(j);
  if (UNIT) {  // the diagonal is 1, whatever A holds there: swap A[i, i] x_i for x_i
    if (lane == 0) s += xv(i) * (1.0 - (double)Ar[i]);
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
  if (lane == 0) r[i] = (float)((double)b[i * incb] - s);
}

// The same for TransA: r[i] involves column i of A.  A workgroup of 8 waves
// owns 64 consecutive columns; wave w takes rows w, w + 8, ... so a wave's
// load is 64 consecutive floats of one row; FP64 column sums reduced through
// LDS in wave order.
template <bool LOWER, bool UNIT>
__global__ __launch_bounds__(512) void residual_cols_kernel(int64_t n, const float* __restrict__ A, int64_t lda,
                                                            const float* __restrict__ b, int64_t incb,
                                                            const float* __restrict__ xh, const float* __restrict__ xl,
                                                            float* r) {
  __shared__ double red[8][64];
  const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
  const int64_t i = (int64_t)blockIdx.x * 64 + lane;  // this lane's column
  const int64_t c0 = (int64_t)blockIdx.x * 64;
  // M(i, j) = A[j, i]: lower A, Trans -> j >= i; upper A, Trans -> j <= i
  const int64_t jl = LOWER ? c0 : 0, jh = LOWER ? n : std::min<int64_t>(n, c0 + 64);
  double s = 0.0;
  if (i < n) {
    for (int64_t j = jl + wave; j < jh; j += 8) {
      const bool in = LOWER ? j >= i : j <= i;
      if (in) s += (double)A[j * lda + i] * ((double)xh[j] + (xl ? (double)xl[j] : 0.0));
    }
    if (UNIT && i >= jl && i < jh && (i - jl) % 8 == wave)
      s += ((double)xh[i] + (xl ? (double)xl[i] : 0.0)) * (1.0 - (double)A[i * lda + i]);
  }
  red[wave][lane] = s;
  __syncthreads();
  if (wave == 0 && i < n) {
    double t = red[0][lane];
#pragma unroll
    for (int w = 1; w < 8; ++w) t += red[w][lane];
    r[i] = (float)((double)b[i * incb] - t);
  }
}

// x[i * incx] = (float)((double)xh[i] + (double)xl[i]); xf[i] = x[i * incx]
// (gather) when xl is null and xh is the destination of a copy.
__global__ void combine_kernel(int64_t n, const float* xh, const float* xl, float* x, int64_t incx) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i < n) x[i * incx] = (float)((double)xh[i] + (double)xl[i]);
}
__global__ void gather_f_kernel(int64_t n, const float* x, int64_t incx, float* out) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i < n) out[i] = x[i * incx];
}
__global__ void acc_f_kernel(int64_t n, float* xl, const float* d, int first) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i < n) xl[i] = first ? d[i] : (float)((double)xl[i] + (double)d[i]);
}

__global__ void widen_kernel(int64_t n, const float* x, int64_t incx, double* y) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i < n) y[i] = (double)x[i * incx];
}
__global__ void add_kernel(int64_t n, double* x, const double* d) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i < n) x[i] += d[i];
}
__global__ void narrow_kernel(int64_t n, const double* x, float* y, int64_t incy) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i < n) y[i * incy] = (float)x[i];
}
__global__ void narrow_f_copy_kernel(int64_t n, const float* xh, float* x, int64_t incx) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i < n) x[i * incx] = xh[i];
}
__global__ void to_bf16_kernel(int64_t rows, int64_t cols, const float* A, int64_t lda, __bf16* out) {
  const int64_t i = blockIdx.y;
  for (int64_t j = (int64_t)blockIdx.x * 256 + threadIdx.x; j < cols; j += (int64_t)gridDim.x * 256)
    out[i * cols + j] = (__bf16)A[i * lda + j];
}

// One refinement fused into the solve (FP32, update = 32, refine = 1): the
// look-back solve of trsv_core.hpp (64-row blocks, G1 = inv(D) M(p, p-1) and
// G2 = inv(D) M(p, p-2) folded ahead of time) runs two chains in one
// launch.  Block p first publishes x_p as the plain solve does; then forms
// its residual rows r_p = b_p - sum_{j <= p} M(p, j) x_j in FP64 (the panel
// part accumulated in the same loop that consumes x_j, from the same loaded
// entries; the last three blocks from raw entries loaded after the panel),
// and solves the correction chain d_p = inv(D)(r_p - sum_{j <= p-3} M(p, j)
// d_j) - G2 d_{p-2} - G1 d_{p-1} with the same inverse and folds; the panel
// loop consumes d_{q-2} beside x_q (each panel block loaded once: a ring of
// the last four blocks' entries), so the correction chain trails the
// solve by about a step and the refined solve costs about one plain solve.
// The result x = x_p + d_p is added in FP64 and rounded once.  x_j and d_j
// are published as tagged words in two areas of the stream's flag space.
template <bool FWD, bool TRANS, bool UNIT>
__global__ __launch_bounds__(kNT) void trsv_refined_kernel(int64_t n, const float* __restrict__ A, int64_t lda,
                                                           float* x, int64_t incx, unsigned* flags, unsigned* gen,
                                                           unsigned* err) {
  using trsv::VecN;
  constexpr int KB = kB, CW = KB / 4, PT = trsv::per_thread<KB>(), VN = 4;
  constexpr int MS = trsv::mb_stride<float, KB>();
  __shared__ __align__(16) trsv::Smem<float> sm;
  __shared__ __align__(16) float mb[KB][MS];   // M(p, p-1), then G1
  __shared__ __align__(16) float mb2[KB][MS];  // M(p, p-2), then G2
  __shared__ __align__(16) float xd[2][KB];    // d blocks of the panel
  __shared__ __align__(16) float xt[4][KB];    // the panel's last d blocks
  __shared__ __align__(16) float xp[KB];       // x_p, natural order within the block
  __shared__ __align__(16) float bs[KB];       // b, natural order
  __shared__ __align__(16) float vd[KB];       // r_p - d panel, local order
  __shared__ double rred[4][KB];
  // raw M(p, p-1), M(p, p-2) (rows in local order, natural columns) and D
  // (local order), kept for the residual
  __shared__ __align__(16) float rb1[KB][MS];
  __shared__ __align__(16) float rb2[KB][MS];
  __shared__ float rd[KB][KB + 1];
  __shared__ __align__(16) float tail0[4][KB][CW];  // wave 0's entries of the panel's last blocks
  __shared__ int s_p;
  __shared__ unsigned s_marker;
  const int nb = (int)((n + KB - 1) / KB);
  const int t = threadIdx.x;
  if (t == 0) {
    unsigned marker;
    s_p = (int)tagged::generation_ticket(flags, (unsigned)nb, gen, &marker);
    s_marker = marker;
  }
  __syncthreads();
  if (s_p >= nb) return;  // only if launches sharing the stream's ticket overlapped
  const int p = s_p;
  const unsigned marker = s_marker;
  unsigned* pubx = flags + kFlagBase;
  unsigned* pubd = pubx + trsv::pub_words<float>(n);
  const int64_t bi = FWD ? p : nb - 1 - p;
  const int64_t r0 = bi * KB;
  auto blk = [&](int q) -> int64_t { return FWD ? q : nb - 1 - q; };  // natural block of step q
  // right-hand side, M(p, p-1), M(p, p-2) and the diagonal block: every load
  // issued before the first is used
  float rhs = 0.f;
  if (t < KB) {
    const int64_t g = r0 + t;  // natural order
    if (g < n) rhs = x[g * incx];
  }
  auto load_prev = [&](int d, float (&m)[PT]) {
    const int64_t cb0 = blk(p - d) * KB;
#pragma unroll
    for (int e = 0; e < PT; ++e) {
      const int idx = t + 4 * KB * e;
      const int a = idx / KB, c2 = idx % KB;
      const int k = TRANS ? c2 : a, c = TRANS ? a : c2;
      const int64_t gk = trsv::row_of<FWD, KB>(r0, k), gc = cb0 + c;
      m[e] = (gk < n && gc < n) ? trsv::mat<float, TRANS>(A, lda, gk, gc) : 0.f;
    }
  };
  auto stage_prev = [&](const float (&m)[PT], float (*dst)[MS]) {
#pragma unroll
    for (int e = 0; e < PT; ++e) {
      const int idx = t + 4 * KB * e;
      const int a = idx / KB, c2 = idx % KB;
      dst[TRANS ? c2 : a][TRANS ? a : c2] = m[e];
    }
  };
  float mv[PT], mv2[PT], dv[PT];
  if (p > 0) load_prev(1, mv);
  if (p > 1) load_prev(2, mv2);
  trsv::load_diag<float, float, FWD, TRANS, UNIT, KB>(n, A, lda, r0, t, dv);
  if (t < KB) {
    bs[t] = rhs;
    sm.vs[FWD ? t : KB - 1 - t] = rhs;  // local order
  }
  if (p > 0) {
    stage_prev(mv, mb);
    stage_prev(mv, rb1);
  }
  if (p > 1) {
    stage_prev(mv2, mb2);
    stage_prev(mv2, rb2);
  }
#pragma unroll
  for (int e = 0; e < PT; ++e) {  // load_diag's (i, j) for this thread's entry e
    const int idx = t + 4 * KB * e;
    const int a = idx / KB, c2 = idx % KB;
    rd[TRANS ? c2 : a][TRANS ? a : c2] = dv[e];
  }
  trsv::invert_loaded<float, TRANS, KB>(sm.inv, t, dv);
  if (p > 0) trsv::fold_g<float, KB>(sm.inv, mb, t);
  if (p > 1) trsv::fold_g<float, KB>(sm.inv, mb2, t);
  // the panel: NoTrans 4 threads per row, CW consecutive columns each; Trans
  // one column of KB rows per KB threads, CW columns per group
  const int r = TRANS ? t % KB : t / 4;
  const int cq = TRANS ? t / KB : t % 4;
  const int64_t gi = r0 + r;
  auto load_m = [&](int q, float (&m)[CW]) {
    const int64_t c0 = blk(q) * KB + cq * CW;
    if (gi < n && c0 + CW <= n) {
#pragma unroll
      for (int e = 0; e < CW; ++e) m[e] = trsv::mat<float, TRANS>(A, lda, gi, c0 + e);
    } else {
#pragma unroll
      for (int e = 0; e < CW; ++e) m[e] = (gi < n && c0 + e < n) ? trsv::mat<float, TRANS>(A, lda, gi, c0 + e) : 0.f;
    }
  };
  const int pp = p > 1 ? p - 2 : 0;  // panel blocks 0 .. p-3
  // The panel loop consumes d_{q-LAG} beside x_q (a second stream of the
  // same entries, LAG blocks behind): the correction chain trails the solve
  // by a few microseconds (a step of it starts after its block's x and
  // residual), and a lag of four blocks keeps the solve's panel from
  // waiting on it.
  // Wave 0 polls before it issues its own loads: a poll waits for every
  // load its wave has in flight.
  constexpr int LAG = 4;
  const int j0 = pp > LAG ? pp - LAG : 0;  // the panel's last LAG blocks
  float acc = 0.f, accd = 0.f;
  double accr = 0.0;
  float m[2][CW], md[2][CW];
  auto prefetch = [&](int qq, int h) {
    if (qq + 1 < pp) load_m(qq + 1, m[(h + 1) & 1]);
    if (qq + 1 >= LAG && qq + 1 < pp) load_m(qq + 1 - LAG, md[(h + 1) & 1]);
  };
  if (pp > 0) load_m(0, m[0]);
  for (int q = 0; q < pp; q += 2) {
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int qq = q + h;
      if (qq < pp) {
        if (t < KB) {  // x_qq and d_{qq-LAG}, both loads in flight together
          const int64_t gx = blk(qq) * KB + t;
          const int64_t gd = blk(qq >= LAG ? qq - LAG : 0) * KB + t;
          float xv, dv;
          tagged::read2<float>(pubx, gx, gx < n, pubd, gd, qq >= LAG && gd < n, marker, err, xv, dv);
          sm.xs[h][t] = xv;
          if (qq >= LAG) xd[h][t] = dv;
          prefetch(qq, h);
        }
        __syncthreads();
        // (waves 1-3 issue theirs after the barrier, so the barrier does not
        // wait for their issue)
        if (t >= KB) prefetch(qq, h);
#pragma unroll
        for (int e = 0; e < CW; ++e) {
          const float xv = sm.xs[h][cq * CW + e];
          acc += m[h][e] * xv;
          accr += (double)m[h][e] * (double)xv;
        }
        if (qq >= LAG) {
#pragma unroll
          for (int e = 0; e < CW; ++e) accd += md[h][e] * xd[h][cq * CW + e];
        }
        if (t < KB && qq >= j0) {
#pragma unroll
          for (int e = 0; e < CW; ++e) tail0[qq - j0][t][e] = m[h][e];
        }
      }
    }
  }
  // the entries of the panel's last LAG blocks, whose d the loop above did
  // not reach: waves 1-3 read them again during the chain, wave 0 kept
  // them in LDS (so its polls never wait behind loads of its own)
  float mtl[LAG][CW];
  auto load_tail = [&]() {
#pragma unroll
    for (int k = 0; k < LAG; ++k) {
      if (j0 + k < pp) {
        load_m(j0 + k, mtl[k]);
      } else {
#pragma unroll
        for (int e = 0; e < CW; ++e) mtl[k][e] = 0.f;
      }
    }
  };
  // v = b - panel sums (vs holds b, local order)
  if constexpr (!TRANS) {
    float ps = acc + __shfl_xor(acc, 1, 64);
    ps += __shfl_xor(ps, 2, 64);
    if (cq == 0 && gi < n) sm.vs[FWD ? r : KB - 1 - r] -= ps;
  } else {
    sm.red[cq][r] = acc;
    __syncthreads();
    if (t < KB) {
      const int64_t g = trsv::row_of<FWD, KB>(r0, t);
      if (g < n) {
        const int nr = (int)(g - r0);
        sm.vs[t] -= (sm.red[0][nr] + sm.red[1][nr]) + (sm.red[2][nr] + sm.red[3][nr]);
      }
    }
  }
  __syncthreads();
  if (t >= KB) load_tail();  // during the chain, which waves 1-3 do not take part in
  // wave 0, lane = local row t: z = inv(D) v - G2 u_{p-2} - G1 u_{p-1} for
  // u = x (first chain) or d (correction chain); v in vv (local order), the
  // two previous blocks through sm.xs[1] / sm.xs[0]
  auto chain = [&](const float* vv, unsigned* pub, const float* pre2) -> float {
    float y[4] = {0.f, 0.f, 0.f, 0.f};
    {
      const VecN<float, VN>* v4p = reinterpret_cast<const VecN<float, VN>*>(vv);
      VecN<float, VN> vr[KB / VN];
#pragma unroll
      for (int c = 0; c < KB / VN; ++c) vr[c] = v4p[c];
#pragma unroll
      for (int c = 0; c < KB / VN; ++c)
#pragma unroll
        for (int e = 0; e < VN; ++e) y[e] += sm.inv[t][c * VN + e] * vr[c].v[e];
    }
    float z = (y[0] + y[1]) + (y[2] + y[3]);
    if (p > 1) {
      const float* xs = pre2;  // the p-2 block, already in LDS, or polled now
      if (!pre2) {
        const int64_t g2 = blk(p - 2) * KB + t;
        sm.xs[1][t] = g2 < n ? tagged::read<float>(pub, g2, marker, err) : 0.f;
        __builtin_amdgcn_wave_barrier();
        xs = sm.xs[1];
      }
      const VecN<float, VN>* x4 = reinterpret_cast<const VecN<float, VN>*>(xs);
      const VecN<float, VN>* g4 = reinterpret_cast<const VecN<float, VN>*>(mb2[t]);
      float u[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int c = 0; c < KB / VN; ++c) {
        const VecN<float, VN> xx = x4[c], gg = g4[c];
#pragma unroll
        for (int e = 0; e < VN; ++e) u[e] += gg.v[e] * xx.v[e];
      }
      z -= (u[0] + u[1]) + (u[2] + u[3]);
    }
    if (p > 0) {
      float gr[KB];  // this row of G1, read before the wait
      {
        const VecN<float, VN>* g4 = reinterpret_cast<const VecN<float, VN>*>(mb[t]);
#pragma unroll
        for (int c = 0; c < KB / VN; ++c) {
          const VecN<float, VN> gg = g4[c];
#pragma unroll
          for (int e = 0; e < VN; ++e) gr[c * VN + e] = gg.v[e];
        }
      }
      const int64_t g1 = blk(p - 1) * KB + t;
      float* xs = sm.xs[0];
      xs[t] = g1 < n ? tagged::read<float>(pub, g1, marker, err) : 0.f;
      __builtin_amdgcn_wave_barrier();
      const VecN<float, VN>* x4 = reinterpret_cast<const VecN<float, VN>*>(xs);
      VecN<float, VN> xr[KB / VN];
#pragma unroll
      for (int c = 0; c < KB / VN; ++c) xr[c] = x4[c];
      float u[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int c = 0; c < KB / VN; ++c)
#pragma unroll
        for (int e = 0; e < VN; ++e) u[e] += gr[c * VN + e] * xr[c].v[e];
      z -= (u[0] + u[1]) + (u[2] + u[3]);
    }
    return z;
  };
  float xmine = 0.f;  // wave 0: x_p of local row t
  const int64_t gmine = trsv::row_of<FWD, KB>(r0, t);
  if (t < KB) {
    xmine = chain(sm.vs, pubx, nullptr);
    if (gmine < n) tagged::publish<float>(pubx, gmine, xmine, marker);
    xp[gmine - r0] = xmine;
  }
  // wave 0 issues the first polls of the tail's d blocks now, so their
  // round trip overlaps the residual (checked after it).  (d_{p-2} is
  // polled later, in the chain: waiting on the newest block before the
  // older work is done lengthens every step of the correction chain.)
  uint64_t* wd64 = reinterpret_cast<uint64_t*>(pubd);
  uint64_t wt[LAG];
  auto tail_g = [&](int k) -> int64_t { return blk(j0 + k) * KB + t; };
  auto tail_on = [&](int k) -> bool { return j0 + k < pp && tail_g(k) < n; };
  if (t < KB) {
#pragma unroll
    for (int k = 0; k < LAG; ++k)
      if (tail_on(k)) wt[k] = __hip_atomic_load(wd64 + tail_g(k), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  __syncthreads();
  // residual rows in FP64: the panel part, then the last three blocks
  {
    double s = accr;
    const int cb = cq * CW;
    // (the x buffers of blocks that do not exist are never read: they hold
    // whatever the LDS held)
    const int kl = FWD ? r : KB - 1 - r;  // local order of this thread's row
    if (p > 1) {
#pragma unroll
      for (int e = 0; e < CW; ++e) s += (double)rb2[kl][cb + e] * (double)sm.xs[1][cb + e];
    }
    if (p > 0) {
#pragma unroll
      for (int e = 0; e < CW; ++e) s += (double)rb1[kl][cb + e] * (double)sm.xs[0][cb + e];
    }
#pragma unroll
    for (int e = 0; e < CW; ++e) s += (double)rd[kl][FWD ? cb + e : KB - 1 - (cb + e)] * (double)xp[cb + e];
    if constexpr (!TRANS) {
      s += __shfl_xor(s, 1, 64);
      s += __shfl_xor(s, 2, 64);
      if (cq == 0) rred[0][r] = s;
    } else {
      rred[cq][r] = s;
    }
  }
  // the tail's d blocks into LDS
  if (t < KB) {
#pragma unroll
    for (int k = 0; k < LAG; ++k) {
      float v = 0.f;
      if (tail_on(k))
        v = __uint_as_float((wt[k] >> 32) == marker ? (unsigned)wt[k]
                                                     : tagged::poll_word(wd64 + tail_g(k), marker, err));
      xt[k][t] = v;
    }
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < LAG; ++k) {
    if (j0 + k < pp) {
#pragma unroll
      for (int e = 0; e < CW; ++e) accd += (t < KB ? tail0[k][t][e] : mtl[k][e]) * xt[k][cq * CW + e];
    }
  }
  // vd = (r_p - d panel), local order
  if constexpr (!TRANS) {
    float pd = accd + __shfl_xor(accd, 1, 64);
    pd += __shfl_xor(pd, 2, 64);
    if (cq == 0) {
      const double rv = (double)bs[r] - rred[0][r];
      vd[FWD ? r : KB - 1 - r] = gi < n ? (float)(rv - (double)pd) : 0.f;
    }
  } else {
    sm.red[cq][r] = accd;
    __syncthreads();
    if (t < KB) {
      const int nr = (int)(gmine - r0);
      const double rv = (double)bs[nr] - ((rred[0][nr] + rred[1][nr]) + (rred[2][nr] + rred[3][nr]));
      const float pd = (sm.red[0][nr] + sm.red[1][nr]) + (sm.red[2][nr] + sm.red[3][nr]);
      vd[t] = gmine < n ? (float)(rv - (double)pd) : 0.f;
    }
  }
  __syncthreads();
  if (t < KB) {
    const float d = chain(vd, pubd, nullptr);
    if (gmine < n) {
      tagged::publish<float>(pubd, gmine, d, marker);
      x[gmine * incx] = (float)((double)xmine + (double)d);
    }
  }
  __syncthreads();
}

template <class TO, bool FWD, bool TRANS>
void solve(Ctx& c, bool unit, int64_t n, const float* A, int64_t lda, const TO* Aoff, int64_t ldo,
           const double* rhs, double* x, float* xf) {
  const int64_t nb = ceil_div(n, kB);
  unsigned epoch = 0;  // (not used: the marker comes from the launch generation)
  unsigned* flags = c.epoch_flags((size_t)trsv::pub_words<float>(n) + kFlagBase, &epoch);
  unsigned* gen = c.counters() + Ctx::kGeneration;
  if (unit)
    hipLaunchKernelGGL((trsv_mixed_kernel<TO, FWD, TRANS, true>), dim3((unsigned)nb), dim3(kNT), 0, c.stream, n, A,
                       lda, Aoff, ldo, rhs, x, xf, flags, gen, c.timeout_word());
  else
    hipLaunchKernelGGL((trsv_mixed_kernel<TO, FWD, TRANS, false>), dim3((unsigned)nb), dim3(kNT), 0, c.stream, n, A,
                       lda, Aoff, ldo, rhs, x, xf, flags, gen, c.timeout_word());
  EXO_LAUNCH_CHECK();
}

template <class TO>
void solve_any(Ctx& c, bool fwd, bool trans, bool unit, int64_t n, const float* A, int64_t lda, const TO* Aoff,
               int64_t ldo, const double* rhs, double* x, float* xf) {
  if (fwd && !trans) solve<TO, true, false>(c, unit, n, A, lda, Aoff, ldo, rhs, x, xf);
  if (fwd && trans) solve<TO, true, true>(c, unit, n, A, lda, Aoff, ldo, rhs, x, xf);
  if (!fwd && !trans) solve<TO, false, false>(c, unit, n, A, lda, Aoff, ldo, rhs, x, xf);
  if (!fwd && trans) solve<TO, false, true>(c, unit, n, A, lda, Aoff, ldo, rhs, x, xf);
}

template <bool FWD, bool TRANS>
void refined_launch(Ctx& c, bool unit, int64_t n, const float* A, int64_t lda, float* x, int64_t incx) {
  const int64_t nb = ceil_div(n, kB);
  if (nb > 0x7fffffff) throw Error(kInvalidSize, "trsv: n too large");
  unsigned epoch = 0;  // (not used: the marker comes from the launch generation)
  unsigned* flags = c.epoch_flags(2 * (size_t)trsv::pub_words<float>(n) + kFlagBase, &epoch);
  unsigned* gen = c.counters() + Ctx::kGeneration;
  if (unit)
    hipLaunchKernelGGL((trsv_refined_kernel<FWD, TRANS, true>), dim3((unsigned)nb), dim3(kNT), 0, c.stream, n, A,
                       lda, x, incx, flags, gen, c.timeout_word());
  else
    hipLaunchKernelGGL((trsv_refined_kernel<FWD, TRANS, false>), dim3((unsigned)nb), dim3(kNT), 0, c.stream, n, A,
                       lda, x, incx, flags, gen, c.timeout_word());
  EXO_LAUNCH_CHECK();
}

void refined(Ctx& c, bool fwd, bool trans, bool unit, int64_t n, const float* A, int64_t lda, float* x,
             int64_t incx) {
  if (fwd && !trans) refined_launch<true, false>(c, unit, n, A, lda, x, incx);
  if (fwd && trans) refined_launch<true, true>(c, unit, n, A, lda, x, incx);
  if (!fwd && !trans) refined_launch<false, false>(c, unit, n, A, lda, x, incx);
  if (!fwd && trans) refined_launch<false, true>(c, unit, n, A, lda, x, incx);
}

void residual(Ctx& c, bool lower, bool trans, bool unit, int64_t n, const float* A, int64_t lda, const double* b,
              const double* x, double* r) {
  const dim3 g((unsigned)ceil_div(n, 4));
#define EXO_RES(L, T, U) \
  hipLaunchKernelGGL((residual_kernel<L, T, U>), g, dim3(256), 0, c.stream, n, A, lda, b, x, r)
  if (lower && !trans && unit) EXO_RES(true, false, true);
  if (lower && !trans && !unit) EXO_RES(true, false, false);
  if (lower && trans && unit) EXO_RES(true, true, true);
  if (lower && trans && !unit) EXO_RES(true, true, false);
  if (!lower && !trans && unit) EXO_RES(false, false, true);
  if (!lower && !trans && !unit) EXO_RES(false, false, false);
  if (!lower && trans && unit) EXO_RES(false, true, true);
  if (!lower && trans && !unit) EXO_RES(false, true, false);
#undef EXO_RES
  EXO_LAUNCH_CHECK();
}

// r = b - op(A)(xh + xl) for the FP32 refinement path.
void residual_f(Ctx& c, bool lower, bool trans, bool unit, int64_t n, const float* A, int64_t lda, const float* b,
                int64_t incb, const float* xh, const float* xl, float* r) {
  if (!trans) {
    const bool vec = (lda % 4) == 0 && aligned16(A);
    const dim3 g((unsigned)ceil_div(n, 4));
#define EXO_RR(L, U, V) \
  hipLaunchKernelGGL((residual_rows_kernel<L, U, V>), g, dim3(256), 0, c.stream, n, A, lda, b, incb, xh, xl, r)
    if (lower) {
      if (unit) { if (vec) EXO_RR(true, true, true); else EXO_RR(true, true, false); }
      else { if (vec) EXO_RR(true, false, true); else EXO_RR(true, false, false); }
    } else {
      if (unit) { if (vec) EXO_RR(false, true, true); else EXO_RR(false, true, false); }
      else { if (vec) EXO_RR(false, false, true); else EXO_RR(false, false, false); }
    }
#undef EXO_RR
  } else {
    const dim3 g((unsigned)ceil_div(n, 64));
#define EXO_RC(L, U) \
  hipLaunchKernelGGL((residual_cols_kernel<L, U>), g, dim3(512), 0, c.stream, n, A, lda, b, incb, xh, xl, r)
    if (lower) { if (unit) EXO_RC(true, true); else EXO_RC(true, false); }
    else { if (unit) EXO_RC(false, true); else EXO_RC(false, false); }
#undef EXO_RC
  }
  EXO_LAUNCH_CHECK();
}

}  // namespace

namespace ops {

void trsv_mixed_pack(Ctx& c, int64_t n, const float* A, int64_t lda, void* packed) {
  __bf16* p = static_cast<__bf16*>(packed);
  for (int64_t r0 = 0; r0 < n; r0 += 65535) {
    const int64_t rows = std::min<int64_t>(65535, n - r0);
    hipLaunchKernelGGL(to_bf16_kernel, dim3((unsigned)std::min<int64_t>(ceil_div(n, 256), 64), (unsigned)rows),
                       dim3(256), 0, c.stream, rows, n, A + r0 * lda, lda, p + r0 * n);
    EXO_LAUNCH_CHECK();
  }
}

void trsv_mixed(Ctx& c, int64_t Uplo, int64_t TransA, int64_t Diag, int64_t n, float* x, int64_t incx,
                const float* A, int64_t lda, int64_t update, int64_t refine, const void* packed) {
  if (n <= 0) return;
  const bool trans = TransA != 111, lower = Uplo == 122, unit = Diag == 132;
  const bool fwd = lower != trans;
  const dim3 gn((unsigned)ceil_div(n, 256));
  if (update == 32) {
    // scratch slot 2: x_hi, x_lo, r / d (FP32); b stays in the caller's x
    // until the final combine overwrites it
    if (refine == 1) {  // the refinement fused into the solve, one launch
      refined(c, fwd, trans, unit, n, A, lda, x, incx);
      return;
    }
    float* xh = static_cast<float*>(c.scratch((size_t)n * 3 * sizeof(float) + 1024, 2));
    float* xl = xh + ceil_div(n, 64) * 64;
    float* rd = xl + ceil_div(n, 64) * 64;
    hipLaunchKernelGGL(gather_f_kernel, gn, dim3(256), 0, c.stream, n, x, incx, xh);
    EXO_LAUNCH_CHECK();
    trsv_rm<float>(c, Uplo, TransA, Diag, n, xh, 1, A, lda);
    if (refine <= 0) {
      hipLaunchKernelGGL(narrow_f_copy_kernel, gn, dim3(256), 0, c.stream, n, xh, x, incx);
      EXO_LAUNCH_CHECK();
      return;
    }
    for (int64_t it = 0; it < refine; ++it) {
      residual_f(c, lower, trans, unit, n, A, lda, x, incx, xh, it == 0 ? nullptr : xl, rd);
      trsv_rm<float>(c, Uplo, TransA, Diag, n, rd, 1, A, lda);
      hipLaunchKernelGGL(acc_f_kernel, gn, dim3(256), 0, c.stream, n, xl, rd, it == 0 ? 1 : 0);
      EXO_LAUNCH_CHECK();
    }
    hipLaunchKernelGGL(combine_kernel, gn, dim3(256), 0, c.stream, n, xh, xl, x, incx);
    EXO_LAUNCH_CHECK();
    return;
  }
  // scratch slot 2: b, x, d, r (FP64) and the FP32 shadow of the solution
  char* ws = static_cast<char*>(c.scratch((size_t)n * (4 * sizeof(double) + sizeof(float)) + 1024, 2));
  double* b = reinterpret_cast<double*>(ws);
  double* xs = b + n;
  double* d = xs + n;
  double* r = d + n;
  float* xf = reinterpret_cast<float*>(r + n);
  const dim3 g1((unsigned)ceil_div(n, 256));
  hipLaunchKernelGGL(widen_kernel, g1, dim3(256), 0, c.stream, n, x, incx, b);
  EXO_LAUNCH_CHECK();
  const __bf16* Abf = nullptr;
  if (update == 16 && packed) {  // the caller's copy (trsv_mixed_pack of this A)
    Abf = static_cast<const __bf16*>(packed);
  } else if (update == 16) {  // BF16 copy of the matrix for the off-diagonal updates (scratch slot 3)
    void* p = c.scratch((size_t)n * n * sizeof(__bf16), 3);
    trsv_mixed_pack(c, n, A, lda, p);
    Abf = static_cast<const __bf16*>(p);
  }
  auto run_solve = [&](const double* rhs, double* out) {
    if (Abf)
      solve_any<__bf16>(c, fwd, trans, unit, n, A, lda, Abf, n, rhs, out, xf);
    else
      solve_any<float>(c, fwd, trans, unit, n, A, lda, A, lda, rhs, out, xf);
  };
  run_solve(b, xs);
  for (int64_t it = 0; it < refine; ++it) {
    residual(c, lower, trans, unit, n, A, lda, b, xs, r);
    run_solve(r, d);
    hipLaunchKernelGGL(add_kernel, g1, dim3(256), 0, c.stream, n, xs, d);
    EXO_LAUNCH_CHECK();
  }
  hipLaunchKernelGGL(narrow_kernel, g1, dim3(256), 0, c.stream, n, xs, x, incx);
  EXO_LAUNCH_CHECK();
}

}  // namespace ops
}  // namespace exoblas

// ------------------------------------------------------------ packed entry points
//
// No counterpart in the reference: the BF16 copy of the update-16 mixed
// solve made once by the caller (exoblas_strsv_mixed_pack_async) and reused
// by any number of solves with the same matrix (exoblas_strsv_mixed_packed_
// async), instead of being remade by every exoblas_strsv_mixed call.
// Arguments and checks as the generated exoblas_strsv_mixed_async.
#include "exoblas_amd_mixed.h"
#include "runtime.hpp"

extern "C" int exoblas_strsv_mixed_pack_async(hipStream_t stream, int_fast32_t n, struct exo_win_2f32c A,
                                              void* packed) {
  exoblas::Call call("exoblas_strsv_mixed_pack_async", exoblas::Call::kAsync, stream);
  if (n < 0) return call.fail_size("n", n);
  if (A.strides[1] != 1) return call.fail_stride("A", A.strides[1]);
  if (n > 1 && A.strides[0] < n) return call.fail_ld("A", A.strides[0]);
  call.window(0, A.data, 0, exoblas::span2_hi(n, n, A.strides[0]), sizeof(float), false, true);
  call.window(1, packed, 0, (int64_t)n * n, sizeof(__bf16), true, true);
  if (!call.stage()) return call.status();
  call.run([&] {
    exoblas::ops::trsv_mixed_pack(call.ctx(), (int64_t)n, call.dev<const float>(0), (int64_t)A.strides[0],
                                  call.dev<void>(1));
  });
  return call.finish();
}

extern "C" int exoblas_strsv_mixed_packed_async(hipStream_t stream, int_fast32_t Uplo, int_fast32_t TransA,
                                                int_fast32_t Diag, int_fast32_t n, struct exo_win_1f32 x,
                                                struct exo_win_2f32c A, const void* packed, int_fast32_t refine) {
  exoblas::Call call("exoblas_strsv_mixed_packed_async", exoblas::Call::kAsync, stream);
  if (!(Uplo == 121 || Uplo == 122)) return call.fail_enum("Uplo", Uplo);
  if (!(TransA == 111 || TransA == 112 || TransA == 113)) return call.fail_enum("TransA", TransA);
  if (!(Diag == 131 || Diag == 132)) return call.fail_enum("Diag", Diag);
  if (n < 0) return call.fail_size("n", n);
  if (refine < 0) return call.fail_size("refine", refine);
  if (A.strides[1] != 1) return call.fail_stride("A", A.strides[1]);
  if (n > 1 && A.strides[0] < n) return call.fail_ld("A", A.strides[0]);
  call.window(0, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(float), true,
              true);
  call.window(1, A.data, 0, exoblas::span2_hi(n, n, A.strides[0]), sizeof(float), false, true);
  call.window(2, packed, 0, (int64_t)n * n, sizeof(__bf16), false, true);
  if (!call.stage()) return call.status();
  call.run([&] {
    exoblas::ops::trsv_mixed(call.ctx(), (int64_t)Uplo, (int64_t)TransA, (int64_t)Diag, (int64_t)n,
                             call.dev<float>(0), (int64_t)x.strides[0], call.dev<const float>(1),
                             (int64_t)A.strides[0], 16, (int64_t)refine, call.dev<const void>(2));
  });
  return call.finish();
}
