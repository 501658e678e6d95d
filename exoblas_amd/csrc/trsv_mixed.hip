// Mixed-precision strsv with iterative refinement.
//
// update = 32 (FP32-stored updates): classical refinement around the plain
// FP32 solve.
//   x_hi = solve(b)                 the FP32 look-back solve (trsv.hip)
//   repeat `refine` times:
//     r = b - op(A) (x_hi + x_lo)   FP64 accumulation over the FP32 matrix,
//                                   b read from the caller's vector, r stored
//                                   in FP32 as the next right-hand side
//     d = solve(r);  x_lo += d      (x_lo kept in FP32: it is ~eps |x|)
//   x = x_hi + x_lo                 added in FP64, rounded once to FP32
// The solution is carried as the FP32 pair (x_hi, x_lo), so no FP64 vector
// is stored; each solve is the plain kernel's chain (the previous block
// folded into the inverse ahead of time), so a refined solve costs two plain
// solves plus a residual pass.
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

void trsv_mixed(Ctx& c, int64_t Uplo, int64_t TransA, int64_t Diag, int64_t n, float* x, int64_t incx,
                const float* A, int64_t lda, int64_t update, int64_t refine) {
  if (n <= 0) return;
  const bool trans = TransA != 111, lower = Uplo == 122, unit = Diag == 132;
  const bool fwd = lower != trans;
  const dim3 gn((unsigned)ceil_div(n, 256));
  if (update == 32) {
    // scratch slot 2: x_hi, x_lo, r / d (FP32); b stays in the caller's x
    // until the final combine overwrites it
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
  if (update == 16) {  // BF16 copy of the matrix for the off-diagonal updates (scratch slot 3)
    __bf16* p = static_cast<__bf16*>(c.scratch((size_t)n * n * sizeof(__bf16), 3));
    for (int64_t r0 = 0; r0 < n; r0 += 65535) {
      const int64_t rows = std::min<int64_t>(65535, n - r0);
      hipLaunchKernelGGL(to_bf16_kernel, dim3((unsigned)std::min<int64_t>(ceil_div(n, 256), 64), (unsigned)rows),
                         dim3(256), 0, c.stream, rows, n, A + r0 * lda, lda, p + r0 * n);
      EXO_LAUNCH_CHECK();
    }
    Abf = p;
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
