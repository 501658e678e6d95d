// Split-precision FP32 products on the BF16 MFMA (v_mfma_f32_16x16x32_bf16).
//
// Each FP32 operand element x is split into BF16 pieces by repeated rounding
// of the exact remainder:  x = h + m + l + r  with h = bf16(x),
// m = bf16(x - h), l = bf16(x - h - m) (every difference is exact in FP32).
//   terms = 3:  two pieces (h, m):     h*h + h*m + m*h        |r| ~ 2^-17 |x|
//   terms = 6:  three pieces (h, m, l): adds m*m + h*l + l*h  |r| ~ 2^-26 |x|
// Products of BF16 pieces are exact in FP32 and the MFMA accumulates in FP32;
// the small terms are issued first so they enter the accumulator before the
// dominant h*h term of the same K step.
//
// Tile: 128 x 128 per workgroup of 4 waves (64 x 64 each, 4 x 4 MFMA tiles of
// 16 x 16), K step 32.  Operand tiles are read from global memory as FP32 with
// the level-3 tile loader (any trans form), split while being written to LDS
// in row-major [row][k] BF16 layout, so a lane's MFMA fragment (8 consecutive
// k of one row) is one 16-byte LDS read per piece.
#include "device.hpp"
#include "gemm_core.hpp"
#include "ops.hpp"

namespace exoblas {
namespace split {

using gemm::Operand;
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));
typedef __bf16 v4bf __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

constexpr int BM = 128, BN = 128, BK = 32, NT = 256, WM = 64, WN = 64;
constexpr int SM = WM / 16, SN = WN / 16;
constexpr int KS = BK + 8;  // LDS row stride in BF16 elements (80 bytes)

using Loader = gemm::TileLoader<float, 128, BK, NT>;

template <int P>
__device__ __forceinline__ void pieces(float x, __bf16 (&out)[P]) {
  float r = x;
#pragma unroll
  for (int i = 0; i < P; ++i) {
    out[i] = (__bf16)r;
    r = r - (float)out[i];
  }
}

// Writes the loader's FP32 registers to LDS as P BF16 planes [P][R][KS].
template <int P, int R>
__device__ __forceinline__ void store_split(const Loader& L, __bf16* lds) {
  constexpr int VN = 4;
  constexpr int NE = Loader::NE;
  const int tid = threadIdx.x;
  if (L.mode == 0) {  // 4 consecutive k of one row
#pragma unroll
    for (int i = 0; i < NE / VN; ++i) {
      const int v = tid + NT * i;
      const int r = v / (BK / VN), kv = v % (BK / VN);
      v4bf q[P];
#pragma unroll
      for (int e = 0; e < VN; ++e) {
        __bf16 pc[P];
        pieces<P>(L.reg[i * VN + e], pc);
#pragma unroll
        for (int p = 0; p < P; ++p) q[p][e] = pc[p];
      }
#pragma unroll
      for (int p = 0; p < P; ++p) *reinterpret_cast<v4bf*>(lds + (p * R + r) * KS + kv * VN) = q[p];
    }
  } else if (L.mode == 1) {  // 4 consecutive rows of one k
#pragma unroll
    for (int i = 0; i < NE / VN; ++i) {
      const int v = tid + NT * i;
      const int k = v / (R / VN), rv = v % (R / VN);
#pragma unroll
      for (int e = 0; e < VN; ++e) {
        __bf16 pc[P];
        pieces<P>(L.reg[i * VN + e], pc);
#pragma unroll
        for (int p = 0; p < P; ++p) lds[(p * R + rv * VN + e) * KS + k] = pc[p];
      }
    }
  } else {
#pragma unroll
    for (int j = 0; j < NE; ++j) {
      const int idx = tid + NT * j;
      const int k = idx / R, r = idx % R;
      __bf16 pc[P];
      pieces<P>(L.reg[j], pc);
#pragma unroll
      for (int p = 0; p < P; ++p) lds[(p * R + r) * KS + k] = pc[p];
    }
  }
}

__device__ __forceinline__ v4f mma(v8bf a, v8bf b, v4f c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

template <int P, int TRI>
__global__ __launch_bounds__(NT) void split_kernel(int64_t M, int64_t N, int64_t K, float alpha, Operand<float> A,
                                                   Operand<float> B, float* Cp, int64_t ldc, int64_t tiles_m,
                                                   int64_t tiles_n) {
  __shared__ __align__(16) __bf16 As[P * BM * KS];
  __shared__ __align__(16) __bf16 Bs[P * BN * KS];
  int64_t tm, tn;
  gemm::tile_coords(blockIdx.x, gridDim.x, tiles_m, tiles_n, tm, tn);
  const int64_t m0 = tm * BM, n0 = tn * BN;
  if (TRI == gemm::kLower && n0 > m0 + BM - 1) return;
  if (TRI == gemm::kUpper && n0 + BN - 1 < m0) return;
  const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
  const int wm0 = (wave / 2) * WM, wn0 = (wave % 2) * WN;
  const int lm = lane % 16, lg = lane / 16;
  v4f acc[SM][SN];
#pragma unroll
  for (int a = 0; a < SM; ++a)
#pragma unroll
    for (int b = 0; b < SN; ++b) acc[a][b] = v4f{0.f, 0.f, 0.f, 0.f};
  Loader la, lb;
  const int64_t nk = (K + BK - 1) / BK;
  la.load(A, m0, 0, M, K);
  lb.load(B, n0, 0, N, K);
  for (int64_t t = 0; t < nk; ++t) {
    __syncthreads();
    store_split<P, BM>(la, As);
    store_split<P, BN>(lb, Bs);
    __syncthreads();
    if (t + 1 < nk) {
      la.load(A, m0, (t + 1) * BK, M, K);
      lb.load(B, n0, (t + 1) * BK, N, K);
    }
    v8bf a[P][SM], b[P][SN];
#pragma unroll
    for (int p = 0; p < P; ++p) {
#pragma unroll
      for (int ta = 0; ta < SM; ++ta)
        a[p][ta] = *reinterpret_cast<const v8bf*>(As + (p * BM + wm0 + ta * 16 + lm) * KS + 8 * lg);
#pragma unroll
      for (int tb = 0; tb < SN; ++tb)
        b[p][tb] = *reinterpret_cast<const v8bf*>(Bs + (p * BN + wn0 + tb * 16 + lm) * KS + 8 * lg);
    }
#pragma unroll
    for (int ta = 0; ta < SM; ++ta)
#pragma unroll
      for (int tb = 0; tb < SN; ++tb) {
        v4f c = acc[ta][tb];
        if (P == 3) {
          c = mma(a[2][ta], b[0][tb], c);
          c = mma(a[0][ta], b[2][tb], c);
          c = mma(a[1][ta], b[1][tb], c);
        }
        c = mma(a[1][ta], b[0][tb], c);
        c = mma(a[0][ta], b[1][tb], c);
        c = mma(a[0][ta], b[0][tb], c);
        acc[ta][tb] = c;
      }
  }
  // D layout of the 16 x 16 MFMA: lane (lm, lg) holds rows 4*lg + q of column lm.
#pragma unroll
  for (int ta = 0; ta < SM; ++ta)
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int64_t row = m0 + wm0 + ta * 16 + 4 * lg + q;
      if (row >= M) continue;
#pragma unroll
      for (int tb = 0; tb < SN; ++tb) {
        const int64_t col = n0 + wn0 + tb * 16 + lm;
        bool in = col < N;
        if (TRI == gemm::kLower) in = in && col <= row;
        if (TRI == gemm::kUpper) in = in && col >= row;
        if (in) Cp[row * ldc + col] += alpha * acc[ta][tb][q];
      }
    }
}

template <int P, int TRI>
void launch(Ctx& c, int64_t M, int64_t N, int64_t K, float alpha, const Operand<float>& A, const Operand<float>& B,
            float* Cp, int64_t ldc) {
  const int64_t tm = ceil_div(M, BM), tn = ceil_div(N, BN);
  hipLaunchKernelGGL((split_kernel<P, TRI>), dim3((unsigned)(tm * tn)), dim3(NT), 0, c.stream, M, N, K, alpha, A, B,
                     Cp, ldc, tm, tn);
  EXO_LAUNCH_CHECK();
}

void run(Ctx& c, int tri, int64_t terms, int64_t M, int64_t N, int64_t K, float alpha, const Operand<float>& A,
         const Operand<float>& B, float* Cp, int64_t ldc) {
  if (M <= 0 || N <= 0 || K <= 0) return;
  const bool six = terms == 6;
  if (tri == gemm::kLower) {
    if (six) launch<3, gemm::kLower>(c, M, N, K, alpha, A, B, Cp, ldc);
    else launch<2, gemm::kLower>(c, M, N, K, alpha, A, B, Cp, ldc);
  } else if (tri == gemm::kUpper) {
    if (six) launch<3, gemm::kUpper>(c, M, N, K, alpha, A, B, Cp, ldc);
    else launch<2, gemm::kUpper>(c, M, N, K, alpha, A, B, Cp, ldc);
  } else {
    if (six) launch<3, gemm::kFull>(c, M, N, K, alpha, A, B, Cp, ldc);
    else launch<2, gemm::kFull>(c, M, N, K, alpha, A, B, Cp, ldc);
  }
}

Operand<float> operand(const float* p, int64_t ld, int form) {
  Operand<float> o;
  o.p = p;
  o.ld = ld;
  o.form = form;
  o.vec = (ld % 4 == 0) && aligned16(p);
  return o;
}

}  // namespace split

namespace ops {

void gemm_split(Ctx& c, int64_t TransA, int64_t TransB, int64_t M, int64_t N, int64_t K, float alpha,
                const float* A, int64_t lda, const float* B, int64_t ldb, const float* AT, int64_t ldat,
                const float* BT, int64_t ldbt, float* Cp, int64_t ldc, int64_t terms) {
  using namespace gemm;
  const Operand<float> a = TransA == 111 ? split::operand(A, lda, kRK) : split::operand(AT, ldat, kKR);
  const Operand<float> b = TransB == 111 ? split::operand(B, ldb, kKR) : split::operand(BT, ldbt, kRK);
  split::run(c, kFull, terms, M, N, K, alpha, a, b, Cp, ldc);
}

void syrk_split(Ctx& c, int64_t Uplo, int64_t Trans, int64_t N, int64_t K, float alpha, const float* A, int64_t lda,
                const float* AT, int64_t ldat, float* Cp, int64_t ldc, int64_t terms) {
  using namespace gemm;
  const int tri = Uplo == 121 ? kUpper : kLower;
  if (Trans == 111) {
    const Operand<float> a = split::operand(A, lda, kRK);
    split::run(c, tri, terms, N, N, K, alpha, a, a, Cp, ldc);
  } else {
    const Operand<float> a = split::operand(AT, ldat, kKR);
    split::run(c, tri, terms, N, N, K, alpha, a, a, Cp, ldc);
  }
}

}  // namespace ops
}  // namespace exoblas
