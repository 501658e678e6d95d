// Split-precision FP32 products on the BF16 MFMA (v_mfma_f32_16x16x32_bf16).
//
// Each FP32 operand element x is split into BF16 pieces by repeated rounding
// of the exact remainder:  x = h + m + l + r  with h = bf16(x),
// m = bf16(x - h), l = bf16(x - h - m) (every difference is exact in FP32).
//   terms = 3:  two pieces (h, m):      h*h + h*m + m*h        |r| ~ 2^-17 |x|
//               (+ m*m on the syrk tiles that meet the diagonal)
//   terms = 4:  two pieces, adds m*m everywhere
//   terms = 6:  three pieces (h, m, l): adds m*m + h*l + l*h   |r| ~ 2^-26 |x|
// Products of BF16 pieces are exact in FP32 and the MFMA accumulates in FP32.
// With wave tiles of 64 x 64 the small products go to a second set of
// accumulators, added once at the end, so the main accumulator takes one
// rounding per K step instead of one per product (the accumulation error of
// the main sum is what the small terms would otherwise add to).
//
// Two passes:
//  1. pack: each operand, in any trans form, is split once into P BF16 planes
//     in the MFMA-fragment layout [plane][K/8][rows][8] (rows padded to the
//     tile, K to 32, zero filled), so 8 consecutive k of one row are 16
//     contiguous bytes;
//  2. product: 256 x 128 workgroup tiles of 8 waves (64 x 64 each, 4 x 4
//     MFMA tiles of 16 x 16) when they give every CU a workgroup, else
//     128 x 128 of 4 waves; K step 32, planes copied global -> LDS by the
//     load-to-LDS path (double-buffered), one 16-byte LDS read per fragment,
//     no bounds checks inside the K loop.
#include "device.hpp"
#include "gemm_core.hpp"
#include "ops.hpp"

namespace exoblas {
namespace split {

using gemm::Operand;
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));
typedef float v4f __attribute__((ext_vector_type(4)));

constexpr int KT = 32;   // K per tile = one MFMA K step
constexpr int G = KT / 8;
constexpr int NT = 256;  // pack kernel threads

__host__ __device__ inline int64_t pad(int64_t v, int64_t m) { return (v + m - 1) / m * m; }

template <int P>
__device__ __forceinline__ void pieces(float x, __bf16 (&out)[P]) {
  float r = x;
#pragma unroll
  for (int i = 0; i < P; ++i) {
    out[i] = (__bf16)r;
    r = r - (float)out[i];
  }
}

// out[((p * (Kp/8) + g) * Rp + r) * 8 + e] = piece p of X(r, 8g + e).
template <int P>
__global__ __launch_bounds__(NT) void pack_kernel(Operand<float> X, int64_t R, int64_t K, int64_t Rp, int64_t Kp,
                                                  __bf16* out) {
  const int64_t groups = Kp / 8;
  const int64_t idx = (int64_t)blockIdx.x * NT + threadIdx.x;
  if (idx >= groups * Rp) return;
  const int64_t g = idx / Rp, r = idx % Rp;
  float v[8];
  if (X.form == gemm::kRK && X.vec && r < R && 8 * g + 8 <= K) {
    const float4 a = *reinterpret_cast<const float4*>(X.p + r * X.ld + 8 * g);
    const float4 b = *reinterpret_cast<const float4*>(X.p + r * X.ld + 8 * g + 4);
    v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w; v[4] = b.x; v[5] = b.y; v[6] = b.z; v[7] = b.w;
  } else {
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      const int64_t k = 8 * g + e;
      v[e] = (r < R && k < K) ? (X.form == gemm::kRK ? X.p[r * X.ld + k] : X.p[k * X.ld + r]) : 0.f;
    }
  }
  v8bf q[P];
#pragma unroll
  for (int e = 0; e < 8; ++e) {
    __bf16 pc[P];
    pieces<P>(v[e], pc);
#pragma unroll
    for (int p = 0; p < P; ++p) q[p][e] = pc[p];
  }
#pragma unroll
  for (int p = 0; p < P; ++p) *reinterpret_cast<v8bf*>(out + (((int64_t)p * groups + g) * Rp + r) * 8) = q[p];
}

__device__ __forceinline__ v4f mma(v8bf a, v8bf b, v4f c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// LDS tile of one operand: [P][G][R + PADV] 16-byte fragments (8 BF16
// each); with PADV = 1 the extra vector per group row shifts the 4 lane
// groups of a fragment read onto different banks.  The packed planes hold
// the R rows of one (plane, k group) contiguously, so the tile is copied
// global -> LDS directly (global_load_lds_dwordx4: each lane moves 16 bytes
// and a wave fills 1 KiB of contiguous LDS), without passing through VGPRs.
template <int P, int R, int NTH, int PADV>
struct Stage {
  static constexpr int GS = R + PADV;
  static constexpr int TV = P * G * GS;     // vectors per tile in LDS (padded)
  static constexpr int CH = P * G * R / 64;  // 64-row chunks: one wave instruction each
  static constexpr int NW = NTH / 64;
  static_assert(R % 64 == 0 && CH % NW == 0, "tile / wave count mismatch");
  // Issues the copy of K tile kt into lds; the caller waits (vmcnt) and
  // synchronises the workgroup before reading it.
  __device__ __forceinline__ static void copy(const __bf16* planes, int64_t groups, int64_t Rp, int64_t r0,
                                              int64_t kt, uint4* lds) {
    const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
#pragma unroll
    for (int i = 0; i < CH / NW; ++i) {
      const int c = wave + NW * i;
      const int pg = c / (R / 64), rr = c % (R / 64);
      const int p = pg / G, gl = pg % G;
      const __bf16* src = planes + (((int64_t)p * groups + kt * G + gl) * Rp + r0 + rr * 64 + lane) * 8;
      __builtin_amdgcn_global_load_lds(src, (__attribute__((address_space(3))) void*)(lds + pg * GS + rr * 64), 16,
                                       0, 0);
    }
  }
};

__device__ __forceinline__ void wait_copies() { asm volatile("s_waitcnt vmcnt(0)" ::: "memory"); }

// Workgroup tile BM x BN, wave tile WM x WN, LDS row padding PADV (the
// 256-wide tile drops it to stay within 128 KiB of LDS).
template <int BM_, int BN_, int WM_, int WN_, int PADV_>
struct Shape {
  static constexpr int BM = BM_, BN = BN_, WM = WM_, WN = WN_, PADV = PADV_;
  static constexpr int WAVES_N = BN / WN;
  static constexpr int NTH = 64 * (BM / WM) * (BN / WN);
  static constexpr int SM = WM / 16, SN = WN / 16;
};
using Small = Shape<128, 128, 64, 64, 1>;
using Wide = Shape<256, 128, 64, 64, 0>;

template <int P, int TRI, class SH>
__global__ __launch_bounds__(SH::NTH) void product_kernel(int64_t M, int64_t N, int64_t Kp, float alpha,
                                                          const __bf16* Ap, int64_t Mp, const __bf16* Bp, int64_t Np,
                                                          float* Cp, int64_t ldc, int64_t tiles_m, int64_t tiles_n,
                                                          int all_mm) {
  using LA = Stage<P, SH::BM, SH::NTH, SH::PADV>;
  using LB = Stage<P, SH::BN, SH::NTH, SH::PADV>;
  constexpr int SM = SH::SM, SN = SH::SN;
  extern __shared__ __align__(16) unsigned char smem[];
  uint4* As = reinterpret_cast<uint4*>(smem);
  uint4* Bs = As + 2 * LA::TV;
  int64_t tm, tn;
  if (TRI == gemm::kFull)
    gemm::tile_coords(blockIdx.x, gridDim.x, tiles_m, tiles_n, tm, tn);
  else
    gemm::tri_tile_coords(TRI, blockIdx.x, gridDim.x, tiles_m, tiles_n, SH::BM, SH::BN, tm, tn);
  if (tm >= tiles_m || tn >= tiles_n) return;
  const int64_t m0 = tm * SH::BM, n0 = tn * SH::BN;
  const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
  const int wm0 = (wave / SH::WAVES_N) * SH::WM, wn0 = (wave % SH::WAVES_N) * SH::WN;
  const int lm = lane % 16, lg = lane / 16;
  const int64_t groups = Kp / 8;
  const int64_t nk = Kp / KT;
  // The m*m product: always with all_mm (the 4-product form); with 3 products
  // still on a syrk tile that meets the diagonal, which multiplies rows by
  // themselves there, so the dropped term would be m^2 >= 0, a bias that
  // grows with K rather than noise.
  const bool diag = all_mm || (TRI != gemm::kFull && m0 < n0 + SH::BN && n0 < m0 + SH::BM);
  // Wave tiles of at most 64 x 64 keep the small products in a second set of
  // accumulators, added once at the end: the main accumulator then takes one
  // rounding per K step (the h*h product) instead of one per product.
  constexpr bool DUAL = SM * SN <= 16;
  v4f acc[SM][SN], acc2[DUAL ? SM : 1][DUAL ? SN : 1];
#pragma unroll
  for (int a = 0; a < SM; ++a)
#pragma unroll
    for (int b = 0; b < SN; ++b) acc[a][b] = v4f{0.f, 0.f, 0.f, 0.f};
  if constexpr (DUAL) {
#pragma unroll
    for (int a = 0; a < SM; ++a)
#pragma unroll
      for (int b = 0; b < SN; ++b) acc2[a][b] = v4f{0.f, 0.f, 0.f, 0.f};
  }
  LA::copy(Ap, groups, Mp, m0, 0, As);
  LB::copy(Bp, groups, Np, n0, 0, Bs);
  wait_copies();
  __syncthreads();
  for (int64_t t = 0; t < nk; ++t) {
    const int cur = (int)(t & 1);
    const v8bf* as = reinterpret_cast<const v8bf*>(As + cur * LA::TV);
    const v8bf* bs = reinterpret_cast<const v8bf*>(Bs + cur * LB::TV);
    v8bf a[P][SM], b[P][SN];
#pragma unroll
    for (int p = 0; p < P; ++p) {
#pragma unroll
      for (int i = 0; i < SM; ++i) a[p][i] = as[(p * G + lg) * LA::GS + wm0 + i * 16 + lm];
#pragma unroll
      for (int i = 0; i < SN; ++i) b[p][i] = bs[(p * G + lg) * LB::GS + wn0 + i * 16 + lm];
    }
    // The other buffer was last read in step t - 1, before the barrier that
    // ended it: the copy of tile t + 1 overlaps this step's MFMAs.  It is
    // issued after this step's fragment reads so those never wait on it.
    if (t + 1 < nk) {
      LA::copy(Ap, groups, Mp, m0, t + 1, As + (cur ^ 1) * LA::TV);
      LB::copy(Bp, groups, Np, n0, t + 1, Bs + (cur ^ 1) * LB::TV);
    }
    // Term-major order: the accumulators are independent, so each term's
    // MFMAs issue back to back and a tile's next term never waits on the
    // previous MFMA into the same accumulator.
    auto term = [&](int pa, int pb) {
#pragma unroll
      for (int ta = 0; ta < SM; ++ta)
#pragma unroll
        for (int tb = 0; tb < SN; ++tb) {
          if (DUAL && (pa | pb))
            acc2[ta % (DUAL ? SM : 1)][tb % (DUAL ? SN : 1)] =
                mma(a[pa][ta], b[pb][tb], acc2[ta % (DUAL ? SM : 1)][tb % (DUAL ? SN : 1)]);
          else
            acc[ta][tb] = mma(a[pa][ta], b[pb][tb], acc[ta][tb]);
        }
    };
    if (P == 3) {
      term(2 % P, 0);
      term(0, 2 % P);
      term(1, 1);
    } else if (diag) {
      term(1, 1);
    }
    term(1, 0);
    term(0, 1);
    term(0, 0);
    wait_copies();
    __syncthreads();
  }
  if constexpr (DUAL) {
#pragma unroll
    for (int a = 0; a < SM; ++a)
#pragma unroll
      for (int b = 0; b < SN; ++b) acc[a][b] += acc2[a][b];
  }
  // D layout of the FP32-accumulating 16 x 16 MFMA: lane (lm, lg) holds rows
  // 4*lg + q of column lm.
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

template <int P>
void pack(Ctx& c, const Operand<float>& X, int64_t R, int64_t K, int64_t Rp, int64_t Kp, __bf16* out) {
  const int64_t items = (Kp / 8) * Rp;
  hipLaunchKernelGGL((pack_kernel<P>), dim3((unsigned)ceil_div(items, NT)), dim3(NT), 0, c.stream, X, R, K, Rp, Kp,
                     out);
  EXO_LAUNCH_CHECK();
}

template <int P, int TRI, class SH>
void product(Ctx& c, int64_t M, int64_t N, int64_t Kp, float alpha, const __bf16* Ap, int64_t Mp, const __bf16* Bp,
             int64_t Np, float* Cp, int64_t ldc, int all_mm) {
  const int64_t tm = Mp / SH::BM, tn = Np / SH::BN;
  const size_t lds =
      2 * 16 * (size_t)(Stage<P, SH::BM, SH::NTH, SH::PADV>::TV + Stage<P, SH::BN, SH::NTH, SH::PADV>::TV);
  static std::atomic<uint64_t> lds_set{0};
  ensure_lds_limit(reinterpret_cast<const void*>(&product_kernel<P, TRI, SH>), lds, c.device, lds_set);
  hipLaunchKernelGGL((product_kernel<P, TRI, SH>), dim3((unsigned)gemm::tri_tile_count(TRI, tm, tn, SH::BM, SH::BN)),
                     dim3(SH::NTH), lds, c.stream, M, N, Kp,
                     alpha, Ap, Mp, Bp, Np, Cp, ldc, tm, tn, all_mm);
  EXO_LAUNCH_CHECK();
}

template <int P, class SH>
void product_tri(Ctx& c, int tri, int64_t M, int64_t N, int64_t Kp, float alpha, const __bf16* Ap, int64_t Mp,
                 const __bf16* Bp, int64_t Np, float* Cp, int64_t ldc, int all_mm) {
  if (tri == gemm::kLower)
    product<P, gemm::kLower, SH>(c, M, N, Kp, alpha, Ap, Mp, Bp, Np, Cp, ldc, all_mm);
  else if (tri == gemm::kUpper)
    product<P, gemm::kUpper, SH>(c, M, N, Kp, alpha, Ap, Mp, Bp, Np, Cp, ldc, all_mm);
  else
    product<P, gemm::kFull, SH>(c, M, N, Kp, alpha, Ap, Mp, Bp, Np, Cp, ldc, all_mm);
}

// C (tri part) += alpha * X_A X_B^T, X_A M x K, X_B N x K; same_operand when
// X_B is X_A (syrk), which is then packed once.
template <int P>
void run_p(Ctx& c, int tri, int64_t M, int64_t N, int64_t K, float alpha, const Operand<float>& A,
           const Operand<float>& B, bool same_operand, float* Cp, int64_t ldc, int all_mm) {
  // 256 x 128 workgroup tiles (8 waves of 64 x 64, small products in their
  // own accumulators) while they give every CU a workgroup, else 128 x 128.
  // Rows are padded to the row tile, columns to the column tile; a syrk
  // operand, packed once for both sides, to the larger.
  const bool wide = gemm::tri_tile_count(tri, ceil_div(M, Wide::BM), ceil_div(N, Wide::BN), Wide::BM, Wide::BN) >=
                    (int64_t)c.num_cus;
  const int64_t TM = wide ? Wide::BM : Small::BM, TN = wide ? Wide::BN : Small::BN;
  const int64_t Kp = pad(K, KT);
  const int64_t Mp = same_operand ? pad(M, std::max(TM, TN)) : pad(M, TM);
  const int64_t Np = same_operand ? Mp : pad(N, TN);
  const size_t a_elems = (size_t)P * Kp * Mp, b_elems = same_operand ? 0 : (size_t)P * Kp * Np;
  __bf16* ws = static_cast<__bf16*>(c.scratch((a_elems + b_elems) * sizeof(__bf16), 5));
  __bf16* Ap = ws;
  __bf16* Bp = same_operand ? ws : ws + a_elems;
  pack<P>(c, A, M, K, Mp, Kp, Ap);
  if (!same_operand) pack<P>(c, B, N, K, Np, Kp, Bp);
  if (wide)
    product_tri<P, Wide>(c, tri, M, N, Kp, alpha, Ap, Mp, Bp, Np, Cp, ldc, all_mm);
  else
    product_tri<P, Small>(c, tri, M, N, Kp, alpha, Ap, Mp, Bp, Np, Cp, ldc, all_mm);
}

void run(Ctx& c, int tri, int64_t terms, int64_t M, int64_t N, int64_t K, float alpha, const Operand<float>& A,
         const Operand<float>& B, bool same_operand, float* Cp, int64_t ldc) {
  if (M <= 0 || N <= 0 || K <= 0) return;
  if (terms == 6)
    run_p<3>(c, tri, M, N, K, alpha, A, B, same_operand, Cp, ldc, 0);
  else
    run_p<2>(c, tri, M, N, K, alpha, A, B, same_operand, Cp, ldc, terms == 4 ? 1 : 0);
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
  split::run(c, kFull, terms, M, N, K, alpha, a, b, false, Cp, ldc);
}

void syrk_split(Ctx& c, int64_t Uplo, int64_t Trans, int64_t N, int64_t K, float alpha, const float* A, int64_t lda,
                const float* AT, int64_t ldat, float* Cp, int64_t ldc, int64_t terms) {
  using namespace gemm;
  const int tri = Uplo == 121 ? kUpper : kLower;
  const Operand<float> a = Trans == 111 ? split::operand(A, lda, kRK) : split::operand(AT, ldat, kKR);
  split::run(c, tri, terms, N, N, K, alpha, a, a, true, Cp, ldc);
}

}  // namespace ops
}  // namespace exoblas
