// Level-3 kernels for gfx950: gemm / symm / syrk / syr2k / trmm on the MFMA
// tile engine (gemm_core.hpp) and the mscal / trmscal beta passes.
//
// Every product is  C[i, j] += alpha * sum_k X_A(i, k) X_B(j, k)  with the
// operand forms chosen on the host from TransA/TransB, Side/Uplo or
// Uplo/Trans; syrk restricts the update to one triangle of C.  beta is not
// applied here: the reference's wrappers scale C first with a separate
// mscal / trmscal call (test/level3/gemm/exo_gemm_wrapper.h:18-20,
// test/level3/syrk/exo_syrk_wrapper.h:18-20), and the same split is kept.
#include "device.hpp"
#include "gemm_core.hpp"
#include "level3.hpp"
#include "ops.hpp"
#include "tiles.hpp"

namespace exoblas {
namespace gemm {

// Tiles [tile0, tile0 + gridDim.x) of the tile list (list_tile), one per
// workgroup.
template <class C, int TRI, int EXT = 0>
__global__ __launch_bounds__(C::NT) void gemm_kernel(int64_t M, int64_t N, int64_t K, typename C::type alpha,
                                                     Operand<typename C::type> A, Operand<typename C::type> B,
                                                     typename C::type* Cp, int64_t ldc, int cvec, int64_t tiles_m,
                                                     int64_t tiles_n, int64_t tile0) {
  using T = typename C::type;
  extern __shared__ __align__(16) unsigned char smem[];
  int64_t tm, tn;
  list_tile(TRI, tile0 + xcd_regroup(blockIdx.x, gridDim.x), tiles_m, tiles_n, C::BM, C::BN, tm, tn);
  if (tm >= tiles_m || tn >= tiles_n) return;
  const int64_t m0 = tm * C::BM, n0 = tn * C::BN;
  typename Mfma<T>::acc acc[C::SM][C::SN];
#pragma unroll
  for (int a = 0; a < C::SM; ++a)
#pragma unroll
    for (int b = 0; b < C::SN; ++b)
#pragma unroll
      for (int q = 0; q < 4; ++q) acc[a][b][q] = T(0);
  int64_t kb = 0, ke = K;
  if constexpr (EXT == 1 || EXT == 3 || EXT == 4) tri_k_window<EXT>(A, m0, C::BM, kb, ke);
  tile_mma<C, EXT>(A, B, M, N, K, m0, n0, kb, ke, acc, reinterpret_cast<T*>(smem));
  tile_store<C, TRI>(Cp, ldc, cvec, M, N, m0, n0, alpha, acc);
}

// Split-K over tiles [tile0, tile0 + ntiles) of the list: workgroup
// s * ntiles + l runs the K slice [s * ksplit, (s + 1) * ksplit) of tile
// tile0 + l and writes its raw sums as a dense BM x BN block
// P[(s * ntiles + l) * BM * BN ...]; splitk_reduce_kernel then adds
// alpha * sum_s of the blocks to C in a fixed order.
template <class C, int TRI, int EXT = 0>
__global__ __launch_bounds__(C::NT) void gemm_splitk_kernel(int64_t M, int64_t N, int64_t K, int64_t ksplit,
                                                            Operand<typename C::type> A, Operand<typename C::type> B,
                                                            typename C::type* P, int64_t tiles_m, int64_t tiles_n,
                                                            int64_t tile0, int64_t ntiles) {
  using T = typename C::type;
  extern __shared__ __align__(16) unsigned char smem[];
  const int64_t s = blockIdx.x / ntiles, l = blockIdx.x % ntiles;
  int64_t tm, tn;
  list_tile(TRI, tile0 + l, tiles_m, tiles_n, C::BM, C::BN, tm, tn);
  if (tm >= tiles_m || tn >= tiles_n) return;
  const int64_t m0 = tm * C::BM, n0 = tn * C::BN;
  typename Mfma<T>::acc acc[C::SM][C::SN];
#pragma unroll
  for (int a = 0; a < C::SM; ++a)
#pragma unroll
    for (int b = 0; b < C::SN; ++b)
#pragma unroll
      for (int q = 0; q < 4; ++q) acc[a][b][q] = T(0);
  int64_t kb = s * ksplit, ke = kb + ksplit < K ? kb + ksplit : K;
  if constexpr (EXT == 1 || EXT == 3 || EXT == 4)
    tri_k_window<EXT>(A, m0, C::BM, kb, ke);  // an empty window leaves the sums zero
  tile_mma<C, EXT>(A, B, M, N, K, m0, n0, kb, ke, acc, reinterpret_cast<T*>(smem));
  tile_store<C, kFull, true>(P + (s * ntiles + l) * C::BM * C::BN, C::BN, 1, C::BM, C::BN, 0, 0, T(1), acc, T(0));
}

// Workgroup (l, r) adds rows [r * 256 / (BN / VN), ...) of tile tile0 + l:
// each thread one 16-byte column group of one row, summed over the S slices
// in order.
template <class T, int TRI, int BM, int BN>
__global__ __launch_bounds__(256) void splitk_reduce_kernel(int64_t M, int64_t N, int S, int64_t tiles_m,
                                                            int64_t tiles_n, int64_t tile0, int64_t ntiles, T alpha,
                                                            const T* __restrict__ P, T* Cp, int64_t ldc) {
  using V = typename Vec16<T>::type;
  constexpr int VN = Vec16<T>::n, GPR = BN / VN, RPW = 256 / GPR;  // groups per row, rows per workgroup
  const int64_t l = blockIdx.x;
  int64_t tm, tn;
  list_tile(TRI, tile0 + l, tiles_m, tiles_n, BM, BN, tm, tn);
  if (tm >= tiles_m || tn >= tiles_n) return;
  const int r = blockIdx.y * RPW + threadIdx.x / GPR, g = threadIdx.x % GPR;
  const int64_t row = tm * BM + r, col0 = tn * BN + g * VN;
  if (r >= BM || row >= M) return;
  T sum[VN];
#pragma unroll
  for (int e = 0; e < VN; ++e) sum[e] = T(0);
  const T* p = P + l * BM * BN + (int64_t)r * BN + g * VN;
  for (int k0 = 0; k0 < S; k0 += 8) {  // 8 slices' loads in flight, summed in slice order
    V v[8];
#pragma unroll
    for (int j = 0; j < 8; ++j)
      if (k0 + j < S) v[j] = ld_stream<T>(p + (int64_t)(k0 + j) * ntiles * BM * BN);
#pragma unroll
    for (int j = 0; j < 8; ++j)
      if (k0 + j < S)
#pragma unroll
        for (int e = 0; e < VN; ++e) sum[e] += vget<T>(v[j], e);
  }
#pragma unroll
  for (int e = 0; e < VN; ++e) {
    const int64_t col = col0 + e;
    if (col >= N || (TRI == kLower && col > row) || (TRI == kUpper && col < row)) continue;
    Cp[row * ldc + col] += alpha * sum[e];
  }
}

// Split-K over tiles [tile0, tile0 + ntiles) of the list, S slices.
template <class C, int TRI, int EXT = 0>
void launch_splitk(Ctx& c, int64_t S, int64_t tile0, int64_t ntiles, int64_t M, int64_t N, int64_t K,
                   typename C::type alpha, const Operand<typename C::type>& A, const Operand<typename C::type>& B,
                   typename C::type* Cp, int64_t ldc) {
  using T = typename C::type;
  const int64_t tm = ceil_div(M, C::BM), tn = ceil_div(N, C::BN);
  const int64_t ksplit = ceil_div(ceil_div(K, S), C::BK) * C::BK;
  S = ceil_div(K, ksplit);
  T* P = static_cast<T*>(c.scratch(sizeof(T) * (size_t)(S * ntiles * C::BM * C::BN), 4));
  static std::atomic<uint64_t> lds_set{0};
  ensure_lds_limit(reinterpret_cast<const void*>(&gemm_splitk_kernel<C, TRI, EXT>), C::LDS_BYTES, c.device, lds_set);
  hipLaunchKernelGGL((gemm_splitk_kernel<C, TRI, EXT>), dim3((unsigned)(S * ntiles)), dim3(C::NT), C::LDS_BYTES,
                     c.stream, M, N, K, ksplit, A, B, P, tm, tn, tile0, ntiles);
  EXO_LAUNCH_CHECK();
  constexpr int RPW = 256 / (C::BN / Vec16<T>::n);
  hipLaunchKernelGGL((splitk_reduce_kernel<T, TRI, C::BM, C::BN>), dim3((unsigned)ntiles, (unsigned)ceil_div(C::BM, RPW)),
                     dim3(256), 0, c.stream, M, N, (int)S, tm, tn, tile0, ntiles, alpha, P, Cp, ldc);
  EXO_LAUNCH_CHECK();
}

// Tiles [tile0, tile0 + ntiles) of the list, one workgroup each.
template <class C, int TRI, int EXT = 0>
void launch_cfg(Ctx& c, int64_t tile0, int64_t ntiles, int64_t M, int64_t N, int64_t K, typename C::type alpha,
                const Operand<typename C::type>& A, const Operand<typename C::type>& B, typename C::type* Cp,
                int64_t ldc) {
  using T = typename C::type;
  const int64_t tm = ceil_div(M, C::BM), tn = ceil_div(N, C::BN);
  const int cvec = (ldc % C::EB == 0) && (reinterpret_cast<uintptr_t>(Cp) % (C::EB * sizeof(T)) == 0);
  static std::atomic<uint64_t> lds_set{0};
  ensure_lds_limit(reinterpret_cast<const void*>(&gemm_kernel<C, TRI, EXT>), C::LDS_BYTES, c.device, lds_set);
  hipLaunchKernelGGL((gemm_kernel<C, TRI, EXT>), dim3((unsigned)ntiles), dim3(C::NT), C::LDS_BYTES, c.stream, M, N, K,
                     alpha, A, B, Cp, ldc, cvec, tm, tn, tile0);
  EXO_LAUNCH_CHECK();
}

// Tile shapes, largest first (measured on 8192^3 NN: FP32 256 x 256 x 32 in
// 8 waves of 128 x 64 127.6 TFLOP/s, 4 waves of 128 x 128 123.8, 128 x 128 x
// 16 114; FP64 128 x 256 x 16 in 8 waves of 64 x 64 57.9 TFLOP/s, 64 x 128 x
// 16 in 4 waves 46.3, 128 x 128 x 16 in 4 waves 39).  The FP32 128 x 128
// tiles run in 8 waves of 32 x 64 with 32-deep K steps: where they are
// chosen (fewer large tiles than CUs) that beats 4 waves of 64 x 64 with
// 16-deep steps (sgemm 3000^3 75.6 -> 90.4, ssyrk 4096 x 1024 85.5 -> 90.9).
template <class T> struct Tiles;
// The 64 x 64 tiles serve small outputs, where a workgroup runs only a few K
// steps and each step costs about one memory latency: deep K tiles (64 FP32
// / 32 FP64) halve or quarter the steps (sgemm 512^3 13 -> 11 us, 1024^3
// 40 -> 60 TFLOP/s on MI355X against 16-deep tiles).
template <> struct Tiles<float> {
  using Huge = Cfg<float, 256, 256, 32, 128, 64>;
  using Big = Cfg<float, 128, 128, 32, 32, 64>;  // 8 waves, 32-deep K steps
  using Small = Cfg<float, 64, 64, 64, 32, 32>;
};
template <> struct Tiles<double> {
  using Huge = Cfg<double, 128, 256, 16, 64, 64>;
  using Big = Cfg<double, 64, 128, 16, 32, 64>;
  using Small = Cfg<double, 64, 64, 32, 32, 32>;
};

// Workgroup tiles of BM x BN an M x N product runs (for a triangular update
// only those that meet the triangle).
template <class C>
int64_t tile_count(int tri, int64_t M, int64_t N) {
  return tri_tile_count(tri, ceil_div(M, C::BM), ceil_div(N, C::BN), C::BM, C::BN);
}

// One configuration's launches: tiles [0, dp) one workgroup each, the rest
// split S ways along K (S = 1: dp is every tile).
template <class C, int TRI, int EXT = 0>
void launch_plan(Ctx& c, int64_t dp, int64_t S, int64_t M, int64_t N, int64_t K, typename C::type alpha,
                 const Operand<typename C::type>& A, const Operand<typename C::type>& B, typename C::type* Cp,
                 int64_t ldc) {
  const int64_t total = tile_count<C>(TRI, M, N);
  if (dp > 0) launch_cfg<C, TRI, EXT>(c, 0, dp, M, N, K, alpha, A, B, Cp, ldc);
  if (dp < total) launch_splitk<C, TRI, EXT>(c, S, dp, total - dp, M, N, K, alpha, A, B, Cp, ldc);
}

// Split factor in [2, smax] that fills the last round of S * tiles units
// over P CUs best (the smallest on ties); 0 when no S >= 2 fills at least
// `need` of it.
inline int64_t best_split(int64_t tiles, int64_t P, int64_t smax, double need) {
  int64_t best = 0;
  double best_fill = need;
  for (int64_t S = 2; S <= smax; ++S) {
    const int64_t units = S * tiles;
    const double fill = (double)units / (double)(ceil_div(units, P) * P);
    if (fill > best_fill + 1e-9) {
      best = S;
      best_fill = fill;
    }
  }
  return best;
}

// Tiles [0, dp) of C's tile list run whole, the remaining ones split along K
// into the stream's free CUs: the full rounds keep every CU busy and the
// last, partial round is spread over all of them.
template <class C, int TRI, int EXT = 0>
void launch_rounds(Ctx& c, int64_t M, int64_t N, int64_t K, typename C::type alpha,
                   const Operand<typename C::type>& A, const Operand<typename C::type>& B, typename C::type* Cp,
                   int64_t ldc) {
  const int64_t total = tile_count<C>(TRI, M, N), P = c.num_cus;
  const int64_t dp = total / P * P, rem = total - dp;
  const int64_t S = rem ? std::min<int64_t>({P / rem, K / (4 * C::BK), 16}) : 1;
  if (S < 2)
    launch_plan<C, TRI, EXT>(c, total, 1, M, N, K, alpha, A, B, Cp, ldc);
  else
    launch_plan<C, TRI, EXT>(c, dp, S, M, N, K, alpha, A, B, Cp, ldc);
}

// Plan choice, from measurements on MI355X (FP32 TFLOP/s, old rule -> this):
// the largest tiles sustain about 80 % of the MFMA peak even at one tile per
// CU (ssymm 4096^2: 105.8 with 128 x 128 tiles -> 125.8), so they run
// whenever they make a full round; a partial last round is split along K
// (ssyrk 8192 x 2048 84.2 -> 111.5, dsyrk 4096 x 1024 34.4 -> 48.5 FP64).
// Fewer large tiles than CUs: 128 x 128 tiles when they give two per CU
// (ssyrk 4096 x 1024 71.1 -> 84.8), else the large tiles split along K when
// that fills the rounds (sgemm 2048^3 81.2 -> 100.2), else 128 x 128 tiles,
// else 64 x 64 tiles split along K for small outputs with a long K (sgemm
// 512^3, 1024^3).
template <class T, int TRI, int EXT = 0>
void run(Ctx& c, int64_t M, int64_t N, int64_t K, T alpha, const Operand<T>& A, const Operand<T>& B, T* Cp,
         int64_t ldc) {
  if (M <= 0 || N <= 0 || K <= 0) return;
  using Tl = Tiles<T>;
  using Huge = typename Tl::Huge;
  using Big = typename Tl::Big;
  using Small = typename Tl::Small;
  const int64_t P = c.num_cus;
  const int64_t huge = tile_count<Huge>(TRI, M, N), big = tile_count<Big>(TRI, M, N);
  if (huge >= P) return launch_rounds<Huge, TRI, EXT>(c, M, N, K, alpha, A, B, Cp, ldc);
  if (big >= 2 * P) return launch_rounds<Big, TRI, EXT>(c, M, N, K, alpha, A, B, Cp, ldc);
  const int64_t sh = best_split(huge, P, std::min<int64_t>(16, K / (4 * Huge::BK)), 0.75);
  if (sh >= 2) return launch_plan<Huge, TRI, EXT>(c, 0, sh, M, N, K, alpha, A, B, Cp, ldc);
  if (big >= P) return launch_rounds<Big, TRI, EXT>(c, M, N, K, alpha, A, B, Cp, ldc);
  // Few output tiles and a long K: split K so every CU gets two workgroups,
  // each K slice at least 128 deep.
  const int64_t small = tile_count<Small>(TRI, M, N);
  const int64_t S = std::min<int64_t>({ceil_div(2 * P, small), K / 128, 16});
  if (S >= 2) return launch_plan<Small, TRI, EXT>(c, 0, S, M, N, K, alpha, A, B, Cp, ldc);
  launch_rounds<Small, TRI, EXT>(c, M, N, K, alpha, A, B, Cp, ldc);
}

// Triangular-operand products (trmm).  With at least half a round of large
// tiles, two launches on the same row tiling: the K range strictly inside the
// triangle with plain operand loads (EXT 3), split four ways along K (row tile
// t of a lower op(A) has t K blocks, of an upper one the rest: the slices
// outside a tile's range exit at once and the dispatcher spreads the uneven
// remainder over the CUs), then each row tile's diagonal block with the
// per-element triangle (EXT 4).  Smaller products: one launch over the whole
// triangle window (EXT 1).
template <class T>
void run_triangular(Ctx& c, int64_t M, int64_t N, int64_t K, T alpha, const Operand<T>& A, const Operand<T>& B,
                    T* Cp, int64_t ldc) {
  if (M <= 0 || N <= 0 || K <= 0) return;
  using Huge = typename Tiles<T>::Huge;
  const int64_t tiles = tile_count<Huge>(kFull, M, N);
  const int64_t S = std::min<int64_t>(4, K / (4 * Huge::BK));  // 8: 1065 -> 1131 us (strmm 4096^2)
  if (2 * tiles >= c.num_cus && S >= 2) {
    launch_plan<Huge, kFull, 3>(c, 0, S, M, N, K, alpha, A, B, Cp, ldc);
    launch_cfg<Huge, kFull, 4>(c, 0, tiles, M, N, K, alpha, A, B, Cp, ldc);
  } else {
    run<T, kFull, 1>(c, M, N, K, alpha, A, B, Cp, ldc);
  }
}

template <class T>
Operand<T> operand(const T* p, int64_t ld, int form) {
  constexpr int VN = 16 / sizeof(T);
  Operand<T> o;
  o.p = p;
  o.ld = ld;
  o.form = form;
  o.vec = (ld % VN == 0) && aligned16(p);
  return o;
}

template <class T>
void product(Ctx& c, int tri, int64_t M, int64_t N, int64_t K, T alpha, const Operand<T>& A, const Operand<T>& B,
             T* Cp, int64_t ldc) {
  // triangular (1) or concatenated (2) operands take the kernels built with that code
  const int ext = (A.form >= kTri || B.form >= kTri) ? 1 : (A.k1 != INT64_MAX || B.k1 != INT64_MAX) ? 2 : 0;
  if (ext == 1) {
    run<T, kFull, 1>(c, M, N, K, alpha, A, B, Cp, ldc);  // trmm: full output
  } else if (ext == 2) {
    if (tri == kLower)
      run<T, kLower, 2>(c, M, N, K, alpha, A, B, Cp, ldc);
    else if (tri == kUpper)
      run<T, kUpper, 2>(c, M, N, K, alpha, A, B, Cp, ldc);
    else
      run<T, kFull, 2>(c, M, N, K, alpha, A, B, Cp, ldc);
  } else if (tri == kLower) {
    run<T, kLower>(c, M, N, K, alpha, A, B, Cp, ldc);
  } else if (tri == kUpper) {
    run<T, kUpper>(c, M, N, K, alpha, A, B, Cp, ldc);
  } else {
    run<T, kFull>(c, M, N, K, alpha, A, B, Cp, ldc);
  }
}

}  // namespace gemm

namespace {

constexpr int kNT = 256;
constexpr int64_t kNoTrans = 111;
constexpr int64_t kUpperV = 121;
constexpr int64_t kLeftV = 141;

// Bc[i, 0:n) = B[i, 0:n) and B[i, 0:n) = 0 for row i = blockIdx.y: trmm
// reads its right-hand operand from the copy and accumulates into B.
template <class T>
__global__ __launch_bounds__(kNT) void copy_clear_kernel(int64_t i0, int64_t n, T* B, int64_t ldb, T* Bc,
                                                         int64_t ldc) {
  const int64_t i = i0 + blockIdx.y;
  T* row = B + i * ldb;
  T* out = Bc + i * ldc;
  for (int64_t j = (int64_t)blockIdx.x * kNT + threadIdx.x; j < n; j += (int64_t)gridDim.x * kNT) {
    out[j] = row[j];
    row[j] = T(0);
  }
}

// A[i, j] *= alpha over the full M x N matrix (TRI 0) or the lower (1) /
// upper (2) triangle of an N x N matrix: the streaming tile update of
// tiles.hpp (each stored element read and written once, 8 rows of loads in
// flight per lane).
template <class T, int TRI>
void launch_mscal(Ctx& c, int64_t M, int64_t N, T alpha, T* A, int64_t lda) {
  tiles::UpdateArgs<T> g{};
  g.alpha = alpha;
  tiles::launch_update<T, 0, TRI>(c.stream, M, N, A, lda, g);
}

}  // namespace

namespace ops {

// src/level3/gemm.py:17-47.  X_A(i, k) = op(A)[i, k]; X_B(j, k) = op(B)[k, j].
template <class T>
void gemm(Ctx& c, int64_t TransA, int64_t TransB, int64_t M, int64_t N, int64_t K, T alpha, const T* A,
          int64_t lda, const T* B, int64_t ldb, const T* AT, int64_t ldat, const T* BT, int64_t ldbt, T* Cp,
          int64_t ldc) {
  using namespace gemm;
  const Operand<T> a = TransA == kNoTrans ? operand(A, lda, kRK) : operand(AT, ldat, kKR);
  const Operand<T> b = TransB == kNoTrans ? operand(B, ldb, kKR) : operand(BT, ldbt, kRK);
  product<T>(c, kFull, M, N, K, alpha, a, b, Cp, ldc);
}

// src/level3/symm.py:16-47.  Left: C += alpha S B (S = Aleft, M x M);
// Right: C += alpha B S (S = Aright, N x N); S read from its Uplo triangle.
template <class T>
void symm_rm(Ctx& c, int64_t Side, int64_t Uplo, int64_t M, int64_t N, T alpha, const T* Aleft, int64_t ldal,
             const T* Aright, int64_t ldar, const T* B, int64_t ldb, T* Cp, int64_t ldc) {
  using namespace gemm;
  const int sym = Uplo == kUpperV ? kSymUpper : kSymLower;
  if (Side == kLeftV)
    product<T>(c, kFull, M, N, M, alpha, operand(Aleft, ldal, sym), operand(B, ldb, kKR), Cp, ldc);
  else
    product<T>(c, kFull, M, N, N, alpha, operand(B, ldb, kRK), operand(Aright, ldar, sym), Cp, ldc);
}

// src/level3/syrk.py:17-32.  NoTrans: C += alpha A A^T; Trans: C += alpha A^T A,
// on the Uplo triangle of C only.
template <class T>
void syrk_rm(Ctx& c, int64_t Uplo, int64_t Trans, int64_t N, int64_t K, T alpha, const T* A, int64_t lda,
             const T* A_, int64_t lda_, const T* AT, int64_t ldat, const T* AT_, int64_t ldat_, T* Cp,
             int64_t ldc) {
  using namespace gemm;
  const int tri = Uplo == kUpperV ? kUpper : kLower;
  if (Trans == kNoTrans)
    product<T>(c, tri, N, N, K, alpha, operand(A, lda, kRK), operand(A_, lda_, kRK), Cp, ldc);
  else
    product<T>(c, tri, N, N, K, alpha, operand(AT, ldat, kKR), operand(AT_, ldat_, kKR), Cp, ldc);
}

// [X | Y] along K: k < K reads x, k >= K reads y at k - K.
template <class T>
gemm::Operand<T> concat(gemm::Operand<T> x, const gemm::Operand<T>& y, int64_t K) {
  x.p2 = y.p;
  x.ld2 = y.ld;
  x.k1 = K;
  x.vec = x.vec && y.vec;
  return x;
}

// src/level3/syr2k.py:14-39.  NoTrans: C += alpha (A B^T + B A^T); Trans:
// C += alpha (A^T B + B^T A), on the Uplo triangle of C: one triangular
// product over 2K with X_A = [A | B] and X_B = [B | A].
template <class T>
void syr2k_rm(Ctx& c, int64_t Uplo, int64_t Trans, int64_t N, int64_t K, T alpha, const T* A, int64_t lda,
              const T* A_, int64_t lda_, const T* B, int64_t ldb, const T* B_, int64_t ldb_, const T* AT,
              int64_t ldat, const T* AT_, int64_t ldat_, const T* BT, int64_t ldbt, const T* BT_, int64_t ldbt_,
              T* Cp, int64_t ldc) {
  using namespace gemm;
  const int tri = Uplo == kUpperV ? kUpper : kLower;
  if (N <= 0 || K <= 0) return;
  if (Trans == kNoTrans)
    product<T>(c, tri, N, N, 2 * K, alpha, concat(operand(A, lda, kRK), operand(B_, ldb_, kRK), K),
               concat(operand(B, ldb, kRK), operand(A_, lda_, kRK), K), Cp, ldc);
  else
    product<T>(c, tri, N, N, 2 * K, alpha, concat(operand(AT, ldat, kKR), operand(BT_, ldbt_, kKR), K),
               concat(operand(BT, ldbt, kKR), operand(AT_, ldat_, kKR), K), Cp, ldc);
}

// src/level3/trmm.py:12-83.  B <- alpha op(A) B, A triangular m x m (unit
// diagonal when Diag != 0), B m x n in place: B is copied to scratch (row
// length padded to 16 bytes) and cleared in one pass, then the product with
// the triangular operand form accumulates into it.  K tiles outside the
// triangle are skipped, and large products split K (run_triangular).
template <class T>
void trmm_left(Ctx& c, bool upper, bool trans, int64_t m, int64_t n, T alpha, const T* A, int64_t lda, T* B,
               int64_t ldb, int64_t Diag) {
  using namespace gemm;
  if (m <= 0 || n <= 0) return;
  constexpr int VN = Vec16<T>::n;
  const int64_t ldp = ceil_div(n, VN) * VN;
  T* Bc = static_cast<T*>(c.scratch(sizeof(T) * (size_t)(m * ldp), 1));
  const unsigned gx = (unsigned)std::max<int64_t>(1, std::min<int64_t>(ceil_div(n, kNT), 64));
  for (int64_t r0 = 0; r0 < m; r0 += 65535) {
    const int64_t rows = std::min<int64_t>(65535, m - r0);
    hipLaunchKernelGGL((copy_clear_kernel<T>), dim3(gx, (unsigned)rows), dim3(kNT), 0, c.stream, r0, n, B, ldb, Bc,
                       ldp);
    EXO_LAUNCH_CHECK();
  }
  // op(A) is lower for (Lower, NoTrans) and (Upper, Trans); X_A(i, k) reads
  // A[i, k] without a transpose and A[k, i] with one.
  const bool lower_op = upper == trans;
  const Operand<T> a = operand(A, lda, tri_form(trans ? kKR : kRK, !lower_op, Diag != 0));
  run_triangular<T>(c, m, n, m, alpha, a, operand<T>(Bc, ldp, kKR), B, ldb);
}

// src/level3/trmm.py:12-23
template <class T>
void trmm_row_major_Left_Upper_NonTrans(Ctx& c, int64_t m, int64_t n, T alpha, const T* A, int64_t lda, T* B,
                                        int64_t ldb, int64_t Diag) {
  trmm_left<T>(c, true, false, m, n, alpha, A, lda, B, ldb, Diag);
}
// src/level3/trmm.py:26-46
template <class T>
void trmm_row_major_Left_Lower_NonTrans(Ctx& c, int64_t m, int64_t n, T alpha, const T* A, int64_t lda, T* B,
                                        int64_t ldb, int64_t Diag) {
  trmm_left<T>(c, false, false, m, n, alpha, A, lda, B, ldb, Diag);
}
// src/level3/trmm.py:49-69
template <class T>
void trmm_row_major_Left_Upper_Trans(Ctx& c, int64_t m, int64_t n, T alpha, const T* A, int64_t lda, T* B,
                                     int64_t ldb, int64_t Diag) {
  trmm_left<T>(c, true, true, m, n, alpha, A, lda, B, ldb, Diag);
}
// src/level3/trmm.py:72-83
template <class T>
void trmm_row_major_Left_Lower_Trans(Ctx& c, int64_t m, int64_t n, T alpha, const T* A, int64_t lda, T* B,
                                     int64_t ldb, int64_t Diag) {
  trmm_left<T>(c, false, true, m, n, alpha, A, lda, B, ldb, Diag);
}

// src/level3/mscal.py:12-17
template <class T>
void mscal_rm(Ctx& c, int64_t M, int64_t N, T alpha, T* A, int64_t lda) {
  launch_mscal<T, 0>(c, M, N, alpha, A, lda);
}

// src/level3/mscal.py:20-27
template <class T>
void trmscal_rm(Ctx& c, int64_t Uplo, int64_t N, T alpha, T* A, int64_t lda) {
  if (Uplo == kUpperV)
    launch_mscal<T, 2>(c, N, N, alpha, A, lda);
  else
    launch_mscal<T, 1>(c, N, N, alpha, A, lda);
}

#define EXO_L3_INST(T)                                                                                        \
  template void gemm<T>(Ctx&, int64_t, int64_t, int64_t, int64_t, int64_t, T, const T*, int64_t, const T*,    \
                        int64_t, const T*, int64_t, const T*, int64_t, T*, int64_t);                          \
  template void mscal_rm<T>(Ctx&, int64_t, int64_t, T, T*, int64_t);                                          \
  template void trmscal_rm<T>(Ctx&, int64_t, int64_t, T, T*, int64_t);                                        \
  template void symm_rm<T>(Ctx&, int64_t, int64_t, int64_t, int64_t, T, const T*, int64_t, const T*, int64_t, \
                           const T*, int64_t, T*, int64_t);                                                   \
  template void syrk_rm<T>(Ctx&, int64_t, int64_t, int64_t, int64_t, T, const T*, int64_t, const T*, int64_t, \
                           const T*, int64_t, const T*, int64_t, T*, int64_t);                                \
  template void syr2k_rm<T>(Ctx&, int64_t, int64_t, int64_t, int64_t, T, const T*, int64_t, const T*, int64_t,\
                            const T*, int64_t, const T*, int64_t, const T*, int64_t, const T*, int64_t,       \
                            const T*, int64_t, const T*, int64_t, T*, int64_t);
EXO_L3_INST(float)
EXO_L3_INST(double)
template void trmm_row_major_Left_Upper_NonTrans<float>(Ctx&, int64_t, int64_t, float, const float*, int64_t,
                                                        float*, int64_t, int64_t);
template void trmm_row_major_Left_Lower_NonTrans<float>(Ctx&, int64_t, int64_t, float, const float*, int64_t,
                                                        float*, int64_t, int64_t);
template void trmm_row_major_Left_Upper_Trans<float>(Ctx&, int64_t, int64_t, float, const float*, int64_t,
                                                     float*, int64_t, int64_t);
template void trmm_row_major_Left_Lower_Trans<float>(Ctx&, int64_t, int64_t, float, const float*, int64_t,
                                                     float*, int64_t, int64_t);

}  // namespace ops

namespace gemm {
// Exported to the other kernel families (batched, split precision).
template void product<float>(Ctx&, int, int64_t, int64_t, int64_t, float, const Operand<float>&,
                             const Operand<float>&, float*, int64_t);
template void product<double>(Ctx&, int, int64_t, int64_t, int64_t, double, const Operand<double>&,
                              const Operand<double>&, double*, int64_t);
template Operand<float> operand<float>(const float*, int64_t, int);
template Operand<double> operand<double>(const double*, int64_t, int);
}  // namespace gemm
}  // namespace exoblas
