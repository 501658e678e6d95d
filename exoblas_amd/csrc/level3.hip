// Level-3 kernels for gfx950: gemm / symm / syrk / syr2k / trmm on the MFMA
// tile engine (gemm_core.hpp) and the mscal / trmscal beta passes.
//
// Every product is  C[i, j] += alpha * sum_k X_A(i, k) X_B(j, k)  with the
// operand forms chosen on the host from TransA/TransB, Side/Uplo or
// Uplo/Trans; syrk restricts the update to one triangle of C.  The exo_*
// entry points apply no beta: the reference's wrappers scale C first with a
// separate mscal / trmscal call (test/level3/gemm/exo_gemm_wrapper.h:18-20,
// test/level3/syrk/exo_syrk_wrapper.h:18-20), and the same split is kept.
// The CBLAS-shaped gemm_ex extension fuses beta into the epilogue instead
// (BETA kernels: C = alpha * sum + beta * C, C not read when beta == 0).
#include "device.hpp"
#include "gemm_core.hpp"
#include "level3.hpp"
#include "ops.hpp"
#include "tiles.hpp"

namespace exoblas {
namespace gemm {

// Tiles [tile0, tile0 + gridDim.x) of the tile list (list_tile), one per
// workgroup.
template <class C, int TRI, int EXT = 0, bool BETA = false>
__global__ __launch_bounds__(C::NT) void gemm_kernel(int64_t M, int64_t N, int64_t K, typename C::type alpha,
                                                     Operand<typename C::type> A, Operand<typename C::type> B,
                                                     typename C::type* Cp, int64_t ldc, int cvec, int64_t tiles_m,
                                                     int64_t tiles_n, int64_t tile0, typename C::type beta) {
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
  tile_store<C, TRI, BETA>(Cp, ldc, cvec, M, N, m0, n0, alpha, acc, beta);
}

// Triangular-operand products (trmm) on row tiles in pairs (q, tm - 1 - q):
// row tile r's K window holds r + 1 K tiles for a lower op(A) and tm - r for
// an upper one, so every pair holds tm + 1 and ceil(tm / 2) x tn workgroups
// of equal work need no split of K and no reduction.
template <class C>
__global__ __launch_bounds__(C::NT) void gemm_pair_kernel(int64_t M, int64_t N, int64_t K, typename C::type alpha,
                                                          Operand<typename C::type> A, Operand<typename C::type> B,
                                                          typename C::type* Cp, int64_t ldc, int cvec,
                                                          int64_t tiles_m, int64_t tiles_n) {
  using T = typename C::type;
  extern __shared__ __align__(16) unsigned char smem[];
  const int64_t id = xcd_regroup(blockIdx.x, gridDim.x);
  const int64_t q = id / tiles_n, n0 = (id % tiles_n) * C::BN;
  if (q >= (tiles_m + 1) / 2) return;
  for (int h = 0; h < 2; ++h) {
    const int64_t tm = h == 0 ? q : tiles_m - 1 - q;
    if (h == 1 && tm == q) break;  // odd tile count: the middle row tile once
    const int64_t m0 = tm * C::BM;
    typename Mfma<T>::acc acc[C::SM][C::SN];
#pragma unroll
    for (int a = 0; a < C::SM; ++a)
#pragma unroll
      for (int b = 0; b < C::SN; ++b)
#pragma unroll
        for (int e = 0; e < 4; ++e) acc[a][b][e] = T(0);
    int64_t kb = 0, ke = K;
    tri_k_window<1>(A, m0, C::BM, kb, ke);
    tile_mma<C, 1>(A, B, M, N, K, m0, n0, kb, ke, acc, reinterpret_cast<T*>(smem));
    tile_store<C, kFull, false>(Cp, ldc, cvec, M, N, m0, n0, alpha, acc, T(0));
    __syncthreads();  // the second tile's first LDS stores follow every read of the first's
  }
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
template <class T, int TRI, int BM, int BN, bool BETA = false>
__global__ __launch_bounds__(256) void splitk_reduce_kernel(int64_t M, int64_t N, int S, int64_t tiles_m,
                                                            int64_t tiles_n, int64_t tile0, int64_t ntiles, T alpha,
                                                            const T* __restrict__ P, T* Cp, int64_t ldc, T beta) {
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
    if (BETA)
      Cp[row * ldc + col] = alpha * sum[e] + (beta != T(0) ? beta * Cp[row * ldc + col] : T(0));
    else
      Cp[row * ldc + col] += alpha * sum[e];
  }
}

// Split-K over tiles [tile0, tile0 + ntiles) of the list, S slices.
template <class C, int TRI, int EXT = 0, bool BETA = false>
void launch_splitk(Ctx& c, int64_t S, int64_t tile0, int64_t ntiles, int64_t M, int64_t N, int64_t K,
                   typename C::type alpha, const Operand<typename C::type>& A, const Operand<typename C::type>& B,
                   typename C::type* Cp, int64_t ldc, typename C::type beta = 0) {
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
  hipLaunchKernelGGL((splitk_reduce_kernel<T, TRI, C::BM, C::BN, BETA>),
                     dim3((unsigned)ntiles, (unsigned)ceil_div(C::BM, RPW)), dim3(256), 0, c.stream, M, N, (int)S, tm,
                     tn, tile0, ntiles, alpha, P, Cp, ldc, beta);
  EXO_LAUNCH_CHECK();
}

// Tiles [tile0, tile0 + ntiles) of the list, one workgroup each.
template <class C, int TRI, int EXT = 0, bool BETA = false>
void launch_cfg(Ctx& c, int64_t tile0, int64_t ntiles, int64_t M, int64_t N, int64_t K, typename C::type alpha,
                const Operand<typename C::type>& A, const Operand<typename C::type>& B, typename C::type* Cp,
                int64_t ldc, typename C::type beta = 0) {
  using T = typename C::type;
  const int64_t tm = ceil_div(M, C::BM), tn = ceil_div(N, C::BN);
  const int cvec = (ldc % C::EB == 0) && (reinterpret_cast<uintptr_t>(Cp) % (C::EB * sizeof(T)) == 0);
  static std::atomic<uint64_t> lds_set{0};
  ensure_lds_limit(reinterpret_cast<const void*>(&gemm_kernel<C, TRI, EXT, BETA>), C::LDS_BYTES, c.device, lds_set);
  hipLaunchKernelGGL((gemm_kernel<C, TRI, EXT, BETA>), dim3((unsigned)ntiles), dim3(C::NT), C::LDS_BYTES, c.stream, M,
                     N, K, alpha, A, B, Cp, ldc, cvec, tm, tn, tile0, beta);
  EXO_LAUNCH_CHECK();
}

// Narrow products: few output tiles and a K long enough to share.  32 x 32
// output tiles, one per workgroup of 8 waves; wave w accumulates the whole
// tile over its K slice [w * ks, (w + 1) * ks) straight from global memory
// (each operand element is read by one wave, so no LDS staging), and the 8
// partial tiles are summed through LDS in wave order, so results do not
// depend on timing.  An MFMA step's 4-deep reduction takes k = k0 + 4 * lg +
// e for lane group lg at step e of a 16-deep k group: a row-major operand
// row gives a lane its 4 values of a group in one 16-byte load.  One launch
// replaces the split-K pair (partial products + reduction) for these shapes
// and carries the beta epilogue.
namespace narrow {
constexpr int TM = 32, TN = 32, NW = 8, NT = 64 * NW, KG = 16, CG = 2;  // CG k groups per pipeline stage
constexpr int RS = TN + 1;                                             // LDS row stride of a partial tile

template <class T>
constexpr size_t lds_bytes() {
  return sizeof(T) * NW * TM * RS;
}

// X(r, k + e), e = 0..3, of a vector operand in form F (whole group in
// range): one 16-byte load of a row (kRK) or 4 rows of one column (kKR).
template <int F, class T>
__device__ __forceinline__ void load4(const Operand<T>& op, int64_t r, int64_t k, T (&v)[4]) {
  if constexpr (F == kRK) {
    const T* p = op.p + r * op.ld + k;
    if constexpr (sizeof(T) == 4) {
      const VecN<T, 4> x = *reinterpret_cast<const VecN<T, 4>*>(p);
#pragma unroll
      for (int e = 0; e < 4; ++e) v[e] = x.v[e];
    } else {
      const VecN<T, 2> x0 = *reinterpret_cast<const VecN<T, 2>*>(p);
      const VecN<T, 2> x1 = *reinterpret_cast<const VecN<T, 2>*>(p + 2);
      v[0] = x0.v[0];
      v[1] = x0.v[1];
      v[2] = x1.v[0];
      v[3] = x1.v[1];
    }
  } else {
    const T* p = op.p + k * op.ld + r;
#pragma unroll
    for (int e = 0; e < 4; ++e) v[e] = p[e * op.ld];
  }
}

// The same with every element checked against rows x [.., ke).
template <int F, class T>
__device__ __forceinline__ void load4_masked(const Operand<T>& op, int64_t r, int64_t k, int64_t rows, int64_t ke,
                                             T (&v)[4]) {
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    const int64_t kk = k + e;
    v[e] = (r < rows && kk < ke) ? (F == kRK ? op.p[r * op.ld + kk] : op.p[kk * op.ld + r]) : T(0);
  }
}

// FA / FB: the operands' forms (kRK or kKR), fixed at compile time so that
// the loads carry no branches (a runtime form test around each load costs
// the 512^3 FP32 product 11.2 -> 6.3 us on MI355X).
template <class T, int FA, int FB, bool BETA>
__global__ __launch_bounds__(NT) void narrow_kernel(int64_t M, int64_t N, int64_t K, int64_t ks, T alpha, T beta,
                                                    Operand<T> A, Operand<T> B, T* Cp, int64_t ldc, int64_t tiles_n) {
  using MF = Mfma<T>;
  extern __shared__ __align__(16) unsigned char smem[];
  T* red = reinterpret_cast<T*>(smem);
  const int64_t t = xcd_regroup(blockIdx.x, gridDim.x);
  const int64_t m0 = (t / tiles_n) * TM, n0 = (t % tiles_n) * TN;
  const int wave = threadIdx.x / 64, lane = threadIdx.x % 64, lm = lane % 16, lg = lane / 16;
  const int64_t kb = wave * ks, ke = kb + ks < K ? kb + ks : K;
  typename MF::acc acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int q = 0; q < 4; ++q) acc[a][b][q] = T(0);
  // one stage: CG k groups of both row groups of A and B
  struct Stage {
    T a[CG][2][4], b[CG][2][4];
  };
  auto mma = [&](const Stage& st) {
#pragma unroll
    for (int g = 0; g < CG; ++g)
#pragma unroll
      for (int e = 0; e < 4; ++e)
#pragma unroll
        for (int ta = 0; ta < 2; ++ta)
#pragma unroll
          for (int tb = 0; tb < 2; ++tb) acc[ta][tb] = MF::step(st.a[g][ta][e], st.b[g][tb][e], acc[ta][tb]);
  };
  // C is read before the product so its latency overlaps the K loop
  constexpr int PT = TM * TN / NT;  // outputs per thread in the epilogue
  T cin[PT];
#pragma unroll
  for (int j = 0; j < PT; ++j) {
    const int idx = threadIdx.x + NT * j;
    const int64_t gr = m0 + idx / TN, gc = n0 + idx % TN;
    cin[j] = (gr < M && gc < N && (!BETA || beta != T(0))) ? Cp[gr * ldc + gc] : T(0);
  }
  const bool fast = A.vec && B.vec && m0 + TM <= M && n0 + TN <= N;
  int64_t k = kb;
  if (fast && ke > kb) {
    // whole stages of CG groups, the next stage's loads issued before this
    // stage's MFMAs
    const int64_t nst = (ke - kb) / (CG * KG);
    auto load = [&](Stage& st, int64_t k0) {
#pragma unroll
      for (int g = 0; g < CG; ++g)
#pragma unroll
        for (int h = 0; h < 2; ++h) {
          load4<FA>(A, m0 + 16 * h + lm, k0 + g * KG + 4 * lg, st.a[g][h]);
          load4<FB>(B, n0 + 16 * h + lm, k0 + g * KG + 4 * lg, st.b[g][h]);
        }
    };
    Stage s0, s1;
    int64_t i = 0;
    if (nst > 0) load(s0, k);
    for (; i + 2 <= nst; i += 2) {
      load(s1, k + CG * KG);
      mma(s0);
      if (i + 2 < nst) load(s0, k + 2 * CG * KG);
      mma(s1);
      k += 2 * CG * KG;
    }
    if (i < nst) {
      mma(s0);
      k += CG * KG;
    }
  }
  // the rest of the slice (and every slice of an edge tile) element-checked
  for (; k < ke; k += KG) {
    T a[2][4], b[2][4];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      load4_masked<FA>(A, m0 + 16 * h + lm, k + 4 * lg, M, ke, a[h]);
      load4_masked<FB>(B, n0 + 16 * h + lm, k + 4 * lg, N, ke, b[h]);
    }
#pragma unroll
    for (int e = 0; e < 4; ++e)
#pragma unroll
      for (int ta = 0; ta < 2; ++ta)
#pragma unroll
        for (int tb = 0; tb < 2; ++tb) acc[ta][tb] = MF::step(a[ta][e], b[tb][e], acc[ta][tb]);
  }
  // partial tiles to LDS, then each thread sums its outputs over the waves in order
  T* mine = red + wave * TM * RS;
#pragma unroll
  for (int ta = 0; ta < 2; ++ta)
#pragma unroll
    for (int tb = 0; tb < 2; ++tb)
#pragma unroll
      for (int q = 0; q < 4; ++q) mine[(16 * ta + MF::out_row(lg, q)) * RS + 16 * tb + lm] = acc[ta][tb][q];
  __syncthreads();
#pragma unroll
  for (int j = 0; j < PT; ++j) {
    const int idx = threadIdx.x + NT * j, row = idx / TN, col = idx % TN;
    T s = T(0);
#pragma unroll
    for (int w = 0; w < NW; ++w) s += red[w * TM * RS + row * RS + col];
    const int64_t gr = m0 + row, gc = n0 + col;
    if (gr < M && gc < N) {
      T* cp = Cp + gr * ldc + gc;
      if (BETA)
        *cp = alpha * s + (beta != T(0) ? beta * cin[j] : T(0));
      else
        *cp = cin[j] + alpha * s;
    }
  }
}

inline int64_t tiles(int64_t M, int64_t N) { return ceil_div(M, TM) * ceil_div(N, TN); }

template <class T, int FA, int FB, bool BETA>
void launch_forms(Ctx& c, int64_t M, int64_t N, int64_t K, T alpha, const Operand<T>& A, const Operand<T>& B,
                  T* Cp, int64_t ldc, T beta) {
  const int64_t tn = ceil_div(N, TN), nt = tiles(M, N);
  const int64_t ks = ceil_div(ceil_div(K, NW), KG) * KG;
  static std::atomic<uint64_t> lds_set{0};
  ensure_lds_limit(reinterpret_cast<const void*>(&narrow_kernel<T, FA, FB, BETA>), lds_bytes<T>(), c.device, lds_set);
  hipLaunchKernelGGL((narrow_kernel<T, FA, FB, BETA>), dim3((unsigned)nt), dim3(NT), lds_bytes<T>(), c.stream, M, N,
                     K, ks, alpha, beta, A, B, Cp, ldc, tn);
  EXO_LAUNCH_CHECK();
}

template <class T, bool BETA>
void launch(Ctx& c, int64_t M, int64_t N, int64_t K, T alpha, const Operand<T>& A, const Operand<T>& B, T* Cp,
            int64_t ldc, T beta) {
  if (A.form == kRK && B.form == kRK)
    launch_forms<T, kRK, kRK, BETA>(c, M, N, K, alpha, A, B, Cp, ldc, beta);
  else if (A.form == kRK)
    launch_forms<T, kRK, kKR, BETA>(c, M, N, K, alpha, A, B, Cp, ldc, beta);
  else if (B.form == kRK)
    launch_forms<T, kKR, kRK, BETA>(c, M, N, K, alpha, A, B, Cp, ldc, beta);
  else
    launch_forms<T, kKR, kKR, BETA>(c, M, N, K, alpha, A, B, Cp, ldc, beta);
}

// Plain products (kRK / kKR operands: not symmetric, triangular or
// concatenated; the caller checks the forms) whose 32 x 32 tiles give at
// least half a round and at most two, with a K that gives every wave a stage
// or more.
inline bool fits(int64_t M, int64_t N, int64_t K, int64_t P) {
  const int64_t nt = tiles(M, N);
  return nt >= P / 2 && nt <= 2 * P && K >= NW * CG * KG;
}
}  // namespace narrow

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
template <class C, int TRI, int EXT = 0, bool BETA = false>
void launch_plan(Ctx& c, int64_t dp, int64_t S, int64_t M, int64_t N, int64_t K, typename C::type alpha,
                 const Operand<typename C::type>& A, const Operand<typename C::type>& B, typename C::type* Cp,
                 int64_t ldc, typename C::type beta = 0) {
  const int64_t total = tile_count<C>(TRI, M, N);
  if (dp > 0) launch_cfg<C, TRI, EXT, BETA>(c, 0, dp, M, N, K, alpha, A, B, Cp, ldc, beta);
  if (dp < total) launch_splitk<C, TRI, EXT, BETA>(c, S, dp, total - dp, M, N, K, alpha, A, B, Cp, ldc, beta);
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
template <class C, int TRI, int EXT = 0, bool BETA = false>
void launch_rounds(Ctx& c, int64_t M, int64_t N, int64_t K, typename C::type alpha,
                   const Operand<typename C::type>& A, const Operand<typename C::type>& B, typename C::type* Cp,
                   int64_t ldc, typename C::type beta = 0) {
  const int64_t total = tile_count<C>(TRI, M, N), P = c.num_cus;
  const int64_t dp = total / P * P, rem = total - dp;
  const int64_t S = rem ? std::min<int64_t>({P / rem, K / (4 * C::BK), 16}) : 1;
  if (S < 2)
    launch_plan<C, TRI, EXT, BETA>(c, total, 1, M, N, K, alpha, A, B, Cp, ldc, beta);
  else
    launch_plan<C, TRI, EXT, BETA>(c, dp, S, M, N, K, alpha, A, B, Cp, ldc, beta);
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
template <class T, int TRI, int EXT = 0, bool BETA = false>
void run(Ctx& c, int64_t M, int64_t N, int64_t K, T alpha, const Operand<T>& A, const Operand<T>& B, T* Cp,
         int64_t ldc, T beta = 0) {
  if (M <= 0 || N <= 0 || K <= 0) return;
  using Tl = Tiles<T>;
  using Huge = typename Tl::Huge;
  using Big = typename Tl::Big;
  using Small = typename Tl::Small;
  const int64_t P = c.num_cus;
  const int64_t huge = tile_count<Huge>(TRI, M, N), big = tile_count<Big>(TRI, M, N);
  if (huge >= P) return launch_rounds<Huge, TRI, EXT, BETA>(c, M, N, K, alpha, A, B, Cp, ldc, beta);
  if (big >= 2 * P) return launch_rounds<Big, TRI, EXT, BETA>(c, M, N, K, alpha, A, B, Cp, ldc, beta);
  const int64_t sh = best_split(huge, P, std::min<int64_t>(16, K / (4 * Huge::BK)), 0.75);
  if (sh >= 2) return launch_plan<Huge, TRI, EXT, BETA>(c, 0, sh, M, N, K, alpha, A, B, Cp, ldc, beta);
  if (big >= P) return launch_rounds<Big, TRI, EXT, BETA>(c, M, N, K, alpha, A, B, Cp, ldc, beta);
  const int64_t small = tile_count<Small>(TRI, M, N);
  if constexpr (TRI == kFull && EXT == 0)
    if (A.form <= kKR && B.form <= kKR && small < P && narrow::fits(M, N, K, P))
      return narrow::launch<T, BETA>(c, M, N, K, alpha, A, B, Cp, ldc, beta);
  // Few output tiles and a long K: split K so every CU gets two workgroups,
  // each K slice at least 128 deep.
  const int64_t S = std::min<int64_t>({ceil_div(2 * P, small), K / 128, 16});
  if (S >= 2) return launch_plan<Small, TRI, EXT, BETA>(c, 0, S, M, N, K, alpha, A, B, Cp, ldc, beta);
  launch_rounds<Small, TRI, EXT, BETA>(c, M, N, K, alpha, A, B, Cp, ldc, beta);
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
  using Big = typename Tiles<T>::Big;
  const int64_t btm = ceil_div(M, Big::BM), btn = ceil_div(N, Big::BN);
  if (K == M && ((btm + 1) / 2) * btn >= c.num_cus) {
    // square triangle and a full round of row-tile pairs: one balanced launch
    const int cvec = (ldc % Big::EB == 0) && (reinterpret_cast<uintptr_t>(Cp) % (Big::EB * sizeof(T)) == 0);
    static std::atomic<uint64_t> lds_set{0};
    ensure_lds_limit(reinterpret_cast<const void*>(&gemm_pair_kernel<Big>), Big::LDS_BYTES, c.device, lds_set);
    hipLaunchKernelGGL((gemm_pair_kernel<Big>), dim3((unsigned)(((btm + 1) / 2) * btn)), dim3(Big::NT),
                       Big::LDS_BYTES, c.stream, M, N, K, alpha, A, B, Cp, ldc, cvec, btm, btn);
    EXO_LAUNCH_CHECK();
    return;
  }
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
constexpr int64_t kRowMajor = 101;

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

// CBLAS-shaped gemm: C = alpha op(A) op(B) + beta C, C not read when beta ==
// 0, beta fused into the product's epilogue.  Order 102 (column-major) is the
// row-major product C^T = op(B)^T op(A)^T on the same arrays: output rows are
// C's columns, X_A reads op(B) and X_B reads op(A), each in the operand form
// its storage gives.  alpha == 0 or K == 0 leaves only beta C (zeros when
// beta == 0).
template <class T>
void gemm_ex_impl(Ctx& c, int64_t Order, int64_t TransA, int64_t TransB, int64_t M, int64_t N, int64_t K, T alpha,
                  const T* A, int64_t lda, const T* B, int64_t ldb, T beta, T* Cp, int64_t ldc) {
  using namespace gemm;
  if (M <= 0 || N <= 0) return;
  const bool row = Order == kRowMajor, ta = TransA != kNoTrans, tb = TransB != kNoTrans;
  const int64_t rows = row ? M : N, cols = row ? N : M;
  if (alpha == T(0) || K <= 0) {
    if (beta == T(1)) return;
    tiles::UpdateArgs<T> g{};
    g.alpha = beta;
    if (beta == T(0))
      tiles::launch_update<T, 3, tiles::kFull>(c.stream, rows, cols, Cp, ldc, g);
    else
      tiles::launch_update<T, 0, tiles::kFull>(c.stream, rows, cols, Cp, ldc, g);
    return;
  }
  if (row)
    run<T, kFull, 0, true>(c, M, N, K, alpha, operand(A, lda, ta ? kKR : kRK), operand(B, ldb, tb ? kRK : kKR), Cp,
                           ldc, beta);
  else
    run<T, kFull, 0, true>(c, N, M, K, alpha, operand(B, ldb, tb ? kKR : kRK), operand(A, lda, ta ? kRK : kKR), Cp,
                           ldc, beta);
}

void gemm_ex(Ctx& c, int64_t Order, int64_t TransA, int64_t TransB, int64_t M, int64_t N, int64_t K, float alpha,
             const float* A, int64_t lda, const float* B, int64_t ldb, float beta, float* Cp, int64_t ldc) {
  gemm_ex_impl<float>(c, Order, TransA, TransB, M, N, K, alpha, A, lda, B, ldb, beta, Cp, ldc);
}
void gemm_ex(Ctx& c, int64_t Order, int64_t TransA, int64_t TransB, int64_t M, int64_t N, int64_t K, double alpha,
             const double* A, int64_t lda, const double* B, int64_t ldb, double beta, double* Cp, int64_t ldc) {
  gemm_ex_impl<double>(c, Order, TransA, TransB, M, N, K, alpha, A, lda, B, ldb, beta, Cp, ldc);
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
