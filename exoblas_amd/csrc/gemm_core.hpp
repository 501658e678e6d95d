// MFMA tile engine for gfx950: C[tile] += alpha * X_A * X_B^T over a K range.
//
// Both operands are described by their logical form X(r, k), r the output
// row (A) or column (B) index, k the reduction index, so gemm / symm / syrk in
// every trans / side / uplo combination are the same kernel with different
// operand descriptors:
//   kRK        X(r, k) = p[r * ld + k]      (k contiguous: A NoTrans, B^T)
//   kKR        X(r, k) = p[k * ld + r]      (r contiguous: A^T, B NoTrans)
//   kSymLower  X(r, k) = S(r, k), S symmetric, lower triangle stored
//   kSymUpper  X(r, k) = S(r, k), upper triangle stored
//   tri_form(base, upper, unit)  X(r, k) = the kRK / kKR value where k <= r
//              (lower) or k >= r (upper), 0 elsewhere, 1 on the diagonal when
//              unit (triangular operands of trmm)
// A symmetric or triangular operand resolves per K tile to a plain kRK or kKR
// copy when the tile lies on one side of the diagonal (or to zeros outside a
// triangle), and to an element-wise select on the diagonal.
//
// Block tile BM x BN, BK-deep K tiles double-buffered in LDS in k-major layout
// ([BK][BM + pad]).  A workgroup of NT = 64 * (BM/WM) * (BN/WN) threads; each
// wave owns a WM x WN sub-tile made of 16 x 16 MFMA tiles
// (v_mfma_f32_16x16x4f32 / v_mfma_f64_16x16x4f64).  A lane's LDS fragment
// read is one 16-byte ds_read covering E consecutive rows; those E values feed
// E different MFMA tiles, so MFMA row index i of tile t maps to wave row
// 16*E*(t/E) + E*i + t%E.  The output mapping is the same permutation, which
// makes each lane's epilogue a run of E-wide vector read-modify-writes.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "device.hpp"

namespace exoblas {
namespace gemm {

enum Form : int { kRK = 0, kKR = 1, kSymLower = 2, kSymUpper = 3, kTri = 4 };

// Triangular form codes kTri + base + 2 * upper + 4 * unit (base kRK / kKR).
__host__ __device__ constexpr int tri_form(int base, bool upper, bool unit) {
  return kTri + base + (upper ? 2 : 0) + (unit ? 4 : 0);
}
enum Tri : int { kFull = 0, kLower = 1, kUpper = 2 };

// An operand may be the concatenation of two along K: X(r, k) reads (p, ld)
// for k < k1 and (p2, ld2) at k - k1 beyond (syr2k's [A | B]); both parts
// share the form.
template <class T>
struct Operand {
  const T* p;
  int64_t ld;
  int form;
  int vec;  // 16-byte loads allowed (aligned bases, lds a multiple of the vector width)
  const T* p2 = nullptr;
  int64_t ld2 = 0;
  int64_t k1 = INT64_MAX;
};

typedef float v4f __attribute__((ext_vector_type(4)));
typedef double v4d __attribute__((ext_vector_type(4)));

// One 16 x 16 x 4 MFMA step and its output layout: lane (lm = lane % 16,
// lg = lane / 16) holds column lm of D in its 4 accumulator registers q, at
// row out_row(lg, q).  The FP32 form keeps 4 consecutive rows per lane group;
// the FP64 form interleaves the lane groups (row 4q + lg), as measured on
// gfx950.
template <class T> struct Mfma;
template <> struct Mfma<float> {
  using acc = v4f;
  static __device__ __forceinline__ acc step(float a, float b, acc c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int out_row(int lg, int q) { return 4 * lg + q; }
};
template <> struct Mfma<double> {
  using acc = v4d;
  static __device__ __forceinline__ acc step(double a, double b, acc c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int out_row(int lg, int q) { return 4 * q + lg; }
};

template <class T, int N> struct alignas(N * sizeof(T)) VecN { T v[N]; };

__host__ __device__ constexpr int cmin(int a, int b) { return a < b ? a : b; }

// Global -> registers -> LDS copy of one R x BK operand tile.
template <class T, int R, int BK, int NT>
struct TileLoader {
  static constexpr int VN = 16 / sizeof(T);
  static constexpr int NE = R * BK / NT;
  static constexpr int LDS_STRIDE = R + VN;  // padded row (one 16-byte vector)
  static_assert(NE % VN == 0 && NE > 0, "tile / thread count mismatch");
  T reg[NE];
  int mode;  // 0: kRK vector, 1: kKR vector, 2: element-wise

  // EXT 1 and 4: the operand may be triangular (trmm), checked per K tile and
  // on the diagonal per element; 3: a triangular operand read only strictly
  // inside its triangle (the kernel clips K to that), so plain loads; 2:
  // concatenated along K (syr2k).  The plain kernels (EXT 0) compile without
  // any of it, which would cost them registers.
  template <int EXT>
  __device__ __forceinline__ static T elem(const Operand<T>& op, int form, int64_t r, int64_t k) {
    // form is kRK or kKR, symmetric (select by triangle) or triangular
    if constexpr (EXT == 3)
      if (form >= kTri) form = (form - kTri) & 1;
    if constexpr (EXT == 1 || EXT == 4) {
      if (form >= kTri) {
        const int f = form - kTri;
        if (r == k && (f & 4)) return T(1);
        if ((f & 2) ? k < r : k > r) return T(0);
        form = f & 1;
      }
    }
    if (form == kSymLower) form = r >= k ? kRK : kKR;
    if (form == kSymUpper) form = r <= k ? kRK : kKR;
    const T* p = op.p;
    int64_t ld = op.ld;
    if constexpr (EXT == 2) {
      if (k >= op.k1) {
        p = op.p2;
        ld = op.ld2;
        k -= op.k1;
      }
    }
    return form == kRK ? p[r * ld + k] : p[k * ld + r];
  }

  template <int EXT>
  __device__ __forceinline__ void load(const Operand<T>& op, int64_t r0, int64_t k0, int64_t rows, int64_t K) {
    int form = op.form;
    if constexpr (EXT == 3)
      if (form >= kTri) form = (form - kTri) & 1;
    if (form == kSymLower) {
      if (r0 >= k0 + BK - 1) form = kRK;
      else if (r0 + R - 1 < k0) form = kKR;
    } else if (form == kSymUpper) {
      if (r0 + R - 1 <= k0) form = kRK;
      else if (r0 > k0 + BK - 1) form = kKR;
    } else if ((EXT == 1 || EXT == 4) && form >= kTri) {
      const int f = form - kTri;
      // strictly inside the triangle: the plain form; wholly outside: zeros
      const bool inside = (f & 2) ? k0 > r0 + R - 1 : k0 + BK - 1 < r0;
      const bool outside = (f & 2) ? k0 + BK - 1 < r0 : k0 > r0 + R - 1;
      if (inside) form = f & 1;
      if (outside) {
        mode = 2;
#pragma unroll
        for (int j = 0; j < NE; ++j) reg[j] = T(0);
        return;
      }
    }
    // a concatenated operand: tiles past k1 read the second part, a tile
    // across k1 goes element-wise
    const T* p = op.p;
    int64_t ld = op.ld, kb = k0;
    bool across = false;
    if constexpr (EXT == 2) {
      if (k0 >= op.k1) {
        p = op.p2;
        ld = op.ld2;
        kb = k0 - op.k1;
      } else if (k0 + BK > op.k1) {
        across = true;
      }
    }
    const bool full = r0 + R <= rows && k0 + BK <= K && !across;
    const int tid = threadIdx.x;
    if (full && op.vec && form == kRK) {
      mode = 0;
      using V = VecN<T, VN>;
#pragma unroll
      for (int i = 0; i < NE / VN; ++i) {
        const int v = tid + NT * i;
        const int r = v / (BK / VN), kv = v % (BK / VN);
        const V x = *reinterpret_cast<const V*>(p + (r0 + r) * ld + kb + kv * VN);
#pragma unroll
        for (int e = 0; e < VN; ++e) reg[i * VN + e] = x.v[e];
      }
    } else if (full && op.vec && form == kKR) {
      mode = 1;
      using V = VecN<T, VN>;
#pragma unroll
      for (int i = 0; i < NE / VN; ++i) {
        const int v = tid + NT * i;
        const int k = v / (R / VN), rv = v % (R / VN);
        const V x = *reinterpret_cast<const V*>(p + (kb + k) * ld + r0 + rv * VN);
#pragma unroll
        for (int e = 0; e < VN; ++e) reg[i * VN + e] = x.v[e];
      }
    } else {
      mode = 2;
#pragma unroll
      for (int j = 0; j < NE; ++j) {
        const int idx = tid + NT * j;
        const int k = idx / R, r = idx % R;
        const int64_t gr = r0 + r, gk = k0 + k;
        reg[j] = (gr < rows && gk < K) ? elem<EXT>(op, form, gr, gk) : T(0);
      }
    }
  }

  __device__ __forceinline__ void store(T* lds) const {
    const int tid = threadIdx.x;
    if (mode == 0) {
#pragma unroll
      for (int i = 0; i < NE / VN; ++i) {
        const int v = tid + NT * i;
        const int r = v / (BK / VN), kv = v % (BK / VN);
#pragma unroll
        for (int e = 0; e < VN; ++e) lds[(kv * VN + e) * LDS_STRIDE + r] = reg[i * VN + e];
      }
    } else if (mode == 1) {
      using V = VecN<T, VN>;
#pragma unroll
      for (int i = 0; i < NE / VN; ++i) {
        const int v = tid + NT * i;
        const int k = v / (R / VN), rv = v % (R / VN);
        V x;
#pragma unroll
        for (int e = 0; e < VN; ++e) x.v[e] = reg[i * VN + e];
        *reinterpret_cast<V*>(lds + k * LDS_STRIDE + rv * VN) = x;
      }
    } else {
#pragma unroll
      for (int j = 0; j < NE; ++j) {
        const int idx = tid + NT * j;
        const int k = idx / R, r = idx % R;
        lds[k * LDS_STRIDE + r] = reg[j];
      }
    }
  }
};

// Tile parameters.
template <class T, int BM_, int BN_, int BK_, int WM_, int WN_>
struct Cfg {
  using type = T;
  static constexpr int BM = BM_, BN = BN_, BK = BK_, WM = WM_, WN = WN_;
  static constexpr int WAVES_M = BM / WM, WAVES_N = BN / WN;
  static constexpr int NT = 64 * WAVES_M * WAVES_N;
  static constexpr int SM = WM / 16, SN = WN / 16;
  static constexpr int VN = 16 / sizeof(T);
  static constexpr int EA = cmin(VN, SM), EB = cmin(VN, SN);
  static constexpr int RA = SM / EA, RB = SN / EB;
  using LA = TileLoader<T, BM, BK, NT>;
  using LB = TileLoader<T, BN, BK, NT>;
  static constexpr int LDS_A = BK * LA::LDS_STRIDE, LDS_B = BK * LB::LDS_STRIDE;
  static constexpr size_t LDS_BYTES = sizeof(T) * 2 * (LDS_A + LDS_B);
  static_assert(BK % 4 == 0, "BK is a multiple of the MFMA K (4)");
};

// The part of [kb, ke) a triangular A operand contributes to rows
// [m0, m0 + BM) of the product.  EXT 1: where it can be nonzero (k <= r,
// lower: up to m0 + BM; k >= r, upper: from m0).  EXT 3: strictly inside the
// triangle for every row (lower: k < m0; upper: k >= m0 + BM).  EXT 4: the
// diagonal block [m0, m0 + BM).  3 and 4 together cover 1's range once.
template <int EXT, class T>
__device__ __forceinline__ void tri_k_window(const Operand<T>& A, int64_t m0, int BM, int64_t& kb, int64_t& ke) {
  const bool upper = (A.form - kTri) & 2;
  int64_t lo = 0, hi = INT64_MAX;
  if (EXT == 1) {
    if (upper) lo = m0;
    else hi = m0 + BM;
  } else if (EXT == 3) {
    if (upper) lo = m0 + BM;
    else hi = m0;
  } else {
    lo = m0;
    hi = m0 + BM;
  }
  if (kb < lo) kb = lo;
  if (ke > hi) ke = hi;
}

// Accumulates one block tile of  sum_k X_A(m, k) X_B(n, k)  for k in [kb, ke)
// into acc; lds points at LDS_BYTES of shared memory.
template <class C, int EXT = 0>
__device__ __forceinline__ void tile_mma(const Operand<typename C::type>& A, const Operand<typename C::type>& B,
                                         int64_t M, int64_t N, int64_t K, int64_t m0, int64_t n0, int64_t kb,
                                         int64_t ke, typename Mfma<typename C::type>::acc (&acc)[C::SM][C::SN],
                                         typename C::type* lds) {
  using T = typename C::type;
  using MF = Mfma<T>;
  T* As = lds;
  T* Bs = lds + 2 * C::LDS_A;
  const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
  const int wm0 = (wave / C::WAVES_N) * C::WM, wn0 = (wave % C::WAVES_N) * C::WN;
  const int lm = lane % 16, lg = lane / 16;
  typename C::LA la;
  typename C::LB lb;
  const int64_t nk = (ke - kb + C::BK - 1) / C::BK;
  if (nk <= 0) return;
  // the K extent the loaders bound-check against is ke (a split's end)
  la.template load<EXT>(A, m0, kb, M, ke);
  lb.template load<EXT>(B, n0, kb, N, ke);
  la.store(As);
  lb.store(Bs);
  __syncthreads();
  for (int64_t t = 0; t < nk; ++t) {
    const int cur = (int)(t & 1);
    if (t + 1 < nk) {
      la.template load<EXT>(A, m0, kb + (t + 1) * C::BK, M, ke);
      lb.template load<EXT>(B, n0, kb + (t + 1) * C::BK, N, ke);
    }
    const T* as = As + cur * C::LDS_A;
    const T* bs = Bs + cur * C::LDS_B;
#pragma unroll
    for (int kk = 0; kk < C::BK / 4; ++kk) {
      const int k = kk * 4 + lg;
      T a[C::SM], b[C::SN];
#pragma unroll
      for (int g = 0; g < C::RA; ++g) {
        const VecN<T, C::EA> v =
            *reinterpret_cast<const VecN<T, C::EA>*>(as + k * C::LA::LDS_STRIDE + wm0 + g * 16 * C::EA + C::EA * lm);
#pragma unroll
        for (int e = 0; e < C::EA; ++e) a[g * C::EA + e] = v.v[e];
      }
#pragma unroll
      for (int g = 0; g < C::RB; ++g) {
        const VecN<T, C::EB> v =
            *reinterpret_cast<const VecN<T, C::EB>*>(bs + k * C::LB::LDS_STRIDE + wn0 + g * 16 * C::EB + C::EB * lm);
#pragma unroll
        for (int e = 0; e < C::EB; ++e) b[g * C::EB + e] = v.v[e];
      }
#pragma unroll
      for (int ta = 0; ta < C::SM; ++ta)
#pragma unroll
        for (int tb = 0; tb < C::SN; ++tb) acc[ta][tb] = MF::step(a[ta], b[tb], acc[ta][tb]);
    }
    if (t + 1 < nk) {
      la.store(As + (cur ^ 1) * C::LDS_A);
      lb.store(Bs + (cur ^ 1) * C::LDS_B);
    }
    __syncthreads();
  }
}

// C[m, n] += alpha * acc for the block tile at (m0, n0), restricted to the
// TRI triangle of C (n <= m for kLower, n >= m for kUpper).  With BETA the
// update is C = alpha * acc + beta * C instead, C not read when beta == 0
// unless READ0 (then beta * C is formed for every beta, as the reference's
// mscal + gemm does: NaN or Inf in C stay NaN with beta == 0).
template <class C, int TRI, bool BETA = false, bool READ0 = false>
__device__ __forceinline__ void tile_store(typename C::type* Cp, int64_t ldc, int cvec, int64_t M, int64_t N,
                                           int64_t m0, int64_t n0, typename C::type alpha,
                                           const typename Mfma<typename C::type>::acc (&acc)[C::SM][C::SN],
                                           typename C::type beta = 0) {
  using T = typename C::type;
  const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
  const int wm0 = (wave / C::WAVES_N) * C::WM, wn0 = (wave % C::WAVES_N) * C::WN;
  const int lm = lane % 16, lg = lane / 16;
  constexpr int E = C::EB;
#pragma unroll
  for (int ta = 0; ta < C::SM; ++ta) {
    const int ga = ta / C::EA, ea = ta % C::EA;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int64_t row = m0 + wm0 + ga * 16 * C::EA + C::EA * Mfma<T>::out_row(lg, q) + ea;
      if (row >= M) continue;
#pragma unroll
      for (int gb = 0; gb < C::RB; ++gb) {
        const int64_t col0 = n0 + wn0 + gb * 16 * E + E * lm;
        T* cp = Cp + row * ldc + col0;
        bool all_in = cvec && col0 + E <= N;
        if (TRI == kLower) all_in = all_in && col0 + E - 1 <= row;
        if (TRI == kUpper) all_in = all_in && col0 >= row;
        if (all_in) {
          VecN<T, E> v;
          if (!BETA || READ0 || beta != T(0)) v = *reinterpret_cast<VecN<T, E>*>(cp);
#pragma unroll
          for (int e = 0; e < E; ++e) {
            if (!BETA)
              v.v[e] += alpha * acc[ta][gb * E + e][q];
            else
              v.v[e] = alpha * acc[ta][gb * E + e][q] + (READ0 || beta != T(0) ? beta * v.v[e] : T(0));
          }
          *reinterpret_cast<VecN<T, E>*>(cp) = v;
        } else {
#pragma unroll
          for (int e = 0; e < E; ++e) {
            const int64_t col = col0 + e;
            bool in = col < N;
            if (TRI == kLower) in = in && col <= row;
            if (TRI == kUpper) in = in && col >= row;
            if (in) {
              if (!BETA)
                cp[e] += alpha * acc[ta][gb * E + e][q];
              else
                cp[e] = alpha * acc[ta][gb * E + e][q] + (READ0 || beta != T(0) ? beta * cp[e] : T(0));
            }
          }
        }
      }
    }
  }
}

// Workgroup ids are dealt round-robin over the 8 XCDs; this regroups the ids
// of an nblocks grid so that each XCD runs a contiguous share of the tile
// list (identity when nblocks is not a multiple of 8).
__host__ __device__ inline int64_t xcd_regroup(int64_t bid, int64_t nblocks) {
  constexpr int kXcds = 8;
  return nblocks % kXcds == 0 ? (bid % kXcds) * (nblocks / kXcds) + bid / kXcds : bid;
}

// Tiles of row tm that meet the TRI triangle (lower: n0 <= m0 + BM - 1,
// upper: n0 + BN - 1 >= m0), as the column range [first, first + count).
__host__ __device__ inline void tri_row(int tri, int64_t tm, int64_t tiles_n, int BM, int BN, int64_t& first,
                                        int64_t& count) {
  const int64_t lo = tm * BM, hi = lo + BM - 1;
  first = tri == kLower ? 0 : lo / BN;
  const int64_t last = tri == kLower ? (hi / BN < tiles_n - 1 ? hi / BN : tiles_n - 1) : tiles_n - 1;
  count = last >= first ? last - first + 1 : 0;
}

// Length of the tile list of a TRI-restricted product (list_tile's domain).
__host__ __device__ inline int64_t tri_tile_count(int tri, int64_t tiles_m, int64_t tiles_n, int BM, int BN) {
  if (tri == kFull) return tiles_m * tiles_n;
  int64_t n = 0, first, count;
  for (int64_t tm = 0; tm < tiles_m; ++tm) {
    tri_row(tri, tm, tiles_n, BM, BN, first, count);
    n += count;
  }
  return n;
}

// Position idx of the tile list -> (tile_m, tile_n).  Full products walk the
// tiles in groups of GROUP_M tile rows, column by column inside a group, so
// the tiles one XCD runs together share A row panels and B column panels in
// its L2.  Triangular updates list only the tiles that meet the triangle, row
// by row (launching the full grid and letting off-triangle workgroups exit
// leaves whole XCDs idle: with the XCD-major order their shares lie entirely
// on one side of the diagonal).  Past the end of the list tn = tiles_n, which
// the callers' bounds checks skip.
__host__ __device__ inline void list_tile(int tri, int64_t idx, int64_t tiles_m, int64_t tiles_n, int BM, int BN,
                                          int64_t& tm, int64_t& tn) {
  if (tri == kFull) {
    constexpr int kGroupM = 8;
    const int64_t per_group = (int64_t)kGroupM * tiles_n;
    const int64_t group = idx / per_group;
    const int64_t first_m = group * kGroupM;
    const int64_t gsz = tiles_m - first_m < kGroupM ? tiles_m - first_m : kGroupM;
    tm = first_m + (idx % per_group) % gsz;
    tn = (idx % per_group) / gsz;
    return;
  }
  int64_t acc = 0, first, count;
  for (tm = 0; tm < tiles_m; ++tm) {
    tri_row(tri, tm, tiles_n, BM, BN, first, count);
    if (idx < acc + count) {
      tn = first + (idx - acc);
      return;
    }
    acc += count;
  }
  tn = tiles_n;
}

// Block id of an nblocks grid that covers the whole list -> tile.
__device__ __forceinline__ void tile_coords(int64_t bid, int64_t nblocks, int64_t tiles_m, int64_t tiles_n,
                                            int64_t& tm, int64_t& tn) {
  list_tile(kFull, xcd_regroup(bid, nblocks), tiles_m, tiles_n, 1, 1, tm, tn);
}
__device__ __forceinline__ void tri_tile_coords(int tri, int64_t bid, int64_t nblocks, int64_t tiles_m,
                                                int64_t tiles_n, int BM, int BN, int64_t& tm, int64_t& tn) {
  list_tile(tri, xcd_regroup(bid, nblocks), tiles_m, tiles_n, BM, BN, tm, tn);
}

}  // namespace gemm
}  // namespace exoblas
