// Tiles of a row-major matrix for the streaming level-2 / level-3 updates
// (ger, syr, syr2, symv, mscal, trmscal) on gfx950.
//
// The matrix is cut into RB x W tiles (RB rows, W = 256 lanes x VN columns:
// one 16-byte vector per lane), and only the tiles that meet the stored part
// are launched: the full matrix, its lower (j <= i) or its upper (j >= i)
// triangle.  Tiles are listed row block by row block; a workgroup finds its
// tile from its id with the closed-form count of tiles before a row block and
// a binary search, so no tile table is needed in memory.  For the triangles W
// is a multiple K of RB and the matrix is square, so row block b meets column
// blocks [0, b / K] (lower) or [b / K, nc) (upper).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "device.hpp"

namespace exoblas {
namespace tiles {

enum Tri : int { kFull = 0, kLower = 1, kUpper = 2 };

struct TileList {
  int64_t nb;  // row blocks
  int64_t nc;  // column blocks
  int64_t K;   // W / RB (triangles)
  int tri;

  // tiles in row blocks [0, b)
  __host__ __device__ int64_t prefix(int64_t b) const {
    if (tri == kFull) return b * nc;
    const int64_t q = b / K, r = b % K;
    const int64_t below = K * (q * (q - 1) / 2) + r * q;  // sum_{i<b} floor(i / K)
    return tri == kLower ? b + below : b * nc - below;
  }
  __host__ __device__ int64_t total() const { return prefix(nb); }
  // first column block of row block b
  __host__ __device__ int64_t first(int64_t b) const { return tri == kUpper ? b / K : 0; }
  // tile u of the list -> (row block b, column block c), in closed form:
  // row blocks come in groups of K with the same tile count (lower: group g
  // has g + 1 tiles per row block, upper: nc - g), so u's group follows from
  // a quadratic (square root, then an exact integer fix-up) and the rest from
  // one 32-bit division; no search, no 64-bit division.  (Tile lists stay
  // below 2^31 entries: the launches check.)
  __host__ __device__ void tile(int64_t u, int64_t& b, int64_t& c) const {
    if (tri == kFull) {
      b = (int64_t)((uint32_t)u / (uint32_t)nc);
      c = u - b * nc;
      return;
    }
    const double Kd = (double)K;
    int64_t g;
    if (tri == kLower) {  // tiles before group g: K g (g + 1) / 2
      auto before = [&](int64_t q) { return K * (q * (q + 1) / 2); };
      g = (int64_t)((sqrt(8.0 * (double)u / Kd + 1.0) - 1.0) * 0.5);
      while (g > 0 && before(g) > u) --g;
      while (before(g + 1) <= u) ++g;
      const uint32_t o = (uint32_t)(u - before(g)), per = (uint32_t)(g + 1);
      b = g * K + o / per;
      c = o % per;
    } else {  // tiles before group g: K (g nc - g (g - 1) / 2)
      auto before = [&](int64_t q) { return K * (q * nc - q * (q - 1) / 2); };
      const double h = (double)nc + 0.5;
      const double disc = h * h - 2.0 * (double)u / Kd;
      g = (int64_t)(h - sqrt(disc > 0.0 ? disc : 0.0));
      if (g < 0) g = 0;
      while (g > 0 && before(g) > u) --g;
      while (before(g + 1) <= u) ++g;
      const uint32_t o = (uint32_t)(u - before(g)), per = (uint32_t)(nc - g);
      b = g * K + o / per;
      c = g + o % per;
    }
  }
};

template <int RB, int W>
__host__ __device__ inline TileList make_list(int tri, int64_t m, int64_t n) {
  static_assert(W % RB == 0, "W is a multiple of RB");
  TileList t;
  t.nb = ceil_div(m, RB);
  t.nc = ceil_div(n, W);
  t.K = W / RB;
  t.tri = tri;
  return t;
}

// Columns [jl, jh) of row i in the stored part.
__host__ __device__ inline void row_range(int tri, int64_t i, int64_t n, int64_t& jl, int64_t& jh) {
  jl = tri == kUpper ? i : 0;
  jh = tri == kLower ? (i + 1 < n ? i + 1 : n) : n;
}

// Arguments of update_kernel: KIND 0 scales (A *= alpha), 1 adds
// (alpha u1_i) c1_j, 2 also adds (alpha u2_i) c2_j, 3 stores zeros (A not
// read).
template <class T>
struct UpdateArgs {
  T alpha;
  const T* u1;
  int64_t iu1;
  const T* c1;
  int64_t ic1;
  const T* u2;
  int64_t iu2;
  const T* c2;
  int64_t ic2;
};

// A[i, j] <- update(A[i, j]) over the TRI part of an m x n matrix, one RB x W
// tile per workgroup.  A lane owns VN consecutive columns of its tile and
// updates the tile's rows that meet them in one round: every load issued
// before the first store.  (Measured on MI355X for read-modify-write streams:
// few loads per lane in many small workgroups beat deep per-lane pipelines,
// 6.55-6.59 TB/s for 2- or 4-row tiles, 6.4 for 8 rows, 5.1-5.6 for 32-64
// rows with two rounds in flight.)  Misaligned matrices go element by
// element.
template <class T, int KIND, int TRI, bool VEC, int RB>
__global__ __launch_bounds__(256) void update_kernel(int64_t m, int64_t n, T* A, int64_t lda, UpdateArgs<T> g,
                                                     TileList tl) {
  using V = typename Vec16<T>::type;
  constexpr int VN = Vec16<T>::n, W = 256 * VN;
  int64_t b, c;
  tl.tile(blockIdx.x, b, c);
  const int64_t j0 = c * W + (int64_t)threadIdx.x * VN;
  if (j0 >= n) return;
  int64_t i0 = b * RB, i1 = i0 + RB < m ? i0 + RB : m;
  if (TRI == kLower && i0 < j0) i0 = j0;            // rows i >= j0 meet columns j <= i
  if (TRI == kUpper && i1 > j0 + VN) i1 = j0 + VN;  // rows i <= j0 + VN - 1 meet columns j >= i
  if (i0 >= i1) return;
  T cv1[VN], cv2[VN];
#pragma unroll
  for (int k = 0; k < VN; ++k) {
    const int64_t j = j0 + k;
    cv1[k] = ((KIND == 1 || KIND == 2) && j < n) ? g.c1[j * g.ic1] : T(0);
    cv2[k] = (KIND == 2 && j < n) ? g.c2[j * g.ic2] : T(0);
  }
  auto upd = [&](T& a, T r1, T r2, int k) {
    if (KIND == 0) a = a * g.alpha;
    if (KIND == 3) a = T(0);
    if (KIND == 1) a += r1 * cv1[k];
    if (KIND == 2) a += r1 * cv1[k] + r2 * cv2[k];
  };
  auto rows = [&](int64_t i, T& r1, T& r2) {
    r1 = (KIND == 1 || KIND == 2) ? g.alpha * g.u1[i * g.iu1] : T(0);
    r2 = KIND == 2 ? g.alpha * g.u2[i * g.iu2] : T(0);
  };
  // element by element over rows [e0, e1) (masked to the stored part)
  auto elementwise = [&](int64_t e0, int64_t e1) {
    for (int64_t i = e0; i < e1; ++i) {
      T r1, r2;
      rows(i, r1, r2);
      T* row = A + i * lda;
#pragma unroll
      for (int k = 0; k < VN; ++k) {
        const int64_t j = j0 + k;
        bool in = j < n;
        if (TRI == kLower) in = in && j <= i;
        if (TRI == kUpper) in = in && j >= i;
        if (in) upd(row[j], r1, r2, k);
      }
    }
  };
  if (!VEC || j0 + VN > n) {
    elementwise(i0, i1);
    return;
  }
  // Whole 16-byte vectors: every row's load is issued before the first
  // store; a row where only some of the lane's columns are stored (the
  // diagonal) stores those elements one by one, without a serial chain.
  V v[RB];
  T r1[RB], r2[RB];
#pragma unroll
  for (int u = 0; u < RB; ++u)
    if (i0 + u < i1) {
      if (KIND != 3) v[u] = ld_stream<T>(A + (i0 + u) * lda + j0);
      rows(i0 + u, r1[u], r2[u]);
    }
#pragma unroll
  for (int u = 0; u < RB; ++u) {
    const int64_t i = i0 + u;
    if (i >= i1) continue;
#pragma unroll
    for (int k = 0; k < VN; ++k) upd(vget<T>(v[u], k), r1[u], r2[u], k);
    const bool whole = TRI == kLower ? j0 + VN - 1 <= i : TRI == kUpper ? j0 >= i : true;
    if (whole) {
      st_stream<T>(A + i * lda + j0, v[u]);
    } else {
#pragma unroll
      for (int k = 0; k < VN; ++k)
        if (TRI == kLower ? j0 + k <= i : j0 + k >= i) A[i * lda + j0 + k] = vget<T>(v[u], k);
    }
  }
}

// Launches update_kernel over the TRI part of an m x n matrix.
template <class T, int KIND, int TRI>
void launch_update(hipStream_t stream, int64_t m, int64_t n, T* A, int64_t lda, const UpdateArgs<T>& g) {
  if (m <= 0 || n <= 0) return;
  constexpr int RB = 4, W = 256 * Vec16<T>::n;
  const TileList tl = make_list<RB, W>(TRI, m, n);
  const int64_t tiles = tl.total();
  if (tiles > 0x7fffffff) throw Error(kInvalidSize, "matrix too large for one launch");
  const bool vec = (lda % Vec16<T>::n) == 0 && aligned16(A);
  if (vec)
    hipLaunchKernelGGL((update_kernel<T, KIND, TRI, true, RB>), dim3((unsigned)tiles), dim3(256), 0, stream, m, n, A,
                       lda, g, tl);
  else
    hipLaunchKernelGGL((update_kernel<T, KIND, TRI, false, RB>), dim3((unsigned)tiles), dim3(256), 0, stream, m, n, A,
                       lda, g, tl);
  EXO_LAUNCH_CHECK();
}

}  // namespace tiles
}  // namespace exoblas
