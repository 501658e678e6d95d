// Level-1 kernels for gfx950: streaming elementwise updates and reductions.
//
// All of these are HBM-bound.  The unit-stride path moves 16 bytes per lane
// per access (global_load_dwordx4), unrolled so each lane keeps several loads
// in flight; a wave64 touches 1 KiB contiguous per access.  Strided and
// misaligned windows go element by element with 8 elements per lane in
// flight.  Reductions are one
// launch: per-workgroup partial sums (wave shuffle + LDS) published as tagged
// words (tagged.hpp), summed in a fixed order by the workgroup that finishes
// last, so results are deterministic.
#include <type_traits>

#include "device.hpp"
#include "ops.hpp"
#include "tagged.hpp"

namespace exoblas {
namespace {

constexpr int kNT = 256;

// ------------------------------------------------------------ elementwise

// U 16-byte vectors per lane per round (lane l of the workgroup takes
// vectors l, l + NT, ..., l + (U-1) NT of the workgroup's NT*U), all loads
// issued before the first store; the grid sweeps the vectors (one round per
// workgroup when the grid covers them all).
// x and y: the 16-byte aligned starts; `head` (< VN) elements before them
// belong to the range too (windows whose start is off alignment by the same
// amount in both vectors), done by the last workgroup with the < VN trailing
// ones.
template <class T, class Op, int NT, int U>
__global__ __launch_bounds__(NT) void ew_vec_kernel(int64_t nvec, int64_t n, T* __restrict__ x, T* __restrict__ y,
                                                    Op op, int head) {
  using V = typename Vec16<T>::type;
  constexpr int VN = Vec16<T>::n;
  for (int64_t i0 = (int64_t)blockIdx.x * NT * U + threadIdx.x; i0 < nvec; i0 += (int64_t)gridDim.x * NT * U) {
    V xv[U], yv[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int64_t i = i0 + u * NT;
      if (U == 1 || i < nvec) {
        if (Op::rx) xv[u] = ld_stream<T>(x + i * VN);
        if (Op::ry) yv[u] = ld_stream<T>(y + i * VN);
      }
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int64_t i = i0 + u * NT;
      if (U == 1 || i < nvec) {
#pragma unroll
        for (int k = 0; k < VN; ++k) op(vget<T>(xv[u], k), vget<T>(yv[u], k));
        if (Op::wx) st_stream<T>(x + i * VN, xv[u]);
        if (Op::wy) st_stream<T>(y + i * VN, yv[u]);
      }
    }
  }
  if (blockIdx.x == gridDim.x - 1) {  // the < VN trailing elements, then the head
    int64_t t = 0;  // a tail element at t >= 0 or a head element at t < 0
    bool mine = false;
    if (threadIdx.x < VN) {
      t = nvec * VN + threadIdx.x;
      mine = t < n;
    } else if (threadIdx.x >= 64 && threadIdx.x < 64 + head) {
      t = (int64_t)threadIdx.x - 64 - head;
      mine = true;
    }
    if (mine) {
      T xs = Op::rx ? x[t] : T(0);
      T ys = Op::ry ? y[t] : T(0);
      op(xs, ys);
      if (Op::wx) x[t] = xs;
      if (Op::wy) y[t] = ys;
    }
  }
}

// Strided or misaligned windows: element by element, U elements per lane in
// flight (lane l of the workgroup takes elements l, l + NT, ... of its NT*U),
// every load of a round issued before the first store, one round per
// workgroup.  Consecutive lanes take consecutive elements, so a wave's access
// spans 64 * |inc| elements: whole cache lines for |inc| = 1, and for larger
// increments the lines the elements lie in (the bytes the hardware must move
// anyway).
template <class T, class Op, int NT, int U>
__global__ __launch_bounds__(NT) void ew_strided_kernel(int64_t n, T* __restrict__ x, int64_t incx,
                                                        T* __restrict__ y, int64_t incy, Op op) {
  for (int64_t b = (int64_t)blockIdx.x * NT * U + threadIdx.x; b < n; b += (int64_t)gridDim.x * NT * U) {
    T xs[U], ys[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int64_t i = b + u * NT;
      xs[u] = (Op::rx && i < n) ? x[i * incx] : T(0);
      ys[u] = (Op::ry && i < n) ? y[i * incy] : T(0);
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int64_t i = b + u * NT;
      if (i < n) {
        op(xs[u], ys[u]);
        if (Op::wx) x[i * incx] = xs[u];
        if (Op::wy) y[i * incy] = ys[u];
      }
    }
  }
}

template <class T> struct AxpyOp;
template <class T> struct Axpy1Op;

template <class T, class Op>
void launch_ew(Ctx& c, int64_t n, T* x, int64_t incx, T* y, int64_t incy, Op op) {
  if (n <= 0) return;
  constexpr int VN = Vec16<T>::n;
  // Every vector the op touches walks down with unit stride: element i of
  // each is at position n - 1 - i of the range ending at element 0, so the
  // op pairs the same elements as a unit-stride call on those ranges.
  constexpr bool ux = Op::rx || Op::wx, uy = Op::ry || Op::wy;
  if ((!ux || incx == -1) && (!uy || incy == -1)) {
    if (ux) x -= n - 1, incx = 1;
    if (uy) y -= n - 1, incy = 1;
  }
  // unit-stride windows off 16-byte alignment by the same amount: the
  // aligned part on the vector path, the head elements before it with the
  // tail
  int head = 0;
  if ((!ux || incx == 1) && (!uy || incy == 1)) {
    const uintptr_t ox = ux ? reinterpret_cast<uintptr_t>(x) % 16 : 0, oy = uy ? reinterpret_cast<uintptr_t>(y) % 16 : 0;
    const uintptr_t off = ux ? ox : oy;
    if ((!ux || ox == off) && (!uy || oy == off) && off % sizeof(T) == 0 && off != 0) {
      const int64_t h = (int64_t)((16 - off) / sizeof(T));
      if (h < n) {
        head = (int)h;
        if (ux) x += h;
        if (uy) y += h;
        n -= h;
      }
    }
  }
  const bool unit = (!(Op::rx || Op::wx) || (incx == 1 && aligned16(x))) &&
                    (!(Op::ry || Op::wy) || (incy == 1 && aligned16(y)));
  if (!unit && head) {  // not aligned after all: undo the peel
    if (ux) x -= head;
    if (uy) y -= head;
    n += head;
    head = 0;
  }
  if (unit) {
    // One round per workgroup of 256 lanes with four 16-byte vectors per lane
    // in flight, one for FP32 axpy: measured on MI355X inside the bandwidth
    // set, with the same builds on one box (2^26 elements, operands
    // alternating between two sets): saxpy 85.5 % of HBM peak with one
    // vector per lane against 80.5 % with four, while scopy (71.5 % against
    // 77.2 %), sswap (71.5 / 76.9 %) and daxpy (76.3 / 79.9 %) lose with one.
    constexpr bool kAxpy = std::is_same<Op, AxpyOp<T>>::value || std::is_same<Op, Axpy1Op<T>>::value;
    constexpr int NT = kNT, U = (kAxpy && sizeof(T) == 4) ? 1 : 4;
    const int64_t nvec = n / VN;
    const int64_t blocks = std::max<int64_t>(1, ceil_div(nvec, (int64_t)NT * U));
    hipLaunchKernelGGL((ew_vec_kernel<T, Op, NT, U>), dim3((unsigned)blocks), dim3(NT), 0, c.stream, nvec, n, x, y,
                       op, head);
  } else {
    constexpr int U = 8;
    const int64_t blocks = std::min<int64_t>(ceil_div(n, (int64_t)kNT * U), 0x7fffffff);
    hipLaunchKernelGGL((ew_strided_kernel<T, Op, kNT, U>), dim3((unsigned)blocks), dim3(kNT), 0, c.stream, n, x,
                       incx, y, incy, op);
  }
  EXO_LAUNCH_CHECK();
}

// Elementwise functors: flags say which of (x, y) are read / written.
template <class T> struct AxpyOp {
  static constexpr bool rx = true, ry = true, wx = false, wy = true;
  T alpha;
  __device__ void operator()(T& x, T& y) const { y += alpha * x; }
};
template <class T> struct Axpy1Op {
  static constexpr bool rx = true, ry = true, wx = false, wy = true;
  __device__ void operator()(T& x, T& y) const { y += x; }
};
template <class T> struct CopyOp {
  static constexpr bool rx = true, ry = false, wx = false, wy = true;
  __device__ void operator()(T& x, T& y) const { y = x; }
};
template <class T> struct SwapOp {
  static constexpr bool rx = true, ry = true, wx = true, wy = true;
  __device__ void operator()(T& x, T& y) const {
    T t = x;
    x = y;
    y = t;
  }
};
template <class T> struct RotOp {
  static constexpr bool rx = true, ry = true, wx = true, wy = true;
  T c, s;
  __device__ void operator()(T& x, T& y) const {
    const T xr = x;
    x = c * xr + s * y;
    y = -s * xr + c * y;
  }
};
// rotm, src/level1/rotm.py: H = {H00, H01, H10, H11}
template <class T> struct RotmNegOneOp {
  static constexpr bool rx = true, ry = true, wx = true, wy = true;
  T h00, h01, h10, h11;
  __device__ void operator()(T& x, T& y) const {
    const T xr = x;
    x = h00 * xr + h01 * y;
    y = h10 * xr + h11 * y;
  }
};
template <class T> struct RotmZeroOp {
  static constexpr bool rx = true, ry = true, wx = true, wy = true;
  T h01, h10;
  __device__ void operator()(T& x, T& y) const {
    const T xr = x;
    x = xr + h01 * y;
    y = h10 * xr + y;
  }
};
template <class T> struct RotmOneOp {
  static constexpr bool rx = true, ry = true, wx = true, wy = true;
  T h00, h11;
  __device__ void operator()(T& x, T& y) const {
    const T xr = x;
    x = h00 * xr + y;
    y = -xr + h11 * y;
  }
};
template <class T> struct ScalOp {
  static constexpr bool rx = true, ry = false, wx = true, wy = false;
  T alpha;
  __device__ void operator()(T& x, T&) const { x = alpha * x; }
};
template <class T> struct ZeroOp {
  static constexpr bool rx = false, ry = false, wx = true, wy = false;
  __device__ void operator()(T& x, T&) const { x = T(0); }
};

// ------------------------------------------------------------ reductions

template <class T> struct AbsTerm {
  using Acc = T;
  __device__ Acc operator()(T x, T) const { return x > T(0) ? x : -x; }
  static constexpr bool ry = false;
};
template <class T> struct DotTerm {
  using Acc = T;
  __device__ Acc operator()(T x, T y) const { return x * y; }
  static constexpr bool ry = true;
};
struct DsdotTerm {
  using Acc = double;
  __device__ Acc operator()(float x, float y) const { return (double)x * (double)y; }
  static constexpr bool ry = true;
};

constexpr int kReduceFlagBase = 2;  // the 8-byte aligned partials from [2]

template <class Acc, class Out> struct Finish {
  Acc add;  // sdsdot's sb
  __device__ Out operator()(Acc s) const { return (Out)(s + add); }
};

// One launch: every workgroup reduces its share and publishes its partial
// as tagged words; the workgroup that takes the last ticket waits for all
// partials and sums them in a fixed order (thread t takes partials t,
// t + kNT, ..., then the workgroup tree), so the result does not depend on
// which workgroup finishes last.  *ticket: the stream's first counter; the
// marker comes from the stream's launch generation (tagged.hpp), so the
// launch can be replayed from a hipGraph; flags: [2 ..] the tagged
// partials; err: the stream's timeout word.
template <class T, class Term, class Out, int U>
__global__ __launch_bounds__(kNT) void reduce_kernel(int64_t n, const T* __restrict__ x, int64_t incx,
                                                     const T* __restrict__ y, int64_t incy, int vec, int head,
                                                     unsigned* ticket, unsigned* gen, unsigned* flags, unsigned* err,
                                                     Out* result,
                                                     Finish<typename Term::Acc, Out> fin, Term term) {
  using Acc = typename Term::Acc;
  using V = typename Vec16<T>::type;
  constexpr int VN = Vec16<T>::n;
  __shared__ Acc red[kNT / kWave];
  // The generation is loaded now and stored for the workgroup only after the
  // sweep, so no wave waits for it before its streaming loads; thread 0 uses
  // it (publication) before taking its ticket.
  __shared__ unsigned s_marker;
  const unsigned mk = threadIdx.x == 0 ? tagged::read_generation(gen) : 0u;
  Acc acc[U];
#pragma unroll
  for (int u = 0; u < U; ++u) acc[u] = Acc(0);

  if (vec) {
    // The grid sweeps the vector front to back: in round j workgroup b reads
    // the kNT*U vectors starting at (j * gridDim.x + b) * kNT*U, so the whole
    // chip streams one contiguous region at a time.
    const int64_t nvec = n / VN;
    const int64_t span = (int64_t)kNT * U;
    int64_t base = (int64_t)blockIdx.x * span;
    for (; base + span <= nvec; base += (int64_t)gridDim.x * span) {
      V xv[U], yv[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        xv[u] = ld_stream<T>(x + (base + u * kNT + threadIdx.x) * VN);
        if (Term::ry) yv[u] = ld_stream<T>(y + (base + u * kNT + threadIdx.x) * VN);
      }
#pragma unroll
      for (int u = 0; u < U; ++u)
#pragma unroll
        for (int k = 0; k < VN; ++k) acc[u] += term(vget<T>(xv[u], k), Term::ry ? vget<T>(yv[u], k) : T(0));
    }
    if (base < nvec) {  // the partial last round
      for (int64_t i = base + threadIdx.x; i < nvec; i += kNT) {
        const V xv = reinterpret_cast<const V*>(x)[i];
        V yv;
        if (Term::ry) yv = reinterpret_cast<const V*>(y)[i];
#pragma unroll
        for (int k = 0; k < VN; ++k) acc[0] += term(vget<T>(xv, k), Term::ry ? vget<T>(yv, k) : T(0));
      }
    }
    if (blockIdx.x == 0) {  // the < VN trailing elements and the `head` ones before x
      if (threadIdx.x < VN) {
        const int64_t t = nvec * VN + threadIdx.x;
        if (t < n) acc[1 % U] += term(x[t], Term::ry ? y[t] : T(0));
      } else if (threadIdx.x >= 64 && threadIdx.x < 64 + head) {
        const int64_t t = (int64_t)threadIdx.x - 64 - head;
        acc[1 % U] += term(x[t], Term::ry ? y[t] : T(0));
      }
    }
  } else {
    // strided or misaligned: element by element, 8 per lane in flight
    // (consecutive lanes on consecutive elements), the grid sweeping rounds
    constexpr int E = 8;
    const int64_t span = (int64_t)kNT * E;
    for (int64_t base = (int64_t)blockIdx.x * span + threadIdx.x; base < n; base += (int64_t)gridDim.x * span) {
      T xs[E], ys[E];
#pragma unroll
      for (int e = 0; e < E; ++e) {
        const int64_t i = base + e * kNT;
        xs[e] = i < n ? x[i * incx] : T(0);
        ys[e] = (Term::ry && i < n) ? y[i * incy] : T(0);
      }
#pragma unroll
      for (int e = 0; e < E; ++e)
        if (base + e * kNT < n) acc[e % U] += term(xs[e], ys[e]);
    }
  }
  Acc s = acc[0];
#pragma unroll
  for (int u = 1; u < U; ++u) s += acc[u];
  if (threadIdx.x == 0) s_marker = mk;
  s = block_sum<kNT>(s, red);
  unsigned* pub = flags + kReduceFlagBase;
  const unsigned marker = s_marker;  // written before block_sum's barriers
  __shared__ int s_last;
  if (threadIdx.x == 0) {
    tagged::publish<Acc>(pub, blockIdx.x, s, marker);
    s_last = tagged::last_ticket(ticket, gridDim.x);
  }
  __syncthreads();
  if (!s_last) return;
  Acc t = Acc(0);
  for (int i = threadIdx.x; i < (int)gridDim.x; i += 4 * kNT) {
    Acc v[4];
    tagged::read_n<Acc, 4>(pub, i, kNT, (int)ceil_div((int)gridDim.x - i, kNT), v, marker, err);
#pragma unroll
    for (int k = 0; k < 4; ++k)
      if (i + k * kNT < (int)gridDim.x) t += v[k];
  }
  t = block_sum<kNT>(t, red);
  if (threadIdx.x == 0) {
    *result = fin(t);
    tagged::advance_generation(gen);  // every workgroup has read it: all took tickets
  }
}

// Launch shape of a reduction over n elements on a device with `cus` CUs:
// the (possibly re-based) operands, whether the 16-byte path applies and
// the grid.  Two workgroups per CU, each lane with U = 2 (two streams) or 4
// (one stream) streaming loads in flight per round: the best of a sweep over
// block size, unroll and grid on MI355X (7.1 TB/s for a 2^26 dot).
template <class T>
struct ReducePlan {
  const T* x;
  int64_t incx;
  const T* y;
  int64_t incy;
  bool vec;
  int64_t blocks;
  int64_t n;  // elements from x on (n minus head)
  int head;   // elements before x and y (unit-stride windows peeled to 16-byte alignment)
};

template <class T, class Term>
ReducePlan<T> plan_reduce(int64_t n, const T* x, int64_t incx, const T* y, int64_t incy, int cus) {
  constexpr int VN = Vec16<T>::n;
  constexpr int U = Term::ry ? 2 : 4;
  // both vectors walking down with unit stride: the same products over the
  // unit-stride ranges ending at element 0 (summed in another order)
  if (incx == -1 && (!Term::ry || incy == -1) && n > 0) {
    x -= n - 1;
    incx = 1;
    if (Term::ry) y -= n - 1, incy = 1;
  }
  // unit-stride windows off 16-byte alignment by the same amount: peel the
  // elements before the aligned start (the kernel adds them)
  int head = 0;
  if (incx == 1 && (!Term::ry || incy == 1)) {
    const uintptr_t ox = reinterpret_cast<uintptr_t>(x) % 16;
    const uintptr_t oy = Term::ry ? reinterpret_cast<uintptr_t>(y) % 16 : ox;
    if (ox == oy && ox != 0 && ox % sizeof(T) == 0) {
      const int64_t h = (int64_t)((16 - ox) / sizeof(T));
      if (h < n) {
        head = (int)h;
        x += h;
        if (Term::ry) y += h;
        n -= h;
      }
    }
  }
  ReducePlan<T> p;
  p.x = x;
  p.incx = incx;
  p.y = y;
  p.incy = incy;
  p.n = n;
  p.head = head;
  p.vec = incx == 1 && aligned16(x) && (!Term::ry || (incy == 1 && aligned16(y)));
  if (p.vec)
    p.blocks = std::min<int64_t>(std::max<int64_t>(1, ceil_div(n, (int64_t)kNT * U * VN)), (int64_t)cus * 2);
  else  // element by element, 8 per lane: four workgroups per CU
    p.blocks = std::min<int64_t>(std::max<int64_t>(1, ceil_div(n, (int64_t)kNT * 8)), (int64_t)cus * 4);
  return p;
}

// Enqueues the one-launch reduction: ticket and gen are the launch's ticket
// counter and launch generation, flags + kReduceFlagBase (8-byte aligned)
// holds the tagged partials, err is the timeout word.
template <class T, class Term, class Out>
void enqueue_reduce(hipStream_t stream, const ReducePlan<T>& p, int64_t n, unsigned* ticket, unsigned* gen,
                    unsigned* flags, unsigned* err, Out* result, typename Term::Acc add) {
  using Acc = typename Term::Acc;
  constexpr int U = Term::ry ? 2 : 4;
  hipLaunchKernelGGL((reduce_kernel<T, Term, Out, U>), dim3((unsigned)p.blocks), dim3(kNT), 0, stream, p.n, p.x,
                     p.incx, p.y, p.incy, p.vec ? 1 : 0, p.head, ticket, gen, flags, err, result,
                     Finish<Acc, Out>{add}, Term{});
}

template <class T, class Term, class Out>
void launch_reduce(Ctx& c, int64_t n, const T* x, int64_t incx, const T* y, int64_t incy, Out* result,
                   typename Term::Acc add) {
  using Acc = typename Term::Acc;
  const ReducePlan<T> p = plan_reduce<T, Term>(n, x, incx, y, incy, c.num_cus);
  unsigned epoch = 0;  // (the host epoch is not used: the marker comes from the launch generation)
  unsigned* flags = c.epoch_flags((size_t)(kReduceFlagBase + tagged::words<Acc>(p.blocks)), &epoch);
  enqueue_reduce<T, Term, Out>(c.stream, p, n, c.counters(), c.counters() + Ctx::kGeneration, flags,
                               c.timeout_word(), result, add);
  EXO_LAUNCH_CHECK();
}

__global__ __launch_bounds__(kNT) void fill_zero_kernel(uint32_t* p, int64_t words) {
  for (int64_t i = (int64_t)blockIdx.x * kNT + threadIdx.x; i < words; i += (int64_t)gridDim.x * kNT) p[i] = 0u;
}

__global__ void store_zero_kernel(char* p, int bytes) {
  if (threadIdx.x < bytes) p[threadIdx.x] = 0;
}

template <class Out>
void write_zero(Ctx& c, Out* result) {
  hipLaunchKernelGGL(store_zero_kernel, dim3(1), dim3(64), 0, c.stream, reinterpret_cast<char*>(result),
                     (int)sizeof(Out));
  EXO_LAUNCH_CHECK();
}

}  // namespace

void fill_zero(hipStream_t stream, void* p, size_t bytes) {
  if (bytes == 0) return;
  if (bytes % 4 != 0 || (reinterpret_cast<uintptr_t>(p) & 3) != 0)
    throw Error(kInvalidSize, "fill_zero: whole, aligned 32-bit words only");
  const int64_t words = (int64_t)(bytes / 4);
  const unsigned grid = (unsigned)std::min<int64_t>(ceil_div(words, kNT), 1024);
  hipLaunchKernelGGL(fill_zero_kernel, dim3(grid), dim3(kNT), 0, stream, static_cast<uint32_t*>(p), words);
  EXO_LAUNCH_CHECK();
}

namespace ops {

// src/level1/asum.py:11-15
template <class T>
void asum(Ctx& c, int64_t n, const T* x, int64_t incx, T* result) {
  if (n <= 0) return write_zero(c, result);
  launch_reduce<T, AbsTerm<T>, T>(c, n, x, incx, nullptr, 0, result, T(0));
}

// src/level1/axpy.py:9-12
template <class T>
void axpy(Ctx& c, int64_t n, T alpha, const T* x, int64_t incx, T* y, int64_t incy) {
  launch_ew(c, n, const_cast<T*>(x), incx, y, incy, AxpyOp<T>{alpha});
}

// src/level1/axpy.py:15-18
template <class T>
void axpy_alpha_1(Ctx& c, int64_t n, const T* x, int64_t incx, T* y, int64_t incy) {
  launch_ew(c, n, const_cast<T*>(x), incx, y, incy, Axpy1Op<T>{});
}

// src/level1/dot.py:9-13
template <class T>
void dot(Ctx& c, int64_t n, const T* x, int64_t incx, const T* y, int64_t incy, T* result) {
  if (n <= 0) return write_zero(c, result);
  launch_reduce<T, DotTerm<T>, T>(c, n, x, incx, y, incy, result, T(0));
}

// src/level1/dsdot.py:11-21
void dsdot(Ctx& c, int64_t n, const float* x, int64_t incx, const float* y, int64_t incy, double* result) {
  if (n <= 0) return write_zero(c, result);
  launch_reduce<float, DsdotTerm, double>(c, n, x, incx, y, incy, result, 0.0);
}

// src/level1/dsdot.py:24-36: FP64 dot, + sb in FP64, one rounding to FP32
void sdsdot(Ctx& c, int64_t n, float sb, const float* x, int64_t incx, const float* y, int64_t incy,
            float* result) {
  if (n <= 0) {
    // d_dot = 0, d_result = sb
    launch_reduce<float, DsdotTerm, float>(c, 0, x, 1, y, 1, result, (double)sb);
    return;
  }
  launch_reduce<float, DsdotTerm, float>(c, n, x, incx, y, incy, result, (double)sb);
}

// src/level1/exo_copy.py:9-12 (bit-exact data movement)
template <class T>
void copy(Ctx& c, int64_t n, const T* x, int64_t incx, T* y, int64_t incy) {
  launch_ew(c, n, const_cast<T*>(x), incx, y, incy, CopyOp<T>{});
}

// src/level1/swap.py:9-15 (bit-exact data movement)
template <class T>
void swap(Ctx& c, int64_t n, T* x, int64_t incx, T* y, int64_t incy) {
  launch_ew(c, n, x, incx, y, incy, SwapOp<T>{});
}

// src/level1/rot.py:9-15
template <class T>
void rot(Ctx& c, int64_t n, T* x, int64_t incx, T* y, int64_t incy, T cc, T s) {
  launch_ew(c, n, x, incx, y, incy, RotOp<T>{cc, s});
}

// src/level1/rotm.py:9-15
template <class T>
void rotm_flag_neg_one(Ctx& c, int64_t n, T* x, int64_t incx, T* y, int64_t incy, const T* H) {
  launch_ew(c, n, x, incx, y, incy, RotmNegOneOp<T>{H[0], H[1], H[2], H[3]});
}

// src/level1/rotm.py:18-24
template <class T>
void rotm_flag_zero(Ctx& c, int64_t n, T* x, int64_t incx, T* y, int64_t incy, const T* H) {
  launch_ew(c, n, x, incx, y, incy, RotmZeroOp<T>{H[1], H[2]});
}

// src/level1/rotm.py:27-33
template <class T>
void rotm_flag_one(Ctx& c, int64_t n, T* x, int64_t incx, T* y, int64_t incy, const T* H) {
  launch_ew(c, n, x, incx, y, incy, RotmOneOp<T>{H[0], H[3]});
}

// src/level1/scal.py:9-12
template <class T>
void scal(Ctx& c, int64_t n, T alpha, T* x, int64_t incx) {
  launch_ew(c, n, x, incx, static_cast<T*>(nullptr), 0, ScalOp<T>{alpha});
}

// src/level1/scal.py:15-19
template <class T>
void scal_alpha_0(Ctx& c, int64_t n, T* x, int64_t incx) {
  launch_ew(c, n, x, incx, static_cast<T*>(nullptr), 0, ZeroOp<T>{});
}

#define EXO_L1_INST(T)                                                                           \
  template void asum<T>(Ctx&, int64_t, const T*, int64_t, T*);                                  \
  template void axpy<T>(Ctx&, int64_t, T, const T*, int64_t, T*, int64_t);                      \
  template void axpy_alpha_1<T>(Ctx&, int64_t, const T*, int64_t, T*, int64_t);                 \
  template void dot<T>(Ctx&, int64_t, const T*, int64_t, const T*, int64_t, T*);                \
  template void copy<T>(Ctx&, int64_t, const T*, int64_t, T*, int64_t);                         \
  template void swap<T>(Ctx&, int64_t, T*, int64_t, T*, int64_t);                               \
  template void rot<T>(Ctx&, int64_t, T*, int64_t, T*, int64_t, T, T);                          \
  template void rotm_flag_neg_one<T>(Ctx&, int64_t, T*, int64_t, T*, int64_t, const T*);        \
  template void rotm_flag_zero<T>(Ctx&, int64_t, T*, int64_t, T*, int64_t, const T*);           \
  template void rotm_flag_one<T>(Ctx&, int64_t, T*, int64_t, T*, int64_t, const T*);            \
  template void scal<T>(Ctx&, int64_t, T, T*, int64_t);                                         \
  template void scal_alpha_0<T>(Ctx&, int64_t, T*, int64_t);

EXO_L1_INST(float)
EXO_L1_INST(double)

}  // namespace ops
}  // namespace exoblas

// ------------------------------------------------------------ allocation-free sdot
//
// exoblas_sdot_noalloc_async: the sdot reduction for callers that cannot
// allocate device memory.  Its scratch (ticket, launch generation, timeout
// word and the tagged per-workgroup partials) is the file-scope array
// `scratch` below instead of the stream's buffers, so a call allocates
// nothing.  The host never names the array in a launch: each device's copy
// is resolved once with hipGetSymbolAddress and that address is what the
// kernel receives.  Every call on a device shares its copy, so calls that
// overlap on one device (on different streams, or replays of graphs that
// overlap) race; serialise them.
#include <array>
#include <atomic>

#include "exoblas_amd.h"
#include "exoblas_amd_noalloc.h"

namespace {

constexpr int kNoAllocHeader = 4;            // [0] ticket, [1] generation, [2] timeout word, [3] unused
constexpr int kNoAllocMaxBlocks = 2048;      // partial slots: plan_reduce's largest grid (4 workgroups per
                                             // CU, strided windows) on up to 512 CUs
// tagged partials: one 64-bit word (two 32-bit units) per FP32 value
constexpr size_t kNoAllocWords = kNoAllocHeader + 2 * kNoAllocMaxBlocks;  // 16,400 bytes

static __device__ __align__(256) unsigned scratch[kNoAllocWords];

struct NoAllocDevice {
  std::atomic<unsigned*> ptr{nullptr};  // this device's copy of `scratch`
  std::atomic<int> cus{0};
};
constexpr int kMaxDevices = 64;
std::array<NoAllocDevice, kMaxDevices> g_noalloc;
std::mutex g_noalloc_mutex;

int noalloc_fail(int code, const char* msg) {
  exoblas::record_error(code, msg);
  return code;
}

// Resolves the device's copy of `scratch` (with `dev` current) on the
// device's first call; later calls read the cached address.
int noalloc_device(int dev, unsigned** ptr, int* cus) {
  NoAllocDevice& d = g_noalloc[dev];
  unsigned* p = d.ptr.load(std::memory_order_acquire);
  if (p == nullptr) {
    std::lock_guard<std::mutex> g(g_noalloc_mutex);
    p = d.ptr.load(std::memory_order_acquire);
    if (p == nullptr) {
      int n = 0;
      if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n <= 0)
        return noalloc_fail(exoblas::kHipError, "exoblas_sdot_noalloc_async: CU count query failed");
      void* q = nullptr;
      if (hipGetSymbolAddress(&q, HIP_SYMBOL(scratch)) != hipSuccess || q == nullptr)
        return noalloc_fail(exoblas::kHipError, "exoblas_sdot_noalloc_async: hipGetSymbolAddress(scratch) failed");
      d.cus.store(n, std::memory_order_relaxed);
      p = static_cast<unsigned*>(q);
      d.ptr.store(p, std::memory_order_release);
    }
  }
  *ptr = p;
  *cus = d.cus.load(std::memory_order_relaxed);
  return exoblas::kOk;
}

}  // namespace

extern "C" int exoblas_sdot_noalloc_async(hipStream_t stream, int_fast32_t n, struct exo_win_1f32c x,
                                          struct exo_win_1f32c y, float* result) {
  using namespace exoblas;
  if (n < 0) return noalloc_fail(kInvalidSize, "exoblas_sdot_noalloc_async: n is negative");
  if (result == nullptr || (n > 0 && (x.data == nullptr || y.data == nullptr)))
    return noalloc_fail(kInvalidPointer, "exoblas_sdot_noalloc_async: null pointer");
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices)
    return noalloc_fail(kHipError, "exoblas_sdot_noalloc_async: no usable current device");
  unsigned* p = nullptr;
  int cus = 0;
  const int st = noalloc_device(dev, &p, &cus);
  if (st != kOk) return st;
  if (n == 0) {
    hipLaunchKernelGGL(store_zero_kernel, dim3(1), dim3(64), 0, stream, reinterpret_cast<char*>(result),
                       (int)sizeof(float));
  } else {
    const ReducePlan<float> plan = plan_reduce<float, DotTerm<float>>(n, x.data, x.strides[0], y.data,
                                                                       y.strides[0], cus);
    if (plan.blocks > kNoAllocMaxBlocks)
      return noalloc_fail(kInvalidSize, "exoblas_sdot_noalloc_async: the grid needs more partial slots than the "
                                        "fixed scratch holds (2048)");
    // ticket p[0], generation p[1], timeout word p[2]; partials from p[4]
    // (the kernel's flags + kReduceFlagBase with flags = p + 2)
    enqueue_reduce<float, DotTerm<float>, float>(stream, plan, n, p, p + 1, p + 2, p + 2, result, 0.f);
  }
  if (hipGetLastError() != hipSuccess) return noalloc_fail(kHipError, "exoblas_sdot_noalloc_async: launch failed");
  return kOk;
}
