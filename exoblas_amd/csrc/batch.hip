// Persistent batched kernel for gfx950: many uneven FP32 gemm / gemv / trsv
// problems in one launch.
//
// Every problem is cut into work items: a gemm into C tiles of the size the
// tuning table picks for its shape (32, 64 or 128 square), a gemv into
// 64-row blocks, a trsv into its 64-row substitution steps.  The items of a
// batch form one queue; a grid sized to what the chip holds at once takes
// items from it with an atomic counter until it is empty, so short and long
// problems pack onto the CUs without a launch per problem.  trsv steps of one
// problem sit in the queue in processing order and wait (bounded) on the
// flags of earlier steps only; an item is taken only by a running workgroup,
// so every step it waits on is already running and the queue cannot
// deadlock.  The last workgroup to finish resets the queue for the next
// launch and advances the launch count that marks this launch's trsv flags,
// so a launch needs no memset and can be replayed from a hipGraph.
#include <algorithm>
#include <cstring>
#include <map>
#include <mutex>

#include "batch.hpp"
#include "device.hpp"
#include "exoblas_amd.h"
#include "gemm_core.hpp"
#include "runtime.hpp"
#include "trsv_core.hpp"

namespace exoblas {
namespace batch {

using gemm::Cfg;
using gemm::Operand;
using Tile0 = Cfg<float, 32, 32, 32, 16, 16>;
using Tile1 = Cfg<float, 64, 64, 16, 32, 32>;
using Tile2 = Cfg<float, 128, 128, 16, 64, 64>;
constexpr int kNT = 256;
static_assert(Tile0::NT == kNT && Tile1::NT == kNT && Tile2::NT == kNT, "one workgroup size");
constexpr int kTileSize[3] = {32, 64, 128};

constexpr size_t cmax(size_t a, size_t b) { return a > b ? a : b; }
constexpr size_t kSmem =
    cmax(cmax(Tile0::LDS_BYTES, Tile1::LDS_BYTES), cmax(Tile2::LDS_BYTES, sizeof(trsv::Smem<float>)));

__device__ __forceinline__ int vec_ok(const float* p, int64_t ld) {
  return (ld % 4 == 0) && ((reinterpret_cast<uintptr_t>(p) & 15) == 0);
}

template <class C>
__device__ __forceinline__ void gemm_item(const Desc& d, uint32_t local, float* lds) {
  const int64_t tm = local / d.tiles_n, tn = local % d.tiles_n;
  const int64_t m0 = tm * C::BM, n0 = tn * C::BN;
  Operand<float> A{d.a, d.lda, d.ta == 111 ? gemm::kRK : gemm::kKR, vec_ok(d.a, d.lda)};
  Operand<float> B{d.b, d.ldb, d.tb == 111 ? gemm::kKR : gemm::kRK, vec_ok(d.b, d.ldb)};
  typename gemm::Mfma<float>::acc acc[C::SM][C::SN];
#pragma unroll
  for (int a = 0; a < C::SM; ++a)
#pragma unroll
    for (int b = 0; b < C::SN; ++b) acc[a][b] = gemm::v4f{0.f, 0.f, 0.f, 0.f};
  gemm::tile_mma<C>(A, B, d.m, d.n, d.k, m0, n0, 0, d.k, acc, lds);
  const int cvec = (d.ldc % C::EB == 0) && ((reinterpret_cast<uintptr_t>(d.c) % (C::EB * 4)) == 0);
  gemm::tile_store<C, gemm::kFull, true, true>(d.c, d.ldc, cvec, d.m, d.n, m0, n0, d.alpha, acc, d.beta);
  __syncthreads();  // LDS reused by the next item
}

// y[r] = alpha * A[r, :] x + beta * y[r] for the 64 rows of item `local`;
// each wave owns 16 rows, lanes stride the columns 16 bytes at a time.
__device__ __forceinline__ void gemv_item(const Desc& d, uint32_t local) {
  constexpr int RW = kRows / (kNT / 64);
  const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
  const int64_t rb = (int64_t)local * kRows + wave * RW;
  const float* x = d.b;
  const int64_t incx = d.ldb;
  float acc[RW];
#pragma unroll
  for (int r = 0; r < RW; ++r) acc[r] = 0.f;
  const bool vec = incx == 1 && vec_ok(d.a, d.lda) && (reinterpret_cast<uintptr_t>(x) & 15) == 0;
  int64_t c0 = 0;
  if (vec) {
    const int64_t nv = d.n / 4;
    // rows past the end read the last row (in bounds) and are not stored, so
    // the loop has no per-row branch; A is streamed (read once)
    const float* arow[RW];
#pragma unroll
    for (int r = 0; r < RW; ++r) arow[r] = d.a + (rb + r < d.m ? rb + r : d.m - 1) * d.lda;
    for (int64_t v = lane; v < nv; v += 64) {
      const float4 xv = reinterpret_cast<const float4*>(x)[v];
      float4 av[RW];
#pragma unroll
      for (int r = 0; r < RW; ++r) av[r] = ld_stream<float>(arow[r] + 4 * v);
#pragma unroll
      for (int r = 0; r < RW; ++r) acc[r] += av[r].x * xv.x + av[r].y * xv.y + av[r].z * xv.z + av[r].w * xv.w;
    }
    c0 = nv * 4;
  }
  for (int64_t col = c0 + lane; col < d.n; col += 64) {
    const float xv = x[col * incx];
#pragma unroll
    for (int r = 0; r < RW; ++r) {
      const int64_t row = rb + r;
      if (row < d.m) acc[r] += d.a[row * d.lda + col] * xv;
    }
  }
#pragma unroll
  for (int r = 0; r < RW; ++r) {
    const float s = wave_sum(acc[r]);
    const int64_t row = rb + r;
    if (lane == 0 && row < d.m) {
      float* yp = d.c + row * d.ldc;
      *yp = d.alpha * s + d.beta * *yp;  // y read for every beta, as the per-call gemv does
    }
  }
}

template <bool FWD, bool TRANS, bool UNIT>
__device__ __forceinline__ void trsv_item(const Desc& d, uint32_t local, unsigned* ctrl, unsigned marker,
                                          unsigned* err, trsv::Smem<float>& sm) {
  trsv::step<float, FWD, TRANS, UNIT, trsv::kB, false>((int)local, d.n, d.a, d.lda, d.c,
                                                       ctrl + kCtrlWords + d.flag_off, marker, err, sm);
}

__global__ __launch_bounds__(kNT) void batch_kernel(const Desc* __restrict__ descs, const uint64_t* __restrict__ order,
                                                    uint32_t total, unsigned* ctrl, unsigned* err) {
  extern __shared__ __align__(16) unsigned char smem[];
  __shared__ uint32_t s_item;
  const unsigned marker = __hip_atomic_load(ctrl + 2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) + 1u;
  for (;;) {
    if (threadIdx.x == 0) s_item = atomicAdd(ctrl, 1u);
    __syncthreads();
    const uint32_t item = s_item;
    __syncthreads();
    if (item >= total) break;
    const uint64_t e = order[item];
    const Desc d = descs[e >> 32];
    const uint32_t local = (uint32_t)e;
    if (d.kind == EXOBLAS_OP_GEMM) {
      float* lds = reinterpret_cast<float*>(smem);
      if (d.tile == 0) gemm_item<Tile0>(d, local, lds);
      else if (d.tile == 1) gemm_item<Tile1>(d, local, lds);
      else gemm_item<Tile2>(d, local, lds);
    } else if (d.kind == EXOBLAS_OP_GEMV) {
      gemv_item(d, local);
    } else {
      auto& sm = *reinterpret_cast<trsv::Smem<float>*>(smem);
      const bool trans = d.ta != 111, lower = d.tb == 122, unit = d.diag == 132;
      const bool fwd = lower != trans;
      if (fwd && !trans) { if (unit) trsv_item<true, false, true>(d, local, ctrl, marker, err, sm); else trsv_item<true, false, false>(d, local, ctrl, marker, err, sm); }
      if (fwd && trans) { if (unit) trsv_item<true, true, true>(d, local, ctrl, marker, err, sm); else trsv_item<true, true, false>(d, local, ctrl, marker, err, sm); }
      if (!fwd && !trans) { if (unit) trsv_item<false, false, true>(d, local, ctrl, marker, err, sm); else trsv_item<false, false, false>(d, local, ctrl, marker, err, sm); }
      if (!fwd && trans) { if (unit) trsv_item<false, true, true>(d, local, ctrl, marker, err, sm); else trsv_item<false, true, false>(d, local, ctrl, marker, err, sm); }
    }
  }
  if (threadIdx.x == 0) {
    const unsigned fin = atomicAdd(ctrl + 1, 1u);
    if (fin == gridDim.x - 1) {  // every workgroup is past its last queue access
      atomicExch(ctrl, 0u);
      atomicExch(ctrl + 1, 0u);
      atomicAdd(ctrl + 2, 1u);
    }
  }
}

// ------------------------------------------------------------ host side

namespace {

std::mutex g_tune_mutex;
int32_t g_table[72];
int32_t g_forced = -1;
bool g_table_init = false;

int size_class(int64_t v) {
  return v <= 32 ? 0 : v <= 64 ? 1 : v <= 128 ? 2 : v <= 256 ? 3 : v <= 512 ? 4 : 5;
}

void init_table_locked() {
  if (g_table_init) return;
  // Default before tuning: the largest tile that still leaves a few tiles
  // per problem (tile count is what feeds the persistent grid).
  for (int kc = 0; kc < 2; ++kc)
    for (int mc = 0; mc < 6; ++mc)
      for (int nc = 0; nc < 6; ++nc) {
        const int lo = std::min(mc, nc);
        g_table[kc * 36 + mc * 6 + nc] = lo >= 4 ? 2 : lo >= 2 ? 1 : 0;
      }
  g_table_init = true;
}

int choose_tile(int64_t m, int64_t n, int64_t k) {
  std::lock_guard<std::mutex> g(g_tune_mutex);
  init_table_locked();
  if (g_forced >= 0) return g_forced;
  return g_table[(k > 256 ? 36 : 0) + size_class(m) * 6 + size_class(n)];
}

int forced_tile() {
  std::lock_guard<std::mutex> g(g_tune_mutex);
  return g_forced;
}

// A batch whose gemms alone give at least this many 128 x 128 tiles keeps the
// grid busy with the largest tiles, which run at the highest rate; the tuned
// table (chosen per problem in isolation, where small tiles win by exposing
// parallelism) is for narrower batches.  Serving batches of 256 mixed
// requests: 13.6 ms for the 12 k gemms of the stream at 128 x 128 against
// 20.5 ms with the per-problem choice.
constexpr int64_t kWideItems = 1024;

int g_max_grid = 0;

int max_grid() {
  if (g_max_grid) return g_max_grid;
  int dev = 0, cus = 256, occ = 1;
  (void)hipGetDevice(&dev);
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, dev) == hipSuccess) cus = prop.multiProcessorCount;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, batch_kernel, kNT, kSmem) != hipSuccess || occ < 1) {
    (void)hipGetLastError();
    occ = 1;
  }
  g_max_grid = cus * occ;
  return g_max_grid;
}

}  // namespace

int64_t operand_bytes(const exoblas_batch_op& op) {
  const int64_t s = 4;
  if (op.kind == EXOBLAS_OP_GEMM) return s * (op.m * op.k + op.k * op.n + op.m * op.n);
  if (op.kind == EXOBLAS_OP_GEMV) return s * (op.m * op.n + op.n + op.m);
  if (op.kind == EXOBLAS_OP_TRSV) return s * (op.m * (op.m + 1) / 2 + op.m);
  return 0;
}

int build(const exoblas_batch_op* ops, int64_t count, Built& out, std::string& err) {
  out = Built();
  if (count < 0) {
    err = "negative count";
    return kInvalidSize;
  }
  struct Item {
    Desc d;
    uint32_t items;
    int group;     // 0 trsv, 1 gemm, 2 gemv
    double cost;   // per-item cost, for longest-first order
    int64_t index;
  };
  std::vector<Item> v;
  v.reserve(count);
  for (int64_t i = 0; i < count; ++i) {
    const exoblas_batch_op& o = ops[i];
    Desc d{};
    d.kind = o.kind;
    d.ta = o.trans_a;
    d.tb = o.trans_b;
    d.diag = o.diag;
    d.a = o.a;
    d.lda = o.lda;
    d.b = o.b;
    d.ldb = o.ldb;
    d.c = o.c;
    d.ldc = o.ldc;
    d.alpha = o.alpha;
    d.beta = o.beta;
    const std::string at = "op " + std::to_string(i) + ": ";
    auto tr_ok = [](int32_t t) { return t == 111 || t == 112; };
    if (o.kind == EXOBLAS_OP_GEMM) {
      if (!tr_ok(o.trans_a) || !tr_ok(o.trans_b)) { err = at + "invalid trans"; return kInvalidEnum; }
      if (o.m < 0 || o.n < 0 || o.k < 0) { err = at + "negative size"; return kInvalidSize; }
      const int64_t acols = o.trans_a == 111 ? o.k : o.m, bcols = o.trans_b == 111 ? o.n : o.k;
      if ((o.lda < std::max<int64_t>(1, acols)) || o.ldb < std::max<int64_t>(1, bcols) ||
          o.ldc < std::max<int64_t>(1, o.n)) { err = at + "leading dimension too small"; return kInvalidLd; }
      if (o.m == 0 || o.n == 0) continue;
      if (!o.c || (o.k > 0 && (!o.a || !o.b))) { err = at + "null pointer"; return kInvalidPointer; }
      d.m = o.m; d.n = o.n; d.k = o.k;
      d.tile = choose_tile(o.m, o.n, o.k);
      const int ts = kTileSize[d.tile];
      const int64_t tm = (o.m + ts - 1) / ts, tn = (o.n + ts - 1) / ts;
      d.tiles_n = (int32_t)tn;
      v.push_back({d, (uint32_t)(tm * tn), 1, (double)ts * ts * (double)std::max<int64_t>(o.k, 1), i});
    } else if (o.kind == EXOBLAS_OP_GEMV) {
      if (o.trans_a != 111) { err = at + "batched gemv is NoTrans only"; return kNotSupported; }
      if (o.m < 0 || o.n < 0) { err = at + "negative size"; return kInvalidSize; }
      if (o.lda < std::max<int64_t>(1, o.n)) { err = at + "leading dimension too small"; return kInvalidLd; }
      if (o.ldb == 0 || o.ldc == 0) { err = at + "zero increment"; return kInvalidStride; }
      if (o.m == 0) continue;
      if (!o.c || (o.n > 0 && (!o.a || !o.b))) { err = at + "null pointer"; return kInvalidPointer; }
      d.m = o.m; d.n = o.n; d.k = 0;
      v.push_back({d, (uint32_t)((o.m + kRows - 1) / kRows), 2, (double)kRows * o.n, i});
    } else if (o.kind == EXOBLAS_OP_TRSV) {
      if (!tr_ok(o.trans_a) || (o.trans_b != 121 && o.trans_b != 122) || (o.diag != 131 && o.diag != 132)) {
        err = at + "invalid enum"; return kInvalidEnum;
      }
      if (o.m < 0) { err = at + "negative size"; return kInvalidSize; }
      if (o.ldc != 1) { err = at + "batched trsv needs a unit-stride x"; return kNotSupported; }
      if (o.lda < std::max<int64_t>(1, o.m)) { err = at + "leading dimension too small"; return kInvalidLd; }
      if (o.m == 0) continue;
      if (!o.a || !o.c) { err = at + "null pointer"; return kInvalidPointer; }
      d.m = o.m; d.n = o.m; d.k = 0;
      const uint32_t nb = (uint32_t)((o.m + kRows - 1) / kRows);
      d.k = trsv::pub_words<float>(o.m);  // publication words this problem needs
      v.push_back({d, nb, 0, (double)nb, i});
    } else {
      err = at + "unknown kind " + std::to_string(o.kind);
      return kInvalidEnum;
    }
  }
  int64_t items128 = 0;
  for (const auto& it : v)
    if (it.d.kind == EXOBLAS_OP_GEMM) items128 += ((it.d.m + 127) / 128) * ((it.d.n + 127) / 128);
  if (items128 >= kWideItems && forced_tile() < 0) {
    for (auto& it : v) {
      if (it.d.kind != EXOBLAS_OP_GEMM) continue;
      it.d.tile = 2;
      const int64_t tm = (it.d.m + 127) / 128, tn = (it.d.n + 127) / 128;
      it.d.tiles_n = (int32_t)tn;
      it.items = (uint32_t)(tm * tn);
      it.cost = 128.0 * 128.0 * (double)std::max<int64_t>(it.d.k, 1);
    }
  }
  // Problems: trsv chains first, the longest first; then gemm and gemv, the
  // most expensive items first.  Stable on the submission index.
  std::stable_sort(v.begin(), v.end(), [](const Item& a, const Item& b) {
    if (a.group != b.group) return a.group < b.group;
    return a.cost > b.cost;
  });
  uint64_t flags = 0, total = 0;
  uint32_t rounds = 0;
  size_t ntrsv = 0;
  for (auto& it : v) {
    if (it.d.kind == EXOBLAS_OP_TRSV) {
      it.d.flag_off = (uint32_t)flags;  // even: every problem's word count is
      flags += (uint64_t)it.d.k;        // even and the region starts 16-byte aligned
      rounds = std::max(rounds, it.items);
      ++ntrsv;
    }
    total += it.items;
    out.descs.push_back(it.d);
  }
  if (total > 0xffffff00ull) {
    err = "batch too large";
    return kInvalidSize;
  }
  // The queue.  A trsv step waits only on the step before it in its chain,
  // which sits in an earlier round, so every wait is on a queue entry some
  // running workgroup already took and the queue cannot deadlock.  Spreading
  // the steps over rounds keeps at most about one waiting workgroup per
  // chain; the gemm / gemv items between rounds keep the rest of the grid
  // busy while the chains advance.
  out.order.reserve((size_t)total);
  auto entry = [](size_t di, uint32_t local) { return ((uint64_t)di << 32) | local; };
  size_t nd = ntrsv, nl = 0;  // next non-trsv problem and item
  uint64_t rest = total;
  for (size_t di = 0; di < ntrsv; ++di) rest -= v[di].items;
  const uint64_t share = rounds ? (rest + rounds - 1) / rounds : rest;
  auto take_rest = [&](uint64_t k) {
    for (; k > 0 && nd < v.size(); --k) {
      out.order.push_back(entry(nd, (uint32_t)nl));
      if (++nl == v[nd].items) {
        ++nd;
        nl = 0;
      }
    }
  };
  for (uint32_t r = 0; r < rounds; ++r) {
    for (size_t di = 0; di < ntrsv; ++di)
      if (v[di].items > r) out.order.push_back(entry(di, r));
    take_rest(share);
  }
  take_rest(rest);
  out.total_items = (uint32_t)out.order.size();
  out.flag_words = (uint32_t)flags;
  return kOk;
}

void DeviceBatch::release() {
  if (desc) (void)hipFree(desc);
  if (order) (void)hipFree(order);
  if (ctrl) (void)hipFree(ctrl);
  desc = nullptr;
  order = nullptr;
  ctrl = nullptr;
  cap_desc = cap_order = cap_flags = 0;
}

void DeviceBatch::upload(const Built& b, hipStream_t stream) {
  const size_t nd = std::max<size_t>(1, b.descs.size());
  const size_t no = std::max<size_t>(1, b.order.size());
  const size_t nf = (size_t)b.flag_words + kCtrlWords;
  if (nd > cap_desc || no > cap_order || nf > cap_flags || !ctrl) {
    EXO_HIP_CHECK(hipStreamSynchronize(stream));
    if (nd > cap_desc) {
      if (desc) EXO_HIP_CHECK(hipFree(desc));
      size_t c = 64;
      while (c < nd) c <<= 1;
      EXO_HIP_CHECK(hipMalloc(&desc, c * sizeof(Desc)));
      cap_desc = c;
    }
    if (no > cap_order) {
      if (order) EXO_HIP_CHECK(hipFree(order));
      size_t c = 1024;
      while (c < no) c <<= 1;
      EXO_HIP_CHECK(hipMalloc(&order, c * sizeof(uint64_t)));
      cap_order = c;
    }
    if (nf > cap_flags || !ctrl) {
      if (ctrl) EXO_HIP_CHECK(hipFree(ctrl));
      size_t c = 1024;
      while (c < nf) c <<= 1;
      EXO_HIP_CHECK(hipMalloc(&ctrl, c * sizeof(unsigned)));
      fill_zero(stream, ctrl, c * sizeof(unsigned));  // ordered before the launch on this stream
      cap_flags = c;
    }
  }
  if (!b.descs.empty()) {
    EXO_HIP_CHECK(hipMemcpyAsync(desc, b.descs.data(), b.descs.size() * sizeof(Desc), hipMemcpyHostToDevice, stream));
    EXO_HIP_CHECK(hipMemcpyAsync(order, b.order.data(), b.order.size() * sizeof(uint64_t), hipMemcpyHostToDevice,
                                 stream));
  }
}

void launch(const DeviceBatch& d, const Built& b, hipStream_t stream) {
  if (b.total_items == 0) return;
  const int grid = (int)std::min<int64_t>(b.total_items, max_grid());
  unsigned* err = stream_timeout_word(stream);
  hipLaunchKernelGGL(batch_kernel, dim3(grid), dim3(kNT), kSmem, stream, d.desc, d.order, b.total_items, d.ctrl,
                     err);
  EXO_LAUNCH_CHECK();
}

namespace {

struct StreamSlots {
  std::mutex m;
  std::map<hipStream_t, DeviceBatch> slots;
};
StreamSlots g_slots;

int fail(int code, const std::string& msg) {
  record_error(code, msg);
  return code;
}

}  // namespace

}  // namespace batch
}  // namespace exoblas

using namespace exoblas;

struct exoblas_batch_plan_s {
  batch::Built built;
  batch::DeviceBatch dev;
};

extern "C" int exoblas_batch_run(hipStream_t stream, const exoblas_batch_op* ops, int64_t count) {
  try {
    batch::Built b;
    std::string err;
    const int st = batch::build(ops, count, b, err);
    if (st != kOk) return batch::fail(st, "exoblas_batch_run: " + err);
    if (b.total_items == 0) return kOk;
    std::lock_guard<std::mutex> g(batch::g_slots.m);
    batch::DeviceBatch& d = batch::g_slots.slots[stream];
    d.upload(b, stream);
    batch::launch(d, b, stream);
  } catch (const Error& e) {
    return batch::fail(e.code, std::string("exoblas_batch_run: ") + e.what());
  }
  return kOk;
}

extern "C" int exoblas_batch_queue(const exoblas_batch_op* ops, int64_t count, uint64_t* order, int32_t* kinds,
                                   int64_t cap, int64_t* n_items, int64_t* n_problems) {
  if (!n_items || !n_problems) return batch::fail(kInvalidPointer, "exoblas_batch_queue: null count pointer");
  try {
    batch::Built b;
    std::string err;
    const int st = batch::build(ops, count, b, err);
    if (st != kOk) return batch::fail(st, "exoblas_batch_queue: " + err);
    *n_items = (int64_t)b.order.size();
    *n_problems = (int64_t)b.descs.size();
    if (*n_items <= cap && *n_problems <= cap && order && kinds) {
      std::copy(b.order.begin(), b.order.end(), order);
      for (size_t i = 0; i < b.descs.size(); ++i) kinds[i] = b.descs[i].kind;
    }
  } catch (const Error& e) {
    return batch::fail(e.code, std::string("exoblas_batch_queue: ") + e.what());
  }
  return kOk;
}

extern "C" int exoblas_batch_plan_create(exoblas_batch_plan* plan, const exoblas_batch_op* ops, int64_t count) {
  if (!plan) return batch::fail(kInvalidPointer, "exoblas_batch_plan_create: null plan");
  auto* p = new exoblas_batch_plan_s();
  try {
    std::string err;
    const int st = batch::build(ops, count, p->built, err);
    if (st != kOk) {
      delete p;
      return batch::fail(st, "exoblas_batch_plan_create: " + err);
    }
    p->dev.upload(p->built, nullptr);
    EXO_HIP_CHECK(hipStreamSynchronize(nullptr));
  } catch (const Error& e) {
    p->dev.release();
    delete p;
    return batch::fail(e.code, std::string("exoblas_batch_plan_create: ") + e.what());
  }
  *plan = p;
  return kOk;
}

extern "C" int exoblas_batch_plan_launch(exoblas_batch_plan plan, hipStream_t stream) {
  if (!plan) return batch::fail(kInvalidPointer, "exoblas_batch_plan_launch: null plan");
  try {
    batch::launch(plan->dev, plan->built, stream);
  } catch (const Error& e) {
    return batch::fail(e.code, std::string("exoblas_batch_plan_launch: ") + e.what());
  }
  return kOk;
}

extern "C" int exoblas_batch_plan_destroy(exoblas_batch_plan plan) {
  if (!plan) return kOk;
  plan->dev.release();
  delete plan;
  return kOk;
}

extern "C" int exoblas_batch_set_tuning(const int32_t* table, int64_t len) {
  if (!table || len != 72) return batch::fail(kInvalidSize, "exoblas_batch_set_tuning: the table has 72 entries");
  for (int64_t i = 0; i < len; ++i)
    if (table[i] < 0 || table[i] > 2) return batch::fail(kInvalidEnum, "exoblas_batch_set_tuning: tile id 0..2");
  std::lock_guard<std::mutex> g(batch::g_tune_mutex);
  std::memcpy(batch::g_table, table, sizeof(batch::g_table));
  batch::g_table_init = true;
  return kOk;
}

extern "C" int exoblas_batch_get_tuning(int32_t* table, int64_t len) {
  if (!table || len != 72) return batch::fail(kInvalidSize, "exoblas_batch_get_tuning: the table has 72 entries");
  std::lock_guard<std::mutex> g(batch::g_tune_mutex);
  batch::init_table_locked();
  std::memcpy(table, batch::g_table, sizeof(batch::g_table));
  return kOk;
}

extern "C" int exoblas_batch_force_tile(int32_t tile) {
  if (tile < -1 || tile > 2) return batch::fail(kInvalidEnum, "exoblas_batch_force_tile: tile id -1..2");
  std::lock_guard<std::mutex> g(batch::g_tune_mutex);
  batch::g_forced = tile;
  return kOk;
}
