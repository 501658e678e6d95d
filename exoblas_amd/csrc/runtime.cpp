// Host runtime of the exo_* C-ABI (declarations and the two calling modes are
// described in runtime.hpp).
#include "runtime.hpp"

#include <algorithm>
#include <condition_variable>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <thread>

#include "copy_pool.hpp"
#include "exoblas_amd.h"

namespace exoblas {

struct StreamState {
  std::mutex mutex;  // held by an _async call for its duration: calls on one stream from several threads take turns
  hipStream_t stream = nullptr;
  int device = 0;
  void* scratch[Ctx::kSlots] = {};
  size_t scratch_cap[Ctx::kSlots] = {};
  unsigned* counters = nullptr;
  unsigned* flags = nullptr;
  size_t flags_cap = 0;
  unsigned epoch = 0;
  // pinned host word the kernels' bounded waits report to (device address
  // timeout_dev)
  unsigned* timeout_host = nullptr;
  unsigned* timeout_dev = nullptr;
  // grown-out scratch and flag areas: kept, not freed, so work queued or
  // captured with the old addresses stays valid (bounded: each growth at
  // least doubles)
  std::vector<void*> retired;
};

namespace {

// Grow-only pool of device buffers used to stage host windows.
struct Pool {
  struct Block {
    void* p;
    size_t cap;
    bool busy;
  };
  std::vector<Block> blocks;
  void* get(size_t bytes, size_t* cap) {
    int best = -1;
    for (size_t i = 0; i < blocks.size(); ++i)
      if (!blocks[i].busy && blocks[i].cap >= bytes &&
          (best < 0 || blocks[i].cap < blocks[best].cap))
        best = (int)i;
    if (best >= 0) {
      blocks[best].busy = true;
      *cap = blocks[best].cap;
      return blocks[best].p;
    }
    size_t c = 1 << 16;
    while (c < bytes) c <<= 1;
    void* p = nullptr;
    if (hipMalloc(&p, c) != hipSuccess) {
      (void)hipGetLastError();
      // drop idle blocks and retry once
      for (auto it = blocks.begin(); it != blocks.end();) {
        if (!it->busy) {
          (void)hipFree(it->p);
          it = blocks.erase(it);
        } else {
          ++it;
        }
      }
      if (hipMalloc(&p, c) != hipSuccess) {
        (void)hipGetLastError();
        throw Error(kAllocError, "staging allocation of " + std::to_string(c) + " bytes failed");
      }
    }
    blocks.push_back({p, c, true});
    *cap = c;
    return p;
  }
  void put(void* p) {
    for (auto& b : blocks)
      if (b.p == p) b.busy = false;
  }
};

// Pinned, chunked staging of large host ranges.  The host side is copied by
// the copy threads into a ring of pinned chunks while the DMA engine moves
// the previous chunk, so a transfer runs near the pinned PCIe rate (57 GB/s
// per direction on MI355X) instead of the pageable path's 17-27 GB/s.  Used
// only under the device's call mutex.
class Stager {
 public:
  static constexpr size_t kChunk = size_t(8) << 20;
  static constexpr int kRing = 4;
  static constexpr size_t kMinBytes = size_t(4) << 20;  // smaller ranges: one plain copy

  ~Stager() {
    // the HIP runtime may already be gone at process exit: leave the pinned
    // chunks and events to it
  }
  void h2d(char* dev, const char* host, size_t bytes, hipStream_t s) {
    ensure();
    for (size_t i = 0, off = 0; off < bytes; ++i, off += kChunk) {
      const int slot = (int)(i % kRing);
      const size_t len = std::min(kChunk, bytes - off);
      EXO_HIP_CHECK(hipEventSynchronize(ev_[slot]));  // its previous DMA is done
      pool().copy(ring_[slot], host + off, len);
      EXO_HIP_CHECK(hipMemcpyAsync(dev + off, ring_[slot], len, hipMemcpyHostToDevice, s));
      EXO_HIP_CHECK(hipEventRecord(ev_[slot], s));
    }
  }
  // Returns after host[0, bytes) holds dev[0, bytes) as of the work queued on s.
  void d2h(char* host, const char* dev, size_t bytes, hipStream_t s) {
    ensure();
    const size_t n = (bytes + kChunk - 1) / kChunk;
    auto issue = [&](size_t i) {
      const int slot = (int)(i % kRing);
      const size_t off = i * kChunk, len = std::min(kChunk, bytes - off);
      EXO_HIP_CHECK(hipEventSynchronize(ev_[slot]));
      EXO_HIP_CHECK(hipMemcpyAsync(ring_[slot], dev + off, len, hipMemcpyDeviceToHost, s));
      EXO_HIP_CHECK(hipEventRecord(ev_[slot], s));
    };
    for (size_t i = 0; i < n && i < (size_t)kRing; ++i) issue(i);
    for (size_t i = 0; i < n; ++i) {
      const int slot = (int)(i % kRing);
      const size_t off = i * kChunk, len = std::min(kChunk, bytes - off);
      EXO_HIP_CHECK(hipEventSynchronize(ev_[slot]));
      pool().copy(host + off, ring_[slot], len);
      if (i + kRing < n) issue(i + kRing);
    }
  }

 private:
  void ensure() {
    if (ready_) return;
    for (int i = 0; i < kRing; ++i) {
      EXO_HIP_CHECK(hipHostMalloc(reinterpret_cast<void**>(&ring_[i]), kChunk, hipHostMallocDefault));
      EXO_HIP_CHECK(hipEventCreateWithFlags(&ev_[i], hipEventDisableTiming));
    }
    ready_ = true;
  }
  static CopyPool& pool() {
    static CopyPool p(4);  // 4 threads: 76 GB/s host copies measured, above PCIe
    return p;
  }
  char* ring_[kRing] = {};
  hipEvent_t ev_[kRing] = {};
  bool ready_ = false;
};

struct DeviceState {
  std::mutex call_mutex;  // serialises kSync calls on this device
  std::mutex map_mutex;   // guards `streams`
  int device = 0;
  int num_cus = 256;
  StreamState* own = nullptr;  // the library's stream for kSync calls
  std::map<hipStream_t, std::unique_ptr<StreamState>> streams;
  Pool pool;
  Stager stager;
};

std::mutex g_devices_mutex;
std::map<int, std::unique_ptr<DeviceState>> g_devices;

thread_local int t_status = kOk;
thread_local char t_message[512] = {0};

bool capturing(hipStream_t s) {
  hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(s, &st) != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  return st != hipStreamCaptureStatusNone;
}

// The state of stream s: its counters (zeroed by a fill kernel on s, so
// ordered before any of the stream's later work without a device-wide
// synchronisation) and its timeout word.
StreamState* make_stream_state(int device, hipStream_t s) {
  if (capturing(s))
    throw Error(kNotSupported,
                "first call on a stream while it is being captured: make one call on the stream before capturing");
  auto ss = std::make_unique<StreamState>();
  ss->stream = s;
  ss->device = device;
  EXO_HIP_CHECK(hipMalloc(&ss->counters, sizeof(unsigned) * Ctx::kCounters));
  fill_zero(s, ss->counters, sizeof(unsigned) * Ctx::kCounters);
  EXO_HIP_CHECK(hipHostMalloc(reinterpret_cast<void**>(&ss->timeout_host), 64,
                              hipHostMallocMapped | hipHostMallocCoherent));
  *ss->timeout_host = 0;
  EXO_HIP_CHECK(hipHostGetDevicePointer(reinterpret_cast<void**>(&ss->timeout_dev), ss->timeout_host, 0));
  return ss.release();
}

DeviceState& device_state(int device) {
  std::lock_guard<std::mutex> g(g_devices_mutex);
  auto it = g_devices.find(device);
  if (it != g_devices.end()) return *it->second;
  auto ds = std::make_unique<DeviceState>();
  ds->device = device;
  hipDeviceProp_t prop;
  EXO_HIP_CHECK(hipGetDeviceProperties(&prop, device));
  ds->num_cus = prop.multiProcessorCount;
  // A blocking stream: the synchronous entry points then order after work
  // already queued on the legacy default stream (a caller that filled its
  // device buffers there, as a torch program does by default), and work
  // queued there later waits for them, as with a BLAS call on that stream.
  hipStream_t s;
  EXO_HIP_CHECK(hipStreamCreateWithFlags(&s, hipStreamDefault));
  ds->own = make_stream_state(device, s);
  DeviceState& ref = *ds;
  g_devices.emplace(device, std::move(ds));
  return ref;
}

StreamState* stream_state(DeviceState& ds, hipStream_t s) {
  std::lock_guard<std::mutex> g(ds.map_mutex);
  auto it = ds.streams.find(s);
  if (it != ds.streams.end()) return it->second.get();
  StreamState* ss = make_stream_state(ds.device, s);
  ds.streams.emplace(s, std::unique_ptr<StreamState>(ss));
  return ss;
}

bool is_device_pointer(const void* p) {
  hipPointerAttribute_t a;
  std::memset(&a, 0, sizeof(a));
  hipError_t e = hipPointerGetAttributes(&a, p);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  return a.type == hipMemoryTypeDevice || a.type == hipMemoryTypeManaged;
}

}  // namespace

void record_error(int code, const std::string& msg) {
  t_status = code;
  std::snprintf(t_message, sizeof(t_message), "%s", msg.c_str());
  if (std::getenv("EXOBLAS_QUIET") == nullptr) std::fprintf(stderr, "exoblas_amd: %s\n", t_message);
}

void* Ctx::scratch(size_t bytes, int slot) {
  if (bytes == 0) bytes = 256;
  if (ss->scratch_cap[slot] < bytes) {
    if (capturing(stream))
      throw Error(kNotSupported, "scratch must grow to " + std::to_string(bytes) +
                                     " bytes while the stream is being captured: make the call once before capturing");
    size_t c = 1 << 20;
    while (c < bytes) c <<= 1;
    void* p = nullptr;
    EXO_HIP_CHECK(hipMalloc(&p, c));
    if (ss->scratch[slot]) ss->retired.push_back(ss->scratch[slot]);
    ss->scratch[slot] = p;
    ss->scratch_cap[slot] = c;
  }
  return ss->scratch[slot];
}

unsigned* Ctx::counters() { return ss->counters; }

unsigned* Ctx::timeout_word() { return ss->timeout_dev; }

unsigned* Ctx::epoch_flags(size_t n, unsigned* epoch) {
  if (ss->flags_cap < n || ss->epoch >= 0x7fffffffu) {
    if (capturing(stream))
      throw Error(kNotSupported, "the flag area must grow while the stream is being captured: make the call "
                                 "once before capturing");
    size_t c = 1024;
    while (c < n) c <<= 1;
    unsigned* f = nullptr;
    EXO_HIP_CHECK(hipMalloc(&f, c * sizeof(unsigned)));
    fill_zero(stream, f, c * sizeof(unsigned));
    if (ss->flags) ss->retired.push_back(ss->flags);
    ss->flags = f;
    ss->flags_cap = c;
    ss->epoch = 0;
  }
  *epoch = ++ss->epoch;
  return ss->flags;
}

unsigned* stream_timeout_word(hipStream_t s) {
  int dev = 0;
  EXO_HIP_CHECK(hipGetDevice(&dev));
  return stream_state(device_state(dev), s)->timeout_dev;
}

bool take_timeout(hipStream_t s) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return false;
  DeviceState& ds = device_state(dev);
  StreamState* ss = nullptr;
  {
    std::lock_guard<std::mutex> g(ds.map_mutex);
    auto it = ds.streams.find(s);
    if (it != ds.streams.end()) ss = it->second.get();
  }
  if (ss == nullptr) return false;
  volatile unsigned* w = ss->timeout_host;
  const unsigned v = *w;
  if (v != 0) *w = 0;
  return v != 0;
}

namespace {
std::mutex g_spin_mutex;
std::vector<int (*)(long long)>& spin_setters() {
  static std::vector<int (*)(long long)> v;
  return v;
}
}  // namespace

void register_spin_setter(int (*setter)(long long)) {
  std::lock_guard<std::mutex> g(g_spin_mutex);
  spin_setters().push_back(setter);
}

int set_spin_limit(long long polls) {
  std::lock_guard<std::mutex> g(g_spin_mutex);
  for (auto f : spin_setters())
    if (f(polls) != 0) return kHipError;
  return kOk;
}

Call::Call(const char* name, Mode mode, void* src) : name_(name), mode_(mode) {
  try {
    int dev = 0;
    EXO_HIP_CHECK(hipGetDevice(&dev));
    DeviceState& ds = device_state(dev);
    ctx_.device = dev;
    ctx_.num_cus = ds.num_cus;
    if (mode == kSync) {
      lock_ = std::unique_lock<std::mutex>(ds.call_mutex);
      ctx_.ss = ds.own;
    } else {
      ctx_.ss = stream_state(ds, static_cast<hipStream_t>(src));
      lock_ = std::unique_lock<std::mutex>(ctx_.ss->mutex);
    }
    ctx_.stream = ctx_.ss->stream;
  } catch (const Error& e) {
    fail(e.code, e.what());
  }
}

Call::~Call() {
  if (!ranges_.empty()) {
    int dev = ctx_.device;
    DeviceState& ds = device_state(dev);
    for (auto& r : ranges_)
      if (r.dev) ds.pool.put(r.dev);
  }
}

int Call::fail(int code, const std::string& msg) {
  if (status_ == kOk) {
    status_ = code;
    record_error(code, std::string(name_) + ": " + msg);
  }
  return status_;
}

int Call::fail_enum(const char* arg, int64_t v) {
  return fail(kInvalidEnum, std::string(arg) + " = " + std::to_string(v) + " is not a valid CBLAS value");
}
int Call::fail_size(const char* arg, int64_t v) {
  return fail(kInvalidSize, std::string(arg) + " = " + std::to_string(v) + " is negative");
}
int Call::fail_stride(const char* arg, int64_t v) {
  return fail(kInvalidStride, std::string(arg) + " has column stride " + std::to_string(v) + " (must be 1)");
}
int Call::fail_ld(const char* arg, int64_t v) {
  return fail(kInvalidLd, std::string(arg) + " has leading dimension " + std::to_string(v) +
                              " smaller than its row length");
}

void Call::window(int idx, const void* p, int64_t lo, int64_t hi, size_t esize, bool write, bool used,
                  bool result_only) {
  Win& w = wins_[idx];
  w.p = static_cast<const char*>(p);
  w.lo = lo;
  w.hi = hi;
  w.esize = esize;
  w.write = write;
  w.used = used && hi > lo;
  w.result_only = result_only;
  w.dev = used ? const_cast<void*>(p) : nullptr;
  nwin_ = std::max(nwin_, idx + 1);
}

bool Call::stage() {
  if (status_ != kOk) return false;
  if (mode_ == kAsync) {
    for (int i = 0; i < nwin_; ++i)
      if (wins_[i].used && wins_[i].p == nullptr) {
        fail(kInvalidPointer, "null data pointer");
        return false;
      }
    return true;
  }
  try {
    // Host windows, merged where their byte ranges overlap.
    struct Item {
      const char* b;
      const char* e;
      int win;
    };
    std::vector<Item> host;
    for (int i = 0; i < nwin_; ++i) {
      Win& w = wins_[i];
      if (!w.used) continue;
      if (w.p == nullptr) {
        fail(kInvalidPointer, "null data pointer");
        return false;
      }
      const char* b = w.p + w.lo * (int64_t)w.esize;
      const char* e = w.p + w.hi * (int64_t)w.esize;
      if (is_device_pointer(b)) continue;
      host.push_back({b, e, i});
    }
    std::sort(host.begin(), host.end(), [](const Item& a, const Item& b) { return a.b < b.b; });
    for (auto& it : host) {
      Win& w = wins_[it.win];
      if (ranges_.empty() || it.b >= ranges_.back().end) {
        ranges_.push_back({it.b, it.e, false, false, nullptr, 0});
      } else {
        ranges_.back().end = std::max(ranges_.back().end, it.e);
      }
      Range& r = ranges_.back();
      r.read = r.read || !w.result_only;
      r.write = r.write || w.write;
      w.range = (int)ranges_.size() - 1;
    }
    DeviceState& ds = device_state(ctx_.device);
    for (auto& r : ranges_) {
      size_t bytes = (size_t)(r.end - r.begin);
      r.dev = ds.pool.get(bytes, &r.cap);
      if (r.read) {
        if (bytes >= Stager::kMinBytes)
          ds.stager.h2d(static_cast<char*>(r.dev), r.begin, bytes, ctx_.stream);
        else
          EXO_HIP_CHECK(hipMemcpyAsync(r.dev, r.begin, bytes, hipMemcpyHostToDevice, ctx_.stream));
      }
    }
    for (int i = 0; i < nwin_; ++i) {
      Win& w = wins_[i];
      if (w.range >= 0) {
        Range& r = ranges_[w.range];
        w.dev = static_cast<char*>(r.dev) + (w.p - r.begin);
      }
    }
  } catch (const Error& e) {
    fail(e.code, e.what());
    return false;
  }
  staged_ = true;
  return true;
}

void Call::read_host(const void* p, void* out, size_t bytes) {
  if (status_ != kOk) return;
  if (p == nullptr) {
    fail(kInvalidPointer, "null scalar pointer");
    return;
  }
  if (mode_ == kSync && is_device_pointer(p)) {
    hipError_t e = hipMemcpy(out, p, bytes, hipMemcpyDeviceToHost);
    if (e != hipSuccess) fail(kHipError, hipGetErrorString(e));
    return;
  }
  std::memcpy(out, p, bytes);
}

int Call::finish() {
  try {
    if (status_ == kOk) EXO_LAUNCH_CHECK();
    if (mode_ == kSync && ctx_.stream) {
      if (status_ == kOk) {
        DeviceState& ds = device_state(ctx_.device);
        for (auto& r : ranges_) {
          if (!r.write) continue;
          const size_t bytes = (size_t)(r.end - r.begin);
          if (bytes >= Stager::kMinBytes)
            ds.stager.d2h(const_cast<char*>(r.begin), static_cast<const char*>(r.dev), bytes, ctx_.stream);
          else
            EXO_HIP_CHECK(hipMemcpyAsync(const_cast<char*>(r.begin), r.dev, bytes, hipMemcpyDeviceToHost,
                                         ctx_.stream));
        }
      }
      EXO_HIP_CHECK(hipStreamSynchronize(ctx_.stream));
      volatile unsigned* w = ctx_.ss->timeout_host;
      if (*w != 0) {
        *w = 0;
        fail(kTimeout, "a bounded wait between workgroups expired: the results are not valid");
      }
    }
  } catch (const Error& e) {
    fail(e.code, e.what());
  }
  return status_;
}

}  // namespace exoblas

// ------------------------------------------------------------ C utilities

extern "C" int exoblas_last_status(void) { return exoblas::t_status; }
extern "C" const char* exoblas_last_error(void) { return exoblas::t_message; }
extern "C" void exoblas_clear_error(void) {
  exoblas::t_status = exoblas::kOk;
  exoblas::t_message[0] = 0;
}

extern "C" int exoblas_init(void) {
  try {
    int dev = 0;
    EXO_HIP_CHECK(hipGetDevice(&dev));
    exoblas::device_state(dev);
  } catch (const exoblas::Error& e) {
    exoblas::record_error(e.code, e.what());
    return e.code;
  }
  return 0;
}

extern "C" int exoblas_stream_check(void* stream_) {
  hipStream_t stream = static_cast<hipStream_t>(stream_);
  try {
    EXO_HIP_CHECK(hipStreamSynchronize(stream));
    if (exoblas::take_timeout(stream)) {
      exoblas::record_error(exoblas::kTimeout,
                            "exoblas_stream_check: a bounded wait between workgroups expired on this stream: "
                            "results of its calls since the last check are not valid");
      return exoblas::kTimeout;
    }
  } catch (const exoblas::Error& e) {
    exoblas::record_error(e.code, e.what());
    return e.code;
  }
  return 0;
}

extern "C" int exoblas_stream_release(void* stream_) {
  using namespace exoblas;
  hipStream_t s = static_cast<hipStream_t>(stream_);
  try {
    if (capturing(s)) throw Error(kNotSupported, "exoblas_stream_release: the stream is being captured");
    std::unique_ptr<StreamState> ss;
    {
      std::lock_guard<std::mutex> gd(g_devices_mutex);
      for (auto& d : g_devices) {
        std::lock_guard<std::mutex> g(d.second->map_mutex);
        auto it = d.second->streams.find(s);
        if (it == d.second->streams.end()) continue;
        ss = std::move(it->second);
        d.second->streams.erase(it);
        break;
      }
    }
    if (!ss) return kOk;
    std::lock_guard<std::mutex> g(ss->mutex);  // a call in progress on the stream finishes first
    int prev = 0;
    EXO_HIP_CHECK(hipGetDevice(&prev));
    EXO_HIP_CHECK(hipSetDevice(ss->device));
    hipError_t e = hipStreamSynchronize(s);
    for (int i = 0; i < Ctx::kSlots && e == hipSuccess; ++i)
      if (ss->scratch[i]) e = hipFree(ss->scratch[i]);
    for (size_t i = 0; i < ss->retired.size() && e == hipSuccess; ++i) e = hipFree(ss->retired[i]);
    if (e == hipSuccess && ss->flags) e = hipFree(ss->flags);
    if (e == hipSuccess && ss->counters) e = hipFree(ss->counters);
    if (e == hipSuccess && ss->timeout_host) e = hipHostFree(ss->timeout_host);
    (void)hipSetDevice(prev);
    EXO_HIP_CHECK(e);
  } catch (const Error& e) {
    record_error(e.code, e.what());
    return e.code;
  }
  return kOk;
}

extern "C" int exoblas_set_spin_limit(long long polls) {
  const int st = exoblas::set_spin_limit(polls);
  if (st != 0) exoblas::record_error(st, "exoblas_set_spin_limit: hipMemcpyToSymbol failed");
  return st;
}

extern "C" int exoblas_synchronize(void) {
  try {
    int dev = 0;
    EXO_HIP_CHECK(hipGetDevice(&dev));
    auto& ds = exoblas::device_state(dev);
    EXO_HIP_CHECK(hipStreamSynchronize(ds.own->stream));
  } catch (const exoblas::Error& e) {
    exoblas::record_error(e.code, e.what());
    return e.code;
  }
  return 0;
}
