// Host runtime of the exo_* C-ABI: per-device state, stream-ordered scratch,
// host-window staging and error reporting.
//
// Two calling modes share one set of kernels:
//   kSync  -- the drop-in exo_* symbols.  Windows may point at host or device
//             memory; host ranges are staged through pooled device buffers and
//             the call returns after results are back in the caller's memory
//             (the reference's functions are synchronous, test/level*/<k>/
//             exo_<k>_wrapper.h).  The `void *ctxt` argument is ignored, as the
//             reference always passes nullptr.
//   kAsync -- the exoblas_*_async symbols.  Windows and results are device
//             pointers; work is enqueued on the caller's stream.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

namespace exoblas {

enum Status : int {
  kOk = 0,
  kInvalidEnum = 1,
  kInvalidSize = 2,
  kInvalidStride = 3,
  kInvalidLd = 4,
  kHipError = 5,
  kAllocError = 6,
  kNotSupported = 7,
  kInvalidPointer = 8,
  kTimeout = 9,
  kCancelled = 10,
};

// Element offset range [lo, hi) covered by a 1-D window of n elements with
// stride s (element 0 at offset 0; negative strides extend downwards).
inline int64_t span_lo(int64_t n, int64_t s) { return (n > 0 && s < 0) ? (n - 1) * s : 0; }
inline int64_t span_hi(int64_t n, int64_t s) { return n > 0 ? (s > 0 ? (n - 1) * s : 0) + 1 : 0; }
// Row-major rows x cols window with leading dimension ld and unit column stride.
inline int64_t span2_hi(int64_t rows, int64_t cols, int64_t ld) {
  return (rows > 0 && cols > 0) ? (rows - 1) * ld + cols : 0;
}

// Raised inside ops; converted to a status by Call::run.
struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define EXO_HIP_CHECK(expr)                                                          \
  do {                                                                               \
    hipError_t _e = (expr);                                                          \
    if (_e != hipSuccess)                                                            \
      throw ::exoblas::Error(::exoblas::kHipError,                                   \
                             std::string(#expr) + ": " + hipGetErrorString(_e));     \
  } while (0)

// Checks the launch that was just enqueued.
#define EXO_LAUNCH_CHECK() EXO_HIP_CHECK(hipGetLastError())

struct StreamState;

// What an op launches with.
struct Ctx {
  hipStream_t stream = nullptr;
  int device = 0;
  int num_cus = 256;
  StreamState* ss = nullptr;

  // Stream-ordered device scratch of at least `bytes`, 256-byte aligned, in
  // one of kSlots independent buffers.  A slot's memory is valid for the
  // work enqueued until the next scratch() call for that slot on this
  // stream.  Growing a slot allocates a larger buffer and retires the old one
  // without freeing it (kernels still queued, and hipGraphs captured
  // earlier, keep valid pointers), so growth never synchronises.  Growth
  // while the stream is being captured raises kNotSupported: make the
  // largest call once before capturing.
  void* scratch(size_t bytes, int slot = 0);
  static constexpr int kSlots = 6;
  // kCounters zero-initialised 32-bit counters owned by this stream.  A kernel
  // that uses them must leave them zero when it ends, except counter
  // kGeneration: the stream's launch generation, which single-launch kernels
  // read for their publication marker and advance once per launch (so their
  // markers change from launch to launch also when a hipGraph replays them).
  // Work that uses them must be serialised on the stream: graphs captured on
  // this stream and replayed on others share these counters, so such
  // replays must not overlap each other or this stream's work.
  unsigned* counters();
  static constexpr int kCounters = 4096;
  static constexpr int kGeneration = kCounters - 1;
  static constexpr int kLaunchTicket = kCounters - 2;  // a launch-wide ticket
  // At least n 32-bit completion flags owned by this stream, plus a fresh
  // epoch value.  A kernel marks item i done by storing the epoch into
  // flags[i]; since every call gets a new epoch the flags never need
  // clearing.  Zeroed (stream-ordered) when (re)allocated; the epoch
  // restarts then.  Host epochs stay below 2^31; markers drawn from the
  // launch generation have the top bit set, so the two never match each
  // other's words.  Grows like scratch(): the old area is retired, not freed.
  unsigned* epoch_flags(size_t n, unsigned* epoch);
  // Device address of this stream's timeout word: a kernel whose bounded wait
  // expires stores its (nonzero) marker there.  The word lives in pinned host
  // memory, so the host reads it after the stream's work completes
  // (take_timeout) without a copy.
  unsigned* timeout_word();
};

void record_error(int code, const std::string& msg);

// Zeroes `bytes` of device memory with a fill kernel enqueued on `stream`
// (a kernel rather than hipMemsetAsync, so the clear is also correct where a
// caller captures it in a hipGraph and replays it).
void fill_zero(hipStream_t stream, void* p, size_t bytes);

// The timeout word of `stream` on the current device (creating the stream's
// state on first use), and whether a wait on it expired since the last
// check (reads and clears the word; call once the stream's work is done).
unsigned* stream_timeout_word(hipStream_t stream);
bool take_timeout(hipStream_t stream);

// Bounded waits: every kernel unit registers a setter of its copy of the
// spin limit; set_spin_limit applies a value (polls; < 0 restores the
// default) to all of them on the current device.
void register_spin_setter(int (*setter)(long long));
int set_spin_limit(long long polls);

class Call {
 public:
  enum Mode { kSync, kAsync };
  Call(const char* name, Mode mode, void* ctxt_or_stream);
  ~Call();

  int fail_enum(const char* arg, int64_t v);
  int fail_size(const char* arg, int64_t v);
  int fail_stride(const char* arg, int64_t v);
  int fail_ld(const char* arg, int64_t v);

  // Registers window `idx`: element range [lo, hi) around p, of esize bytes each.
  void window(int idx, const void* p, int64_t lo, int64_t hi, size_t esize, bool write, bool used,
              bool result_only = false);
  bool stage();

  template <class T>
  T scalar(const T* p) {
    T v{};
    read_host(p, &v, sizeof(T));
    return v;
  }
  template <class T>
  void array(const T* p, T* out, int n) {
    read_host(p, out, sizeof(T) * n);
  }
  template <class T>
  T* dev(int idx) {
    return static_cast<T*>(wins_[idx].dev);
  }

  template <class F>
  void run(F&& f) {
    if (status_ != kOk) return;
    try {
      f();
    } catch (const Error& e) {
      fail(e.code, e.what());
    } catch (const std::exception& e) {
      fail(kHipError, e.what());
    }
  }

  bool ok() const { return status_ == kOk; }
  int status() const { return status_; }
  int finish();
  Ctx& ctx() { return ctx_; }

 private:
  struct Win {
    const char* p = nullptr;
    int64_t lo = 0, hi = 0;
    size_t esize = 0;
    bool write = false, used = false, result_only = false;
    void* dev = nullptr;
    int range = -1;
  };
  struct Range {
    const char* begin;
    const char* end;
    bool read, write;
    void* dev;
    size_t cap;
  };
  int fail(int code, const std::string& msg);
  void read_host(const void* p, void* out, size_t bytes);

  const char* name_;
  Mode mode_;
  int status_ = kOk;
  Ctx ctx_;
  std::unique_lock<std::mutex> lock_;
  Win wins_[16];
  int nwin_ = 0;
  std::vector<Range> ranges_;
  bool staged_ = false;
};

}  // namespace exoblas
