// Serving loop: continuous batching of device-resident FP32 gemm / gemv /
// trsv requests under an HBM budget (exoblas_amd_batch.h).
//
// Submitters append requests to a FIFO.  One worker thread owns `streams`
// launch slots; whenever a slot is free it admits requests from the head of
// the FIFO while the operand bytes of everything in flight stay within the
// budget (a request that can never fit is rejected at submission), packs them
// into one persistent-kernel batch and launches it on the slot's stream, so
// requests that arrive while other batches run are admitted as soon as a slot
// and memory free up and their work overlaps on the device with the earlier
// batches.  Completion is observed by polling each slot's event; a request's
// latency runs from its submission to that observation.  mode 1 launches
// every admitted request on its own through the plain exo_* device entry
// points instead (gemm as mscal + gemm, as the reference wrapper does), the
// one-launch-per-call comparison.
//
// Failure handling: a request is validated on its own at submission (a
// malformed one is rejected with its status and never joins a batch); a
// queued request can be cancelled (it is skipped at admission; work already
// launched always runs to completion, so kernels never stop early); a batch
// whose launch fails, or whose bounded waits expire (the stream's timeout
// word), marks its own requests failed and the loop carries on with the next
// ones.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <mutex>
#include <thread>
#include <vector>

#include "batch.hpp"
#include "exoblas_amd.h"
#include "exoblas_amd_async.h"
#include "runtime.hpp"

namespace exoblas {
namespace {

using Clock = std::chrono::steady_clock;

double now_us() {
  return std::chrono::duration<double, std::micro>(Clock::now().time_since_epoch()).count();
}

struct Req {
  exoblas_batch_op op;
  int64_t bytes;
  double t_submit = 0, t_done = 0;
  int state = 0;  // 0 queued, 1 in flight, 2 done, 3 failed, 4 cancelled
};

struct Slot {
  hipStream_t stream = nullptr;
  hipEvent_t done = nullptr;
  batch::DeviceBatch dev;
  batch::Built built;
  std::vector<int64_t> ids;
  std::vector<exoblas_batch_op> ops;
  int64_t bytes = 0;
  bool busy = false;
};

}  // namespace
}  // namespace exoblas

using namespace exoblas;

struct exoblas_server_s {
  std::mutex m;
  std::condition_variable cv;       // worker wake-up / drain
  std::deque<Req> reqs;             // by id
  std::deque<int64_t> fifo;         // queued ids
  int64_t budget = 0, in_flight = 0, peak = 0;
  int64_t completed = 0, rejected = 0, batches = 0, failed = 0, cancelled = 0;
  double t_first = 0, t_last = 0;
  int mode = 0, max_batch = 0;
  std::vector<Slot> slots;
  std::thread worker;
  bool stop = false;
  int device = 0;
  std::string error;

  void run();
  bool launch_slot(Slot& s);
};

// Launches the admitted requests of slot `s`; returns false on failure.
bool exoblas_server_s::launch_slot(Slot& s) {
  try {
    if (mode == 0) {
      std::string err;
      const int st = batch::build(s.ops.data(), (int64_t)s.ops.size(), s.built, err);
      if (st != kOk) throw Error(st, err);
      s.dev.upload(s.built, s.stream);
      batch::launch(s.dev, s.built, s.stream);
    } else {
      for (const auto& o : s.ops) {
        int st = kOk;
        if (o.kind == EXOBLAS_OP_GEMM) {
          const exo_win_2f32 c{o.c, {o.ldc, 1}};
          st = exoblas_smscal_rm_stride_1_async(s.stream, o.m, o.n, &o.beta, c);
          if (st == kOk) {
            const exo_win_2f32c a{o.a, {o.lda, 1}}, b{o.b, {o.ldb, 1}};
            st = exoblas_sgemm_stride_1_async(s.stream, o.trans_a, o.trans_b, o.m, o.n, o.k, &o.alpha, a, b, a, b,
                                              c);
          }
        } else if (o.kind == EXOBLAS_OP_GEMV) {
          const exo_win_2f32c a{o.a, {o.lda, 1}};
          st = exoblas_sgemv_rm_stride_any_async(s.stream, 111, o.m, o.n, &o.alpha, &o.beta, a, a,
                                                 exo_win_1f32c{o.b, {o.ldb}}, exo_win_1f32{o.c, {o.ldc}});
        } else {
          const exo_win_2f32c a{o.a, {o.lda, 1}};
          st = exoblas_strsv_rm_stride_1_async(s.stream, o.trans_b, o.trans_a, o.diag, o.m,
                                               exo_win_1f32{o.c, {o.ldc}}, a);
        }
        if (st != kOk) throw Error(st, exoblas_last_error());
      }
    }
    EXO_HIP_CHECK(hipEventRecord(s.done, s.stream));
  } catch (const Error& e) {
    std::lock_guard<std::mutex> g(m);
    error = e.what();
    return false;
  }
  return true;
}

void exoblas_server_s::run() {
  (void)hipSetDevice(device);
  for (;;) {
    bool progressed = false;
    // reap finished slots
    for (auto& s : slots) {
      if (!s.busy) continue;
      const hipError_t q = hipEventQuery(s.done);
      if (q == hipErrorNotReady) continue;
      // a bounded wait that expired in the slot's work invalidates its results
      bool timed_out = false;
      try {
        timed_out = q == hipSuccess && take_timeout(s.stream);
      } catch (const Error&) {
      }
      const bool ok = q == hipSuccess && !timed_out;
      const double t = now_us();
      std::lock_guard<std::mutex> g(m);
      for (int64_t id : s.ids) {
        reqs[id].state = ok ? 2 : 3;
        reqs[id].t_done = t;
      }
      if (!ok) {
        failed += (int64_t)s.ids.size();
        error = timed_out ? "a bounded wait expired in a batch" : hipGetErrorString(q);
      }
      completed += (int64_t)s.ids.size();
      t_last = t;
      in_flight -= s.bytes;
      s.busy = false;
      s.ids.clear();
      progressed = true;
      cv.notify_all();
    }
    // admit into free slots
    for (auto& s : slots) {
      if (s.busy) continue;
      {
        std::lock_guard<std::mutex> g(m);
        s.ops.clear();
        s.ids.clear();
        s.bytes = 0;
        while (!fifo.empty() && (int64_t)s.ids.size() < max_batch) {
          if (reqs[fifo.front()].state == 4) {  // cancelled while queued
            fifo.pop_front();
            continue;
          }
          Req& r = reqs[fifo.front()];
          if (in_flight + s.bytes + r.bytes > budget) break;  // FIFO: wait for memory
          s.bytes += r.bytes;
          s.ids.push_back(fifo.front());
          s.ops.push_back(r.op);
          r.state = 1;
          fifo.pop_front();
        }
        if (s.ids.empty()) break;
        in_flight += s.bytes;
        peak = std::max(peak, in_flight);
        ++batches;
      }
      if (launch_slot(s)) {
        s.busy = true;
      } else {
        std::lock_guard<std::mutex> g(m);
        const double t = now_us();
        for (int64_t id : s.ids) {
          reqs[id].state = 3;
          reqs[id].t_done = t;
        }
        failed += (int64_t)s.ids.size();
        completed += (int64_t)s.ids.size();
        in_flight -= s.bytes;
        s.ids.clear();
        cv.notify_all();
      }
      progressed = true;
    }
    if (!progressed) {
      std::unique_lock<std::mutex> g(m);
      bool any_busy = false;
      for (auto& s : slots) any_busy = any_busy || s.busy;
      if (stop && fifo.empty() && !any_busy) return;
      if (!any_busy && fifo.empty())
        cv.wait_for(g, std::chrono::microseconds(200));
      else
        g.unlock(), std::this_thread::yield();
    }
  }
}

extern "C" int exoblas_server_create(exoblas_server* srv, int64_t hbm_budget, int32_t streams, int32_t mode,
                                     int32_t max_batch) {
  if (!srv) return kInvalidPointer;
  if (hbm_budget <= 0 || streams < 1 || streams > 16 || (mode != 0 && mode != 1) || max_batch < 1) {
    record_error(kInvalidSize, "exoblas_server_create: invalid budget / streams / mode / max_batch");
    return kInvalidSize;
  }
  auto* s = new exoblas_server_s();
  s->budget = hbm_budget;
  s->mode = mode;
  s->max_batch = max_batch;
  (void)hipGetDevice(&s->device);
  s->slots.resize(streams);
  for (auto& sl : s->slots) {
    if (hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&sl.done, hipEventDisableTiming) != hipSuccess) {
      record_error(kHipError, "exoblas_server_create: stream / event creation failed");
      for (auto& t : s->slots) {
        if (t.stream) (void)hipStreamDestroy(t.stream);
        if (t.done) (void)hipEventDestroy(t.done);
      }
      delete s;
      return kHipError;
    }
  }
  s->worker = std::thread([s] { s->run(); });
  *srv = s;
  return kOk;
}

extern "C" int64_t exoblas_server_submit(exoblas_server s, const exoblas_batch_op* op) {
  if (!s || !op) return -kInvalidPointer;
  Req r;
  r.op = *op;
  r.bytes = batch::operand_bytes(*op);
  r.t_submit = now_us();
  // the request on its own: a malformed one never joins a batch
  std::string why;
  batch::Built probe;
  const int valid = batch::build(op, 1, probe, why);
  std::lock_guard<std::mutex> g(s->m);
  if (s->reqs.empty()) s->t_first = r.t_submit;
  const int64_t id = (int64_t)s->reqs.size();
  if (valid != kOk || r.bytes > s->budget) {
    r.state = 3;
    r.t_done = r.t_submit;
    s->reqs.push_back(r);
    ++s->rejected;
    if (valid != kOk) {
      record_error(valid, "exoblas_server_submit: " + why);
      return -valid;
    }
    record_error(kAllocError, "exoblas_server_submit: request larger than the HBM budget");
    return -kAllocError;
  }
  s->reqs.push_back(r);
  s->fifo.push_back(id);
  s->cv.notify_all();
  return id;
}

extern "C" int exoblas_server_replay(exoblas_server s, const exoblas_batch_op* ops, const double* arrival_us,
                                     int64_t count) {
  if (!s || (count > 0 && (!ops || !arrival_us))) return kInvalidPointer;
  const double t0 = now_us();
  for (int64_t i = 0; i < count; ++i) {
    while (now_us() - t0 < arrival_us[i]) {
    }
    const int64_t id = exoblas_server_submit(s, ops + i);
    if (id < 0 && id != -kAllocError) return (int)-id;
  }
  return kOk;
}

extern "C" int exoblas_server_drain(exoblas_server s) {
  if (!s) return kInvalidPointer;
  std::unique_lock<std::mutex> g(s->m);
  s->cv.wait(g, [s] { return s->completed + s->rejected + s->cancelled >= (int64_t)s->reqs.size(); });
  return s->failed ? kHipError : kOk;
}

extern "C" int exoblas_server_cancel(exoblas_server s, int64_t id) {
  if (!s) return -kInvalidPointer;
  std::lock_guard<std::mutex> g(s->m);
  if (id < 0 || id >= (int64_t)s->reqs.size()) return -kInvalidSize;
  Req& r = s->reqs[id];
  if (r.state != 0) return 0;  // launched, finished, rejected or already cancelled
  r.state = 4;
  r.t_done = now_us();
  ++s->cancelled;
  s->cv.notify_all();
  return 1;
}

extern "C" int exoblas_server_done(exoblas_server s, int64_t id) {
  if (!s) return -kInvalidPointer;
  std::lock_guard<std::mutex> g(s->m);
  if (id < 0 || id >= (int64_t)s->reqs.size()) return -kInvalidSize;
  const int st = s->reqs[id].state;
  return st == 2 ? 1 : st == 3 ? -kHipError : st == 4 ? -kCancelled : 0;
}

extern "C" int exoblas_server_stats_get(exoblas_server s, exoblas_server_stats* out) {
  if (!s || !out) return kInvalidPointer;
  std::lock_guard<std::mutex> g(s->m);
  std::vector<double> lat;
  for (const auto& r : s->reqs)
    if (r.state == 2) lat.push_back(r.t_done - r.t_submit);
  std::sort(lat.begin(), lat.end());
  std::memset(out, 0, sizeof(*out));
  out->submitted = (int64_t)s->reqs.size();
  out->completed = s->completed;
  out->rejected = s->rejected;
  out->batches = s->batches;
  out->peak_bytes_in_flight = s->peak;
  out->cancelled = s->cancelled;
  out->failed = s->failed;
  if (!lat.empty()) {
    auto pct = [&](double p) {
      const size_t i = (size_t)std::min<double>(lat.size() - 1, std::max(0.0, p * lat.size() - 1e-9));
      return lat[i];
    };
    out->p50_us = pct(0.50);
    out->p99_us = pct(0.99);
    double sum = 0;
    for (double v : lat) sum += v;
    out->mean_us = sum / lat.size();
    out->max_us = lat.back();
    out->wall_s = (s->t_last - s->t_first) * 1e-6;
  }
  return s->failed ? kHipError : kOk;
}

extern "C" int64_t exoblas_server_latencies(exoblas_server s, double* out, int64_t cap) {
  if (!s || !out) return -kInvalidPointer;
  std::lock_guard<std::mutex> g(s->m);
  int64_t n = 0;
  for (const auto& r : s->reqs) {
    if (n >= cap) break;
    out[n++] = r.state == 2 ? r.t_done - r.t_submit : -1.0;
  }
  return n;
}

extern "C" int exoblas_server_destroy(exoblas_server s) {
  if (!s) return kOk;
  {
    std::lock_guard<std::mutex> g(s->m);
    s->stop = true;
    s->cv.notify_all();
  }
  if (s->worker.joinable()) s->worker.join();
  for (auto& sl : s->slots) {
    (void)hipStreamSynchronize(sl.stream);
    sl.dev.release();
    (void)exoblas_stream_release(sl.stream);  // the library's state for the slot's stream
    (void)hipStreamDestroy(sl.stream);
    (void)hipEventDestroy(sl.done);
  }
  delete s;
  return kOk;
}
