// Host-side copy helper of the staging path: a few persistent threads that
// split one memcpy between them (plain C++; no HIP).
#pragma once

#include <algorithm>
#include <condition_variable>
#include <cstdint>
#include <cstring>
#include <mutex>
#include <thread>
#include <vector>

namespace exoblas {

// A few persistent threads that split a host memcpy between them.
class CopyPool {
 public:
  explicit CopyPool(int nthreads) : n_(std::max(1, nthreads)) {
    for (int i = 1; i < n_; ++i) threads_.emplace_back([this, i] { loop(i); });
  }
  ~CopyPool() {
    {
      std::lock_guard<std::mutex> g(m_);
      stop_ = true;
    }
    cv_.notify_all();
    for (auto& t : threads_) t.join();
  }
  // dst[0, bytes) = src[0, bytes), the calling thread taking part 0.  Copies
  // from several host threads (calls on different devices) run one at a time.
  void copy(void* dst, const void* src, size_t bytes) {
    std::lock_guard<std::mutex> serial(serial_);
    {
      std::lock_guard<std::mutex> g(m_);
      dst_ = static_cast<char*>(dst);
      src_ = static_cast<const char*>(src);
      bytes_ = bytes;
      pending_ = n_ - 1;
      ++gen_;
    }
    cv_.notify_all();
    part(0);
    std::unique_lock<std::mutex> g(m_);
    done_.wait(g, [this] { return pending_ == 0; });
  }

 private:
  void part(int i) {
    const size_t per = ((bytes_ + n_ - 1) / n_ + 63) & ~size_t(63);  // n_ parts cover bytes_
    const size_t b = std::min(bytes_, per * i), e = std::min(bytes_, b + per);
    if (e > b) std::memcpy(dst_ + b, src_ + b, e - b);
  }
  void loop(int i) {
    uint64_t seen = 0;
    for (;;) {
      {
        std::unique_lock<std::mutex> g(m_);
        cv_.wait(g, [&] { return stop_ || gen_ != seen; });
        if (stop_) return;
        seen = gen_;
      }
      part(i);
      {
        std::lock_guard<std::mutex> g(m_);
        if (--pending_ == 0) done_.notify_one();
      }
    }
  }
  int n_;
  std::vector<std::thread> threads_;
  std::mutex serial_;  // one copy at a time
  std::mutex m_;
  std::condition_variable cv_, done_;
  char* dst_ = nullptr;
  const char* src_ = nullptr;
  size_t bytes_ = 0;
  int pending_ = 0;
  uint64_t gen_ = 0;
  bool stop_ = false;
};

}  // namespace exoblas
