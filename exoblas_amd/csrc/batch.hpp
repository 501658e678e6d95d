// Batched execution: problem descriptors, the work-item decomposition and the
// device buffers of one batch (shared by the batch API and the serving loop).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "exoblas_amd_batch.h"

namespace exoblas {
namespace batch {

// One problem as the kernel sees it.
struct Desc {
  int32_t kind, ta, tb, diag;
  int64_t m, n, k;
  const float* a;
  int64_t lda;
  const float* b;
  int64_t ldb;
  float* c;
  int64_t ldc;
  float alpha, beta;
  int32_t tile;        // gemm tile id
  int32_t tiles_n;     // gemm tiles along N
  uint32_t flag_off;   // trsv: offset of its publication words
  uint32_t pad_;
};

// ctrl[0] work queue, ctrl[1] finished workgroups, ctrl[2] launch count,
// ctrl[3] unused, ctrl[kCtrlWords + i] trsv block flags.  Bounded waits
// report to the timeout word of the stream the batch runs on.
constexpr int kCtrlWords = 4;
constexpr int kRows = 64;  // gemv / trsv rows per work item

// order[q] = (problem index << 32) | item index within the problem: the
// work queue, taken front to back by the persistent kernel.
struct Built {
  std::vector<Desc> descs;
  std::vector<uint64_t> order;
  uint32_t total_items = 0;
  uint32_t flag_words = 0;
};

// Validates `count` problems, splits them into work items and orders the
// queue: the trsv chains' steps round by round (step r of every chain, the
// longest chains first), with an even share of the gemm / gemv items (the
// most expensive first) after each round.  Returns 0 or a status, with `err`
// set.
int build(const exoblas_batch_op* ops, int64_t count, Built& out, std::string& err);

// Bytes of device memory a problem's operands occupy (admission accounting).
int64_t operand_bytes(const exoblas_batch_op& op);

// Device-side buffers of one batch slot.
struct DeviceBatch {
  Desc* desc = nullptr;
  uint64_t* order = nullptr;
  unsigned* ctrl = nullptr;
  size_t cap_desc = 0, cap_order = 0, cap_flags = 0;
  void release();
  // Grows the buffers (synchronising `stream` first if they are in use) and
  // enqueues the copies of `b` on `stream`.
  void upload(const Built& b, hipStream_t stream);
};

// Enqueues the persistent kernel for an uploaded batch (timeouts go to the
// stream's timeout word: take_timeout(stream) after completion).
void launch(const DeviceBatch& d, const Built& b, hipStream_t stream);

}  // namespace batch
}  // namespace exoblas
