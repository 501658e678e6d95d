/* exoblas_amd batched execution: many uneven FP32 gemm / gemv / trsv problems
 * in one persistent kernel launch, and the serving loop built on it.
 *
 * No counterpart in the reference, whose exo_* functions run one problem per
 * call on one CPU thread.  All pointers are device memory.  Return values are
 * 0 or an EXOBLAS_STATUS_* code (exoblas_amd.h).
 */
#ifndef EXOBLAS_AMD_BATCH_H
#define EXOBLAS_AMD_BATCH_H

#include <hip/hip_runtime_api.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define EXOBLAS_OP_GEMM 1 /* C = alpha op(A) op(B) + beta C   (row-major; C read for every beta)  */
#define EXOBLAS_OP_GEMV 2 /* y = alpha A x + beta y           (A m x n row-major, NoTrans; y too) */
#define EXOBLAS_OP_TRSV 3 /* x = op(A)^-1 x                   (A n x n row-major, one triangle)  */

typedef struct exoblas_batch_op {
  int32_t kind;    /* EXOBLAS_OP_* */
  int32_t trans_a; /* gemm TransA / trsv TransA (111, 112); gemv: 111 */
  int32_t trans_b; /* gemm TransB (111, 112); trsv Uplo (121, 122) */
  int32_t diag;    /* trsv Diag (131, 132) */
  int64_t m, n, k; /* gemm M, N, K; gemv rows m, columns n; trsv order n */
  const float *a;  /* gemm A / gemv A / trsv A */
  int64_t lda;
  const float *b;  /* gemm B / gemv x */
  int64_t ldb;     /* gemm ldb / gemv incx (1) */
  float *c;        /* gemm C / gemv y / trsv x */
  int64_t ldc;     /* gemm ldc / gemv incy / trsv incx (1) */
  float alpha, beta;
} exoblas_batch_op;

/* Runs `count` problems with one persistent kernel on `stream` (the
 * descriptors are copied before the call returns). */
int exoblas_batch_run(hipStream_t stream, const exoblas_batch_op *ops, int64_t count);

/* The work queue a batch of these problems would run, without running it
 * (host only; the operand pointers are not dereferenced): *n_items entries
 * order[q] = (problem << 32) | item of that problem, and kinds[p] the
 * EXOBLAS_OP_* of problem p of the queue's numbering (*n_problems of them).
 * The arrays are filled when both counts fit in cap. */
int exoblas_batch_queue(const exoblas_batch_op *ops, int64_t count, uint64_t *order, int32_t *kinds, int64_t cap,
                        int64_t *n_items, int64_t *n_problems);

/* A fixed batch with device-resident descriptors: launching it is one kernel
 * launch with no copies or memsets, so it can be captured in a hipGraph and
 * replayed.  Launches of one plan must not overlap in time. */
typedef struct exoblas_batch_plan_s *exoblas_batch_plan;
int exoblas_batch_plan_create(exoblas_batch_plan *plan, const exoblas_batch_op *ops, int64_t count);
int exoblas_batch_plan_launch(exoblas_batch_plan plan, hipStream_t stream);
int exoblas_batch_plan_destroy(exoblas_batch_plan plan);

/* gemm tile choice per size class, written by the autotuner
 * (exoblas_amd/autotune.py).  table[kc * 36 + mc * 6 + nc] is a tile id
 * (0: 32 x 32, 1: 64 x 64, 2: 128 x 128), mc / nc the classes of M / N
 * (<= 32, <= 64, <= 128, <= 256, <= 512, larger), kc the class of K
 * (<= 256, larger).  len must be 72. */
int exoblas_batch_set_tuning(const int32_t *table, int64_t len);
int exoblas_batch_get_tuning(int32_t *table, int64_t len);
/* Forces every gemm onto one tile id (-1 restores the table); for tuning. */
int exoblas_batch_force_tile(int32_t tile);

/* ---- serving loop --------------------------------------------------------
 * A host thread admits submitted problems while the bytes of their operands
 * in flight stay within `hbm_budget`, packs whatever is admitted into a batch
 * as soon as one of `streams` launch slots is free (continuous batching: new
 * requests join the next launch while earlier ones still run) and records
 * per-request latency.  mode 0 = batched persistent kernels, mode 1 = one
 * kernel launch per request (the comparison). */
typedef struct exoblas_server_s *exoblas_server;
typedef struct exoblas_server_stats {
  int64_t submitted, completed, rejected, batches;
  double p50_us, p99_us, mean_us, max_us; /* submit -> completion observed */
  double wall_s;                          /* first submit -> last completion */
  int64_t peak_bytes_in_flight;
  int64_t cancelled; /* cancelled while queued (exoblas_server_cancel) */
  int64_t failed;    /* launched in a batch that failed (launch error or an expired wait) */
} exoblas_server_stats;

int exoblas_server_create(exoblas_server *srv, int64_t hbm_budget, int32_t streams, int32_t mode,
                          int32_t max_batch);
/* Queues one problem; returns its request id (>= 0) or -status.  The problem
 * is validated on its own first (a malformed one is rejected with its
 * status: -EXOBLAS_STATUS_INVALID_*); a problem larger than the whole budget
 * is rejected (EXOBLAS_STATUS_ALLOC_ERROR).  Rejected requests still take an
 * id. */
int64_t exoblas_server_submit(exoblas_server srv, const exoblas_batch_op *op);
/* Submits ops[i] at arrival_us[i] microseconds after the call starts (times
 * ascending), pacing from the calling thread; returns once all are submitted.
 * Request ids are assigned in order.  Returns 0 or a status. */
int exoblas_server_replay(exoblas_server srv, const exoblas_batch_op *ops, const double *arrival_us, int64_t count);
/* Blocks until every submitted request has completed, been rejected or been
 * cancelled. */
int exoblas_server_drain(exoblas_server srv);
/* Cancels request `id` if it is still queued: returns 1 (it will never run
 * and its output is not touched), 0 if it was already launched, finished,
 * rejected or cancelled (launched work always runs to completion), -status
 * for an unknown id. */
int exoblas_server_cancel(exoblas_server srv, int64_t id);
/* 1 if request `id` has completed, 0 if not, -EXOBLAS_STATUS_CANCELLED if it
 * was cancelled, -EXOBLAS_STATUS_HIP_ERROR if it failed or was rejected. */
int exoblas_server_done(exoblas_server srv, int64_t id);
int exoblas_server_stats_get(exoblas_server srv, exoblas_server_stats *out);
/* Latency of each completed request in submission order (us); returns the
 * number written. */
int64_t exoblas_server_latencies(exoblas_server srv, double *out, int64_t cap);
int exoblas_server_destroy(exoblas_server srv);

#ifdef __cplusplus
}
#endif

#endif /* EXOBLAS_AMD_BATCH_H */
