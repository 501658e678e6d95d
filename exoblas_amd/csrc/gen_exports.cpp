// Generated by exoblas_amd/generator.py -- do not edit.
// extern "C" definitions of every exo_* symbol and its _async form.
#include "runtime.hpp"
#include "ops.hpp"
#include "exo_asum.h"
#include "exo_axpy.h"
#include "exo_dot.h"
#include "exo_dsdot.h"
#include "exo_copy.h"
#include "exo_swap.h"
#include "exo_rot.h"
#include "exo_rotm.h"
#include "exo_scal.h"
#include "exo_gemv.h"
#include "exo_ger.h"
#include "exo_symv.h"
#include "exo_syr.h"
#include "exo_syr2.h"
#include "exo_trmv.h"
#include "exo_trsv.h"
#include "exo_gbmv.h"
#include "exo_sbmv.h"
#include "exo_tbmv.h"
#include "exo_tbsv.h"
#include "exo_gemm.h"
#include "exo_mscal.h"
#include "exo_symm.h"
#include "exo_syrk.h"
#include "exo_syr2k.h"
#include "exo_trmm.h"
#include "exoblas_amd_async.h"
#include "exoblas_amd_ext.h"

// src/level1/asum.py:11-15
extern "C" void exo_sasum_stride_any(void *ctxt, int_fast32_t n, struct exo_win_1f32c x, float *result) {
  exoblas::Call call("exo_sasum_stride_any", exoblas::Call::kSync, ctxt);
  if (n < 0) return (void)call.fail_size("n", n);
  call.window(0, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(float), false, true);
  call.window(1, result, 0, 1, sizeof(float), true, true, true);
  if (!call.stage()) return (void)call.status();
  if (!call.ok()) return (void)call.status();
  call.run([&] { exoblas::ops::asum<float>(call.ctx(), (int64_t)n, call.dev<const float>(0), (int64_t)x.strides[0], call.dev<float>(1)); });
  return (void)call.finish();
}
extern "C" int exoblas_sasum_stride_any_async(hipStream_t stream, int_fast32_t n, struct exo_win_1f32c x, float *result) {
  exoblas::Call call("exoblas_sasum_stride_any_async", exoblas::Call::kAsync, stream);
  if (n < 0) return call.fail_size("n", n);
  call.window(0, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(float), false, true);
  call.window(1, result, 0, 1, sizeof(float), true, true, true);
  if (!call.stage()) return call.status();
  if (!call.ok()) return call.status();
  call.run([&] { exoblas::ops::asum<float>(call.ctx(), (int64_t)n, call.dev<const float>(0), (int64_t)x.strides[0], call.dev<float>(1)); });
  return call.finish();
}

// src/level1/asum.py:11-15
extern "C" void exo_sasum_stride_1(void *ctxt, int_fast32_t n, struct exo_win_1f32c x, float *result) {
  exoblas::Call call("exo_sasum_stride_1", exoblas::Call::kSync, ctxt);
  if (n < 0) return (void)call.fail_size("n", n);
  call.window(0, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(float), false, true);
  call.window(1, result, 0, 1, sizeof(float), true, true, true);
  if (!call.stage()) return (void)call.status();
  if (!call.ok()) return (void)call.status();
  call.run([&] { exoblas::ops::asum<float>(call.ctx(), (int64_t)n, call.dev<const float>(0), (int64_t)x.strides[0], call.dev<float>(1)); });
  return (void)call.finish();
}
extern "C" int exoblas_sasum_stride_1_async(hipStream_t stream, int_fast32_t n, struct exo_win_1f32c x, float *result) {
  exoblas::Call call("exoblas_sasum_stride_1_async", exoblas::Call::kAsync, stream);
  if (n < 0) return call.fail_size("n", n);
  call.window(0, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(float), false, true);
  call.window(1, result, 0, 1, sizeof(float), true, true, true);
  if (!call.stage()) return call.status();
  if (!call.ok()) return call.status();
  call.run([&] { exoblas::ops::asum<float>(call.ctx(), (int64_t)n, call.dev<const float>(0), (int64_t)x.strides[0], call.dev<float>(1)); });
  return call.finish();
}

// src/level1/asum.py:11-15
extern "C" void exo_dasum_stride_any(void *ctxt, int_fast32_t n, struct exo_win_1f64c x, double *result) {
  exoblas::Call call("exo_dasum_stride_any", exoblas::Call::kSync, ctxt);
  if (n < 0) return (void)call.fail_size("n", n);
  call.window(0, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(double), false, true);
  call.window(1, result, 0, 1, sizeof(double), true, true, true);
  if (!call.stage()) return (void)call.status();
  if (!call.ok()) return (void)call.status();
  call.run([&] { exoblas::ops::asum<double>(call.ctx(), (int64_t)n, call.dev<const double>(0), (int64_t)x.strides[0], call.dev<double>(1)); });
  return (void)call.finish();
}
extern "C" int exoblas_dasum_stride_any_async(hipStream_t stream, int_fast32_t n, struct exo_win_1f64c x, double *result) {
  exoblas::Call call("exoblas_dasum_stride_any_async", exoblas::Call::kAsync, stream);
  if (n < 0) return call.fail_size("n", n);
  call.window(0, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(double), false, true);
  call.window(1, result, 0, 1, sizeof(double), true, true, true);
  if (!call.stage()) return call.status();
  if (!call.ok()) return call.status();
  call.run([&] { exoblas::ops::asum<double>(call.ctx(), (int64_t)n, call.dev<const double>(0), (int64_t)x.strides[0], call.dev<double>(1)); });
  return call.finish();
}

// src/level1/asum.py:11-15
extern "C" void exo_dasum_stride_1(void *ctxt, int_fast32_t n, struct exo_win_1f64c x, double *result) {
  exoblas::Call call("exo_dasum_stride_1", exoblas::Call::kSync, ctxt);
  if (n < 0) return (void)call.fail_size("n", n);
  call.window(0, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(double), false, true);
  call.window(1, result, 0, 1, sizeof(double), true, true, true);
  if (!call.stage()) return (void)call.status();
  if (!call.ok()) return (void)call.status();
  call.run([&] { exoblas::ops::asum<double>(call.ctx(), (int64_t)n, call.dev<const double>(0), (int64_t)x.strides[0], call.dev<double>(1)); });
  return (void)call.finish();
}
extern "C" int exoblas_dasum_stride_1_async(hipStream_t stream, int_fast32_t n, struct exo_win_1f64c x, double *result) {
  exoblas::Call call("exoblas_dasum_stride_1_async", exoblas::Call::kAsync, stream);
  if (n < 0) return call.fail_size("n", n);
  call.window(0, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(double), false, true);
  call.window(1, result, 0, 1, sizeof(double), true, true, true);
  if (!call.stage()) return call.status();
  if (!call.ok()) return call.status();
  call.run([&] { exoblas::ops::asum<double>(call.ctx(), (int64_t)n, call.dev<const double>(0), (int64_t)x.strides[0], call.dev<double>(1)); });
  return call.finish();
}

// src/level1/axpy.py:9-12
extern "C" void exo_saxpy_stride_any(void *ctxt, int_fast32_t n, const float *alpha, struct exo_win_1f32c x, struct exo_win_1f32 y) {
  exoblas::Call call("exo_saxpy_stride_any", exoblas::Call::kSync, ctxt);
  if (n < 0) return (void)call.fail_size("n", n);
  call.window(0, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(float), false, true);
  call.window(1, y.data, exoblas::span_lo(n, y.strides[0]), exoblas::span_hi(n, y.strides[0]), sizeof(float), true, true);
  if (!call.stage()) return (void)call.status();
  const float alpha_v = call.scalar(alpha);
  if (!call.ok()) return (void)call.status();
  call.run([&] { exoblas::ops::axpy<float>(call.ctx(), (int64_t)n, alpha_v, call.dev<const float>(0), (int64_t)x.strides[0], call.dev<float>(1), (int64_t)y.strides[0]); });
  return (void)call.finish();
}
extern "C" int exoblas_saxpy_stride_any_async(hipStream_t stream, int_fast32_t n, const float *alpha, struct exo_win_1f32c x, struct exo_win_1f32 y) {
  exoblas::Call call("exoblas_saxpy_stride_any_async", exoblas::Call::kAsync, stream);
  if (n < 0) return call.fail_size("n", n);
  call.window(0, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(float), false, true);
  call.window(1, y.data, exoblas::span_lo(n, y.strides[0]), exoblas::span_hi(n, y.strides[0]), sizeof(float), true, true);
  if (!call.stage()) return call.status();
  const float alpha_v = call.scalar(alpha);
  if (!call.ok()) return call.status();
  call.run([&] { exoblas::ops::axpy<float>(call.ctx(), (int64_t)n, alpha_v, call.dev<const float>(0), (int64_t)x.strides[0], call.dev<float>(1), (int64_t)y.strides[0]); });
  return call.finish();
}

// src/level1/axpy.py:9-12
extern "C" void exo_saxpy_stride_1(void *ctxt, int_fast32_t n, const float *alpha, struct exo_win_1f32c x, struct exo_win_1f32 y) {
  exoblas::Call call("exo_saxpy_stride_1", exoblas::Call::kSync, ctxt);
  if (n < 0) return (void)call.fail_size("n", n);
  call.window(0, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(float), false, true);
  call.window(1, y.data, exoblas::span_lo(n, y.strides[0]), exoblas::span_hi(n, y.strides[0]), sizeof(float), true, true);
  if (!call.stage()) return (void)call.status();
  const float alpha_v = call.scalar(alpha);
  if (!call.ok()) return (void)call.status();
  call.run([&] { exoblas::ops::axpy<float>(call.ctx(), (int64_t)n, alpha_v, call.dev<const float>(0), (int64_t)x.strides[0], call.dev<float>(1), (int64_t)y.strides[0]); });
  return (void)call.finish();
}
extern "C" int exoblas_saxpy_stride_1_async(hipStream_t stream, int_fast32_t n, const float *alpha, struct exo_win_1f32c x, struct exo_win_1f32 y) {
  exoblas::Call call("exoblas_saxpy_stride_1_async", exoblas::Call::kAsync, stream);
  if (n < 0) return call.fail_size("n", n);
  call.window(0, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(float), false, true);
  call.window(1, y.data, exoblas::span_lo(n, y.strides[0]), exoblas::span_hi(n, y.strides[0]), sizeof(float), true, true);
  if (!call.stage()) return call.status();
  const float alpha_v = call.scalar(alpha);
  if (!call.ok()) return call.status();
  call.run([&] { exoblas::ops::axpy<float>(call.ctx(), (int64_t)n, alpha_v, call.dev<const float>(0), (int64_t)x.strides[0], call.dev<float>(1), (int64_t)y.strides[0]); });
  return call.finish();
}

// src/level1/axpy.py:9-12
extern "C" void exo_daxpy_stride_any(void *ctxt, int_fast32_t n, const double *alpha, struct exo_win_1f64c x, struct exo_win_1f64 y) {
  exoblas::Call call("exo_daxpy_stride_any", exoblas::Call::kSync, ctxt);
  if (n < 0) return (void)call.fail_size("n", n);
  call.window(0, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(double), false, true);
  call.window(1, y.data, exoblas::span_lo(n, y.strides[0]), exoblas::span_hi(n, y.strides[0]), sizeof(double), true, true);
  if (!call.stage()) return (void)call.status();
  const double alpha_v = call.scalar(alpha);
  if (!call.ok()) return (void)call.status();
  call.run([&] { exoblas::ops::axpy<double>(call.ctx(), (int64_t)n, alpha_v, call.dev<const double>(0), (int64_t)x.strides[0], call.dev<double>(1), (int64_t)y.strides[0]); });
  return (void)call.finish();
}
extern "C" int exoblas_daxpy_stride_any_async(hipStream_t stream, int_fast32_t n, const double *alpha, struct exo_win_1f64c x, struct exo_win_1f64 y) {
  exoblas::Call call("exoblas_daxpy_stride_any_async", exoblas::Call::kAsync, stream);
  if (n < 0) return call.fail_size("n", n);
  call.window(0, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(double), false, true);
  call.window(1, y.data, exoblas::span_lo(n, y.strides[0]), exoblas::span_hi(n, y.strides[0]), sizeof(double), true, true);
  if (!call.stage()) return call.status();
  const double alpha_v = call.scalar(alpha);
  if (!call.ok()) return call.status();
  call.run([&] { exoblas::ops::axpy<double>(call.ctx(), (int64_t)n, alpha_v, call.dev<const double>(0), (int64_t)x.strides[0], call.dev<double>(1), (int64_t)y.strides[0]); });
  return call.finish();
}

// src/level1/axpy.py:9-12
extern "C" void exo_daxpy_stride_1(void *ctxt, int_fast32_t n, const double *alpha, struct exo_win_1f64c x, struct exo_win_1f64 y) {
  exoblas::Call call("exo_daxpy_stride_1", exoblas::Call::kSync, ctxt);
  if (n < 0) return (void)call.fail_size("n", n);
  call.window(0, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(double), false, true);
  call.window(1, y.data, exoblas::span_lo(n, y.strides[0]), exoblas::span_hi(n, y.strides[0]), sizeof(double), true, true);
  if (!call.stage()) return (void)call.status();
  const double alpha_v = call.scalar(alpha);
  if (!call.ok()) return (void)call.status();
  call.run([&] { exoblas::ops::axpy<double>(call.ctx(), (int64_t)n, alpha_v, call.dev<const double>(0), (int64_t)x.strides[0], call.dev<double>(1), (int64_t)y.strides[0]); });
  return (void)call.finish();
}
extern "C" int exoblas_daxpy_stride_1_async(hipStream_t stream, int_fast32_t n, const double *alpha, struct exo_win_1f64c x, struct exo_win_1f64 y) {
  exoblas::Call call("exoblas_daxpy_stride_1_async", exoblas::Call::kAsync, stream);
  if (n < 0) return call.fail_size("n", n);
  call.window(0, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(double), false, true);
  call.window(1, y.data, exoblas::span_lo(n, y.strides[0]), exoblas::span_hi(n, y.strides[0]), sizeof(double), true, true);
  if (!call.stage()) return call.status();
  const double alpha_v = call.scalar(alpha);
  if (!call.ok()) return call.status();
  call.run([&] { exoblas::ops::axpy<double>(call.ctx(), (int64_t)n, alpha_v, call.dev<const double>(0), (int64_t)x.strides[0], call.dev<double>(1), (int64_t)y.strides[0]); });
  return call.finish();
}

// src/level1/axpy.py:15-18
extern "C" void exo_saxpy_alpha_1_stride_any(void *ctxt, int_fast32_t n, struct exo_win_1f32c x, struct exo_win_1f32 y) {
  exoblas::Call call("exo_saxpy_alpha_1_stride_any", exoblas::Call::kSync, ctxt);
  if (n < 0) return (void)call.fail_size("n", n);
  call.window(0, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(float), false, true);
  call.window(1, y.data, exoblas::span_lo(n, y.strides[0]), exoblas::span_hi(n, y.strides[0]), sizeof(float), true, true);
  if (!call.stage()) return (void)call.status();
  if (!call.ok()) return (void)call.status();
  call.run([&] { exoblas::ops::axpy_alpha_1<float>(call.ctx(), (int64_t)n, call.dev<const float>(0), (int64_t)x.strides[0], call.dev<float>(1), (int64_t)y.strides[0]); });
  return (void)call.finish();
}
extern "C" int exoblas_saxpy_alpha_1_stride_any_async(hipStream_t stream, int_fast32_t n, struct exo_win_1f32c x, struct exo_win_1f32 y) {
  exoblas::Call call("exoblas_saxpy_alpha_1_stride_any_async", exoblas::Call::kAsync, stream);
  if (n < 0) return call.fail_size("n", n);
  call.window(0, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(float), false, true);
  call.window(1, y.data, exoblas::span_lo(n, y.strides[0]), exoblas::span_hi(n, y.strides[0]), sizeof(float), true, true);
  if (!call.stage()) return call.status();
  if (!call.ok()) return call.status();
  call.run([&] { exoblas::ops::axpy_alpha_1<float>(call.ctx(), (int64_t)n, call.dev<const float>(0), (int64_t)x.strides[0], call.dev<float>(1), (int64_t)y.strides[0]); });
  return call.finish();
}

// src/level1/axpy.py:15-18
extern "C" void exo_saxpy_alpha_1_stride_1(void *ctxt, int_fast32_t n, struct exo_win_1f32c x, struct exo_win_1f32 y) {
  exoblas::Call call("exo_saxpy_alpha_1_stride_1", exoblas::Call::kSync, ctxt);
  if (n < 0) return (void)call.fail_size("n", n);
  call.window(0, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(float), false, true);
  call.window(1, y.data, exoblas::span_lo(n, y.strides[0]), exoblas::span_hi(n, y.strides[0]), sizeof(float), true, true);
  if (!call.stage()) return (void)call.status();
  if (!call.ok()) return (void)call.status();
  call.run([&] { exoblas::ops::axpy_alpha_1<float>(call.ctx(), (int64_t)n, call.dev<const float>(0), (int64_t)x.strides[0], call.dev<float>(1), (int64_t)y.strides[0]); });
  return (void)call.finish();
}
extern "C" int exoblas_saxpy_alpha_1_stride_1_async(hipStream_t stream, int_fast32_t n, struct exo_win_1f32c x, struct exo_win_1f32 y) {
  exoblas::Call call("exoblas_saxpy_alpha_1_stride_1_async", exoblas::Call::kAsync, stream);
  if (n < 0) return call.fail_size("n", n);
  call.window(0, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(float), false, true);
  call.window(1, y.data, exoblas::span_lo(n, y.strides[0]), exoblas::span_hi(n, y.strides[0]), sizeof(float), true, true);
  if (!call.stage()) return call.status();
  if (!call.ok()) return call.status();
  call.run([&] { exoblas::ops::axpy_alpha_1<float>(call.ctx(), (int64_t)n, call.dev<const float>(0), (int64_t)x.strides[0], call.dev<float>(1), (int64_t)y.strides[0]); });
  return call.finish();
}

// src/level1/axpy.py:15-18
extern "C" void exo_daxpy_alpha_1_stride_any(void *ctxt, int_fast32_t n, struct exo_win_1f64c x, struct exo_win_1f64 y) {
  exoblas::Call call("exo_daxpy_alpha_1_stride_any", exoblas::Call::kSync, ctxt);
  if (n < 0) return (void)call.fail_size("n", n);
  call.window(0, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(double), false, true);
  call.window(1, y.data, exoblas::span_lo(n, y.strides[0]), exoblas::span_hi(n, y.strides[0]), sizeof(double), true, true);
  if (!call.stage()) return (void)call.status();
  if (!call.ok()) return (void)call.status();
  call.run([&] { exoblas::ops::axpy_alpha_1<double>(call.ctx(), (int64_t)n, call.dev<const double>(0), (int64_t)x.strides[0], call.dev<double>(1), (int64_t)y.strides[0]); });
  return (void)call.finish();
}
extern "C" int exoblas_daxpy_alpha_1_stride_any_async(hipStream_t stream, int_fast32_t n, struct exo_win_1f64c x, struct exo_win_1f64 y) {
  exoblas::Call call("exoblas_daxpy_alpha_1_stride_any_async", exoblas::Call::kAsync, stream);
  if (n < 0) return call.fail_size("n", n);
  call.window(0, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(double), false, true);
  call.window(1, y.data, exoblas::span_lo(n, y.strides[0]), exoblas::span_hi(n, y.strides[0]), sizeof(double), true, true);
  if (!call.stage()) return call.status();
  if (!call.ok()) return call.status();
  call.run([&] { exoblas::ops::axpy_alpha_1<double>(call.ctx(), (int64_t)n, call.dev<const double>(0), (int64_t)x.strides[0], call.dev<double>(1), (int64_t)y.strides[0]); });
  return call.finish();
}

// src/level1/axpy.py:15-18
extern "C" void exo_daxpy_alpha_1_stride_1(void *ctxt, int_fast32_t n, struct exo_win_1f64c x, struct exo_win_1f64 y) {
  exoblas::Call call("exo_daxpy_alpha_1_stride_1", exoblas::Call::kSync, ctxt);
  if (n < 0) return (void)call.fail_size("n", n);
  call.window(0, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(double), false, true);
  call.window(1, y.data, exoblas::span_lo(n, y.strides[0]), exoblas::span_hi(n, y.strides[0]), sizeof(double), true, true);
  if (!call.stage()) return (void)call.status();
  if (!call.ok()) return (void)call.status();
  call.run([&] { exoblas::ops::axpy_alpha_1<double>(call.ctx(), (int64_t)n, call.dev<const double>(0), (int64_t)x.strides[0], call.dev<double>(1), (int64_t)y.strides[0]); });
  return (void)call.finish();
}
extern "C" int exoblas_daxpy_alpha_1_stride_1_async(hipStream_t stream, int_fast32_t n, struct exo_win_1f64c x, struct exo_win_1f64 y) {
  exoblas::Call call("exoblas_daxpy_alpha_1_stride_1_async", exoblas::Call::kAsync, stream);
  if (n < 0) return call.fail_size("n", n);
  call.window(0, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(double), false, true);
  call.window(1, y.data, exoblas::span_lo(n, y.strides[0]), exoblas::span_hi(n, y.strides[0]), sizeof(double), true, true);
  if (!call.stage()) return call.status();
  if (!call.ok()) return call.status();
  call.run([&] { exoblas::ops::axpy_alpha_1<double>(call.ctx(), (int64_t)n, call.dev<const double>(0), (int64_t)x.strides[0], call.dev<double>(1), (int64_t)y.strides[0]); });
  return call.finish();
}

// src/level1/dot.py:9-13
extern "C" void exo_sdot_stride_any(void *ctxt, int_fast32_t n, struct exo_win_1f32c x, struct exo_win_1f32c y, float *result) {
  exoblas::Call call("exo_sdot_stride_any", exoblas::Call::kSync, ctxt);
  if (n < 0) return (void)call.fail_size("n", n);
  call.window(0, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(float), false, true);
  call.window(1, y.data, exoblas::span_lo(n, y.strides[0]), exoblas::span_hi(n, y.strides[0]), sizeof(float), false, true);
  call.window(2, result, 0, 1, sizeof(float), true, true, true);
  if (!call.stage()) return (void)call.status();
  if (!call.ok()) return (void)call.status();
  call.run([&] { exoblas::ops::dot<float>(call.ctx(), (int64_t)n, call.dev<const float>(0), (int64_t)x.strides[0], call.dev<const float>(1), (int64_t)y.strides[0], call.dev<float>(2)); });
  return (void)call.finish();
}
extern "C" int exoblas_sdot_stride_any_async(hipStream_t stream, int_fast32_t n, struct exo_win_1f32c x, struct exo_win_1f32c y, float *result) {
  exoblas::Call call("exoblas_sdot_stride_any_async", exoblas::Call::kAsync, stream);
  if (n < 0) return call.fail_size("n", n);
  call.window(0, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(float), false, true);
  call.window(1, y.data, exoblas::span_lo(n, y.strides[0]), exoblas::span_hi(n, y.strides[0]), sizeof(float), false, true);
  call.window(2, result, 0, 1, sizeof(float), true, true, true);
  if (!call.stage()) return call.status();
  if (!call.ok()) return call.status();
  call.run([&] { exoblas::ops::dot<float>(call.ctx(), (int64_t)n, call.dev<const float>(0), (int64_t)x.strides[0], call.dev<const float>(1), (int64_t)y.strides[0], call.dev<float>(2)); });
  return call.finish();
}

// src/level1/dot.py:9-13
extern "C" void exo_sdot_stride_1(void *ctxt, int_fast32_t n, struct exo_win_1f32c x, struct exo_win_1f32c y, float *result) {
  exoblas::Call call("exo_sdot_stride_1", exoblas::Call::kSync, ctxt);
  if (n < 0) return (void)call.fail_size("n", n);
  call.window(0, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(float), false, true);
  call.window(1, y.data, exoblas::span_lo(n, y.strides[0]), exoblas::span_hi(n, y.strides[0]), sizeof(float), false, true);
  call.window(2, result, 0, 1, sizeof(float), true, true, true);
  if (!call.stage()) return (void)call.status();
  if (!call.ok()) return (void)call.status();
  call.run([&] { exoblas::ops::dot<float>(call.ctx(), (int64_t)n, call.dev<const float>(0), (int64_t)x.strides[0], call.dev<const float>(1), (int64_t)y.strides[0], call.dev<float>(2)); });
  return (void)call.finish();
}
extern "C" int exoblas_sdot_stride_1_async(hipStream_t stream, int_fast32_t n, struct exo_win_1f32c x, struct exo_win_1f32c y, float *result) {
  exoblas::Call call("exoblas_sdot_stride_1_async", exoblas::Call::kAsync, stream);
  if (n < 0) return call.fail_size("n", n);
  call.window(0, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(float), false, true);
  call.window(1, y.data, exoblas::span_lo(n, y.strides[0]), exoblas::span_hi(n, y.strides[0]), sizeof(float), false, true);
  call.window(2, result, 0, 1, sizeof(float), true, true, true);
  if (!call.stage()) return call.status();
  if (!call.ok()) return call.status();
  call.run([&] { exoblas::ops::dot<float>(call.ctx(), (int64_t)n, call.dev<const float>(0), (int64_t)x.strides[0], call.dev<const float>(1), (int64_t)y.strides[0], call.dev<float>(2)); });
  return call.finish();
}

// src/level1/dot.py:9-13
extern "C" void exo_ddot_stride_any(void *ctxt, int_fast32_t n, struct exo_win_1f64c x, struct exo_win_1f64c y, double *result) {
  exoblas::Call call("exo_ddot_stride_any", exoblas::Call::kSync, ctxt);
  if (n < 0) return (void)call.fail_size("n", n);
  call.window(0, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(double), false, true);
  call.window(1, y.data, exoblas::span_lo(n, y.strides[0]), exoblas::span_hi(n, y.strides[0]), sizeof(double), false, true);
  call.window(2, result, 0, 1, sizeof(double), true, true, true);
  if (!call.stage()) return (void)call.status();
  if (!call.ok()) return (void)call.status();
  call.run([&] { exoblas::ops::dot<double>(call.ctx(), (int64_t)n, call.dev<const double>(0), (int64_t)x.strides[0], call.dev<const double>(1), (int64_t)y.strides[0], call.dev<double>(2)); });
  return (void)call.finish();
}
extern "C" int exoblas_ddot_stride_any_async(hipStream_t stream, int_fast32_t n, struct exo_win_1f64c x, struct exo_win_1f64c y, double *result) {
  exoblas::Call call("exoblas_ddot_stride_any_async", exoblas::Call::kAsync, stream);
  if (n < 0) return call.fail_size("n", n);
  call.window(0, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(double), false, true);
  call.window(1, y.data, exoblas::span_lo(n, y.strides[0]), exoblas::span_hi(n, y.strides[0]), sizeof(double), false, true);
  call.window(2, result, 0, 1, sizeof(double), true, true, true);
  if (!call.stage()) return call.status();
  if (!call.ok()) return call.status();
  call.run([&] { exoblas::ops::dot<double>(call.ctx(), (int64_t)n, call.dev<const double>(0), (int64_t)x.strides[0], call.dev<const double>(1), (int64_t)y.strides[0], call.dev<double>(2)); });
  return call.finish();
}

// src/level1/dot.py:9-13
extern "C" void exo_ddot_stride_1(void *ctxt, int_fast32_t n, struct exo_win_1f64c x, struct exo_win_1f64c y, double *result) {
  exoblas::Call call("exo_ddot_stride_1", exoblas::Call::kSync, ctxt);
  if (n < 0) return (void)call.fail_size("n", n);
  call.window(0, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(double), false, true);
  call.window(1, y.data, exoblas::span_lo(n, y.strides[0]), exoblas::span_hi(n, y.strides[0]), sizeof(double), false, true);
  call.window(2, result, 0, 1, sizeof(double), true, true, true);
  if (!call.stage()) return (void)call.status();
  if (!call.ok()) return (void)call.status();
  call.run([&] { exoblas::ops::dot<double>(call.ctx(), (int64_t)n, call.dev<const double>(0), (int64_t)x.strides[0], call.dev<const double>(1), (int64_t)y.strides[0], call.dev<double>(2)); });
  return (void)call.finish();
}
extern "C" int exoblas_ddot_stride_1_async(hipStream_t stream, int_fast32_t n, struct exo_win_1f64c x, struct exo_win_1f64c y, double *result) {
  exoblas::Call call("exoblas_ddot_stride_1_async", exoblas::Call::kAsync, stream);
  if (n < 0) return call.fail_size("n", n);
  call.window(0, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(double), false, true);
  call.window(1, y.data, exoblas::span_lo(n, y.strides[0]), exoblas::span_hi(n, y.strides[0]), sizeof(double), false, true);
  call.window(2, result, 0, 1, sizeof(double), true, true, true);
  if (!call.stage()) return call.status();
  if (!call.ok()) return call.status();
  call.run([&] { exoblas::ops::dot<double>(call.ctx(), (int64_t)n, call.dev<const double>(0), (int64_t)x.strides[0], call.dev<const double>(1), (int64_t)y.strides[0], call.dev<double>(2)); });
  return call.finish();
}

// src/level1/dsdot.py:11-21
extern "C" void exo_dsdot_stride_any(void *ctxt, int_fast32_t n, struct exo_win_1f32c x, struct exo_win_1f32c y, double *result) {
  exoblas::Call call("exo_dsdot_stride_any", exoblas::Call::kSync, ctxt);
  if (n < 0) return (void)call.fail_size("n", n);
  call.window(0, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(float), false, true);
  call.window(1, y.data, exoblas::span_lo(n, y.strides[0]), exoblas::span_hi(n, y.strides[0]), sizeof(float), false, true);
  call.window(2, result, 0, 1, sizeof(double), true, true, true);
  if (!call.stage()) return (void)call.status();
  if (!call.ok()) return (void)call.status();
  call.run([&] { exoblas::ops::dsdot(call.ctx(), (int64_t)n, call.dev<const float>(0), (int64_t)x.strides[0], call.dev<const float>(1), (int64_t)y.strides[0], call.dev<double>(2)); });
  return (void)call.finish();
}
extern "C" int exoblas_dsdot_stride_any_async(hipStream_t stream, int_fast32_t n, struct exo_win_1f32c x, struct exo_win_1f32c y, double *result) {
  exoblas::Call call("exoblas_dsdot_stride_any_async", exoblas::Call::kAsync, stream);
  if (n < 0) return call.fail_size("n", n);
  call.window(0, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(float), false, true);
  call.window(1, y.data, exoblas::span_lo(n, y.strides[0]), exoblas::span_hi(n, y.strides[0]), sizeof(float), false, true);
  call.window(2, result, 0, 1, sizeof(double), true, true, true);
  if (!call.stage()) return call.status();
  if (!call.ok()) return call.status();
  call.run([&] { exoblas::ops::dsdot(call.ctx(), (int64_t)n, call.dev<const float>(0), (int64_t)x.strides[0], call.dev<const float>(1), (int64_t)y.strides[0], call.dev<double>(2)); });
  return call.finish();
}

// src/level1/dsdot.py:11-21
extern "C" void exo_dsdot_stride_1(void *ctxt, int_fast32_t n, struct exo_win_1f32c x, struct exo_win_1f32c y, double *result) {
  exoblas::Call call("exo_dsdot_stride_1", exoblas::Call::kSync, ctxt);
  if (n < 0) return (void)call.fail_size("n", n);
  call.window(0, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(float), false, true);
  call.window(1, y.data, exoblas::span_lo(n, y.strides[0]), exoblas::span_hi(n, y.strides[0]), sizeof(float), false, true);
  call.window(2, result, 0, 1, sizeof(double), true, true, true);
  if (!call.stage()) return (void)call.status();
  if (!call.ok()) return (void)call.status();
  call.run([&] { exoblas::ops::dsdot(call.ctx(), (int64_t)n, call.dev<const float>(0), (int64_t)x.strides[0], call.dev<const float>(1), (int64_t)y.strides[0], call.dev<double>(2)); });
  return (void)call.finish();
}
extern "C" int exoblas_dsdot_stride_1_async(hipStream_t stream, int_fast32_t n, struct exo_win_1f32c x, struct exo_win_1f32c y, double *result) {
  exoblas::Call call("exoblas_dsdot_stride_1_async", exoblas::Call::kAsync, stream);
  if (n < 0) return call.fail_size("n", n);
  call.window(0, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(float), false, true);
  call.window(1, y.data, exoblas::span_lo(n, y.strides[0]), exoblas::span_hi(n, y.strides[0]), sizeof(float), false, true);
  call.window(2, result, 0, 1, sizeof(double), true, true, true);
  if (!call.stage()) return call.status();
  if (!call.ok()) return call.status();
  call.run([&] { exoblas::ops::dsdot(call.ctx(), (int64_t)n, call.dev<const float>(0), (int64_t)x.strides[0], call.dev<const float>(1), (int64_t)y.strides[0], call.dev<double>(2)); });
  return call.finish();
}

// src/level1/dsdot.py:24-36
extern "C" void exo_sdsdot_stride_any(void *ctxt, int_fast32_t n, const float *sb, struct exo_win_1f32c x, struct exo_win_1f32c y, float *result) {
  exoblas::Call call("exo_sdsdot_stride_any", exoblas::Call::kSync, ctxt);
  if (n < 0) return (void)call.fail_size("n", n);
  call.window(0, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(float), false, true);
  call.window(1, y.data, exoblas::span_lo(n, y.strides[0]), exoblas::span_hi(n, y.strides[0]), sizeof(float), false, true);
  call.window(2, result, 0, 1, sizeof(float), true, true, true);
  if (!call.stage()) return (void)call.status();
  const float sb_v = call.scalar(sb);
  if (!call.ok()) return (void)call.status();
  call.run([&] { exoblas::ops::sdsdot(call.ctx(), (int64_t)n, sb_v, call.dev<const float>(0), (int64_t)x.strides[0], call.dev<const float>(1), (int64_t)y.strides[0], call.dev<float>(2)); });
  return (void)call.finish();
}
extern "C" int exoblas_sdsdot_stride_any_async(hipStream_t stream, int_fast32_t n, const float *sb, struct exo_win_1f32c x, struct exo_win_1f32c y, float *result) {
  exoblas::Call call("exoblas_sdsdot_stride_any_async", exoblas::Call::kAsync, stream);
  if (n < 0) return call.fail_size("n", n);
  call.window(0, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(float), false, true);
  call.window(1, y.data, exoblas::span_lo(n, y.strides[0]), exoblas::span_hi(n, y.strides[0]), sizeof(float), false, true);
  call.window(2, result, 0, 1, sizeof(float), true, true, true);
  if (!call.stage()) return call.status();
  const float sb_v = call.scalar(sb);
  if (!call.ok()) return call.status();
  call.run([&] { exoblas::ops::sdsdot(call.ctx(), (int64_t)n, sb_v, call.dev<const float>(0), (int64_t)x.strides[0], call.dev<const float>(1), (int64_t)y.strides[0], call.dev<float>(2)); });
  return call.finish();
}

// src/level1/dsdot.py:24-36
extern "C" void exo_sdsdot_stride_1(void *ctxt, int_fast32_t n, const float *sb, struct exo_win_1f32c x, struct exo_win_1f32c y, float *result) {
  exoblas::Call call("exo_sdsdot_stride_1", exoblas::Call::kSync, ctxt);
  if (n < 0) return (void)call.fail_size("n", n);
  call.window(0, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(float), false, true);
  call.window(1, y.data, exoblas::span_lo(n, y.strides[0]), exoblas::span_hi(n, y.strides[0]), sizeof(float), false, true);
  call.window(2, result, 0, 1, sizeof(float), true, true, true);
  if (!call.stage()) return (void)call.status();
  const float sb_v = call.scalar(sb);
  if (!call.ok()) return (void)call.status();
  call.run([&] { exoblas::ops::sdsdot(call.ctx(), (int64_t)n, sb_v, call.dev<const float>(0), (int64_t)x.strides[0], call.dev<const float>(1), (int64_t)y.strides[0], call.dev<float>(2)); });
  return (void)call.finish();
}
extern "C" int exoblas_sdsdot_stride_1_async(hipStream_t stream, int_fast32_t n, const float *sb, struct exo_win_1f32c x, struct exo_win_1f32c y, float *result) {
  exoblas::Call call("exoblas_sdsdot_stride_1_async", exoblas::Call::kAsync, stream);
  if (n < 0) return call.fail_size("n", n);
  call.window(0, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(float), false, true);
  call.window(1, y.data, exoblas::span_lo(n, y.strides[0]), exoblas::span_hi(n, y.strides[0]), sizeof(float), false, true);
  call.window(2, result, 0, 1, sizeof(float), true, true, true);
  if (!call.stage()) return call.status();
  const float sb_v = call.scalar(sb);
  if (!call.ok()) return call.status();
  call.run([&] { exoblas::ops::sdsdot(call.ctx(), (int64_t)n, sb_v, call.dev<const float>(0), (int64_t)x.strides[0], call.dev<const float>(1), (int64_t)y.strides[0], call.dev<float>(2)); });
  return call.finish();
}

// src/level1/exo_copy.py:9-12
extern "C" void exo_scopy_stride_any(void *ctxt, int_fast32_t n, struct exo_win_1f32c x, struct exo_win_1f32 y) {
  exoblas::Call call("exo_scopy_stride_any", exoblas::Call::kSync, ctxt);
  if (n < 0) return (void)call.fail_size("n", n);
  call.window(0, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(float), false, true);
  call.window(1, y.data, exoblas::span_lo(n, y.strides[0]), exoblas::span_hi(n, y.strides[0]), sizeof(float), true, true);
  if (!call.stage()) return (void)call.status();
  if (!call.ok()) return (void)call.status();
  call.run([&] { exoblas::ops::copy<float>(call.ctx(), (int64_t)n, call.dev<const float>(0), (int64_t)x.strides[0], call.dev<float>(1), (int64_t)y.strides[0]); });
  return (void)call.finish();
}
extern "C" int exoblas_scopy_stride_any_async(hipStream_t stream, int_fast32_t n, struct exo_win_1f32c x, struct exo_win_1f32 y) {
  exoblas::Call call("exoblas_scopy_stride_any_async", exoblas::Call::kAsync, stream);
  if (n < 0) return call.fail_size("n", n);
  call.window(0, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(float), false, true);
  call.window(1, y.data, exoblas::span_lo(n, y.strides[0]), exoblas::span_hi(n, y.strides[0]), sizeof(float), true, true);
  if (!call.stage()) return call.status();
  if (!call.ok()) return call.status();
  call.run([&] { exoblas::ops::copy<float>(call.ctx(), (int64_t)n, call.dev<const float>(0), (int64_t)x.strides[0], call.dev<float>(1), (int64_t)y.strides[0]); });
  return call.finish();
}

// src/level1/exo_copy.py:9-12
extern "C" void exo_scopy_stride_1(void *ctxt, int_fast32_t n, struct exo_win_1f32c x, struct exo_win_1f32 y) {
  exoblas::Call call("exo_scopy_stride_1", exoblas::Call::kSync, ctxt);
  if (n < 0) return (void)call.fail_size("n", n);
  call.window(0, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(float), false, true);
  call.window(1, y.data, exoblas::span_lo(n, y.strides[0]), exoblas::span_hi(n, y.strides[0]), sizeof(float), true, true);
  if (!call.stage()) return (void)call.status();
  if (!call.ok()) return (void)call.status();
  call.run([&] { exoblas::ops::copy<float>(call.ctx(), (int64_t)n, call.dev<const float>(0), (int64_t)x.strides[0], call.dev<float>(1), (int64_t)y.strides[0]); });
  return (void)call.finish();
}
extern "C" int exoblas_scopy_stride_1_async(hipStream_t stream, int_fast32_t n, struct exo_win_1f32c x, struct exo_win_1f32 y) {
  exoblas::Call call("exoblas_scopy_stride_1_async", exoblas::Call::kAsync, stream);
  if (n < 0) return call.fail_size("n", n);
  call.window(0, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(float), false, true);
  call.window(1, y.data, exoblas::span_lo(n, y.strides[0]), exoblas::span_hi(n, y.strides[0]), sizeof(float), true, true);
  if (!call.stage()) return call.status();
  if (!call.ok()) return call.status();
  call.run([&] { exoblas::ops::copy<float>(call.ctx(), (int64_t)n, call.dev<const float>(0), (int64_t)x.strides[0], call.dev<float>(1), (int64_t)y.strides[0]); });
  return call.finish();
}

// src/level1/exo_copy.py:9-12
extern "C" void exo_dcopy_stride_any(void *ctxt, int_fast32_t n, struct exo_win_1f64c x, struct exo_win_1f64 y) {
  exoblas::Call call("exo_dcopy_stride_any", exoblas::Call::kSync, ctxt);
  if (n < 0) return (void)call.fail_size("n", n);
  call.window(0, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(double), false, true);
  call.window(1, y.data, exoblas::span_lo(n, y.strides[0]), exoblas::span_hi(n, y.strides[0]), sizeof(double), true, true);
  if (!call.stage()) return (void)call.status();
  if (!call.ok()) return (void)call.status();
  call.run([&] { exoblas::ops::copy<double>(call.ctx(), (int64_t)n, call.dev<const double>(0), (int64_t)x.strides[0], call.dev<double>(1), (int64_t)y.strides[0]); });
  return (void)call.finish();
}
extern "C" int exoblas_dcopy_stride_any_async(hipStream_t stream, int_fast32_t n, struct exo_win_1f64c x, struct exo_win_1f64 y) {
  exoblas::Call call("exoblas_dcopy_stride_any_async", exoblas::Call::kAsync, stream);
  if (n < 0) return call.fail_size("n", n);
  call.window(0, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(double), false, true);
  call.window(1, y.data, exoblas::span_lo(n, y.strides[0]), exoblas::span_hi(n, y.strides[0]), sizeof(double), true, true);
  if (!call.stage()) return call.status();
  if (!call.ok()) return call.status();
  call.run([&] { exoblas::ops::copy<double>(call.ctx(), (int64_t)n, call.dev<const double>(0), (int64_t)x.strides[0], call.dev<double>(1), (int64_t)y.strides[0]); });
  return call.finish();
}

// src/level1/exo_copy.py:9-12
extern "C" void exo_dcopy_stride_1(void *ctxt, int_fast32_t n, struct exo_win_1f64c x, struct exo_win_1f64 y) {
  exoblas::Call call("exo_dcopy_stride_1", exoblas::Call::kSync, ctxt);
  if (n < 0) return (void)call.fail_size("n", n);
  call.window(0, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(double), false, true);
  call.window(1, y.data, exoblas::span_lo(n, y.strides[0]), exoblas::span_hi(n, y.strides[0]), sizeof(double), true, true);
  if (!call.stage()) return (void)call.status();
  if (!call.ok()) return (void)call.status();
  call.run([&] { exoblas::ops::copy<double>(call.ctx(), (int64_t)n, call.dev<const double>(0), (int64_t)x.strides[0], call.dev<double>(1), (int64_t)y.strides[0]); });
  return (void)call.finish();
}
extern "C" int exoblas_dcopy_stride_1_async(hipStream_t stream, int_fast32_t n, struct exo_win_1f64c x, struct exo_win_1f64 y) {
  exoblas::Call call("exoblas_dcopy_stride_1_async", exoblas::Call::kAsync, stream);
  if (n < 0) return call.fail_size("n", n);
  call.window(0, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(double), false, true);
  call.window(1, y.data, exoblas::span_lo(n, y.strides[0]), exoblas::span_hi(n, y.strides[0]), sizeof(double), true, true);
  if (!call.stage()) return call.status();
  if (!call.ok()) return call.status();
  call.run([&] { exoblas::ops::copy<double>(call.ctx(), (int64_t)n, call.dev<const double>(0), (int64_t)x.strides[0], call.dev<double>(1), (int64_t)y.strides[0]); });
  return call.finish();
}

// src/level1/swap.py:9-15
extern "C" void exo_sswap_stride_any(void *ctxt, int_fast32_t n, struct exo_win_1f32 x, struct exo_win_1f32 y) {
  exoblas::Call call("exo_sswap_stride_any", exoblas::Call::kSync, ctxt);
  if (n < 0) return (void)call.fail_size("n", n);
  call.window(0, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(float), true, true);
  call.window(1, y.data, exoblas::span_lo(n, y.strides[0]), exoblas::span_hi(n, y.strides[0]), sizeof(float), true, true);
  if (!call.stage()) return (void)call.status();
  if (!call.ok()) return (void)call.status();
  call.run([&] { exoblas::ops::swap<float>(call.ctx(), (int64_t)n, call.dev<float>(0), (int64_t)x.strides[0], call.dev<float>(1), (int64_t)y.strides[0]); });
  return (void)call.finish();
}
extern "C" int exoblas_sswap_stride_any_async(hipStream_t stream, int_fast32_t n, struct exo_win_1f32 x, struct exo_win_1f32 y) {
  exoblas::Call call("exoblas_sswap_stride_any_async", exoblas::Call::kAsync, stream);
  if (n < 0) return call.fail_size("n", n);
  call.window(0, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(float), true, true);
  call.window(1, y.data, exoblas::span_lo(n, y.strides[0]), exoblas::span_hi(n, y.strides[0]), sizeof(float), true, true);
  if (!call.stage()) return call.status();
  if (!call.ok()) return call.status();
  call.run([&] { exoblas::ops::swap<float>(call.ctx(), (int64_t)n, call.dev<float>(0), (int64_t)x.strides[0], call.dev<float>(1), (int64_t)y.strides[0]); });
  return call.finish();
}

// src/level1/swap.py:9-15
extern "C" void exo_sswap_stride_1(void *ctxt, int_fast32_t n, struct exo_win_1f32 x, struct exo_win_1f32 y) {
  exoblas::Call call("exo_sswap_stride_1", exoblas::Call::kSync, ctxt);
  if (n < 0) return (void)call.fail_size("n", n);
  call.window(0, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(float), true, true);
  call.window(1, y.data, exoblas::span_lo(n, y.strides[0]), exoblas::span_hi(n, y.strides[0]), sizeof(float), true, true);
  if (!call.stage()) return (void)call.status();
  if (!call.ok()) return (void)call.status();
  call.run([&] { exoblas::ops::swap<float>(call.ctx(), (int64_t)n, call.dev<float>(0), (int64_t)x.strides[0], call.dev<float>(1), (int64_t)y.strides[0]); });
  return (void)call.finish();
}
extern "C" int exoblas_sswap_stride_1_async(hipStream_t stream, int_fast32_t n, struct exo_win_1f32 x, struct exo_win_1f32 y) {
  exoblas::Call call("exoblas_sswap_stride_1_async", exoblas::Call::kAsync, stream);
  if (n < 0) return call.fail_size("n", n);
  call.window(0, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(float), true, true);
  call.window(1, y.data, exoblas::span_lo(n, y.strides[0]), exoblas::span_hi(n, y.strides[0]), sizeof(float), true, true);
  if (!call.stage()) return call.status();
  if (!call.ok()) return call.status();
  call.run([&] { exoblas::ops::swap<float>(call.ctx(), (int64_t)n, call.dev<float>(0), (int64_t)x.strides[0], call.dev<float>(1), (int64_t)y.strides[0]); });
  return call.finish();
}

// src/level1/swap.py:9-15
extern "C" void exo_dswap_stride_any(void *ctxt, int_fast32_t n, struct exo_win_1f64 x, struct exo_win_1f64 y) {
  exoblas::Call call("exo_dswap_stride_any", exoblas::Call::kSync, ctxt);
  if (n < 0) return (void)call.fail_size("n", n);
  call.window(0, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(double), true, true);
  call.window(1, y.data, exoblas::span_lo(n, y.strides[0]), exoblas::span_hi(n, y.strides[0]), sizeof(double), true, true);
  if (!call.stage()) return (void)call.status();
  if (!call.ok()) return (void)call.status();
  call.run([&] { exoblas::ops::swap<double>(call.ctx(), (int64_t)n, call.dev<double>(0), (int64_t)x.strides[0], call.dev<double>(1), (int64_t)y.strides[0]); });
  return (void)call.finish();
}
extern "C" int exoblas_dswap_stride_any_async(hipStream_t stream, int_fast32_t n, struct exo_win_1f64 x, struct exo_win_1f64 y) {
  exoblas::Call call("exoblas_dswap_stride_any_async", exoblas::Call::kAsync, stream);
  if (n < 0) return call.fail_size("n", n);
  call.window(0, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(double), true, true);
  call.window(1, y.data, exoblas::span_lo(n, y.strides[0]), exoblas::span_hi(n, y.strides[0]), sizeof(double), true, true);
  if (!call.stage()) return call.status();
  if (!call.ok()) return call.status();
  call.run([&] { exoblas::ops::swap<double>(call.ctx(), (int64_t)n, call.dev<double>(0), (int64_t)x.strides[0], call.dev<double>(1), (int64_t)y.strides[0]); });
  return call.finish();
}

// src/level1/swap.py:9-15
extern "C" void exo_dswap_stride_1(void *ctxt, int_fast32_t n, struct exo_win_1f64 x, struct exo_win_1f64 y) {
  exoblas::Call call("exo_dswap_stride_1", exoblas::Call::kSync, ctxt);
  if (n < 0) return (void)call.fail_size("n", n);
  call.window(0, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(double), true, true);
  call.window(1, y.data, exoblas::span_lo(n, y.strides[0]), exoblas::span_hi(n, y.strides[0]), sizeof(double), true, true);
  if (!call.stage()) return (void)call.status();
  if (!call.ok()) return (void)call.status();
  call.run([&] { exoblas::ops::swap<double>(call.ctx(), (int64_t)n, call.dev<double>(0), (int64_t)x.strides[0], call.dev<double>(1), (int64_t)y.strides[0]); });
  return (void)call.finish();
}
extern "C" int exoblas_dswap_stride_1_async(hipStream_t stream, int_fast32_t n, struct exo_win_1f64 x, struct exo_win_1f64 y) {
  exoblas::Call call("exoblas_dswap_stride_1_async", exoblas::Call::kAsync, stream);
  if (n < 0) return call.fail_size("n", n);
  call.window(0, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(double), true, true);
  call.window(1, y.data, exoblas::span_lo(n, y.strides[0]), exoblas::span_hi(n, y.strides[0]), sizeof(double), true, true);
  if (!call.stage()) return call.status();
  if (!call.ok()) return call.status();
  call.run([&] { exoblas::ops::swap<double>(call.ctx(), (int64_t)n, call.dev<double>(0), (int64_t)x.strides[0], call.dev<double>(1), (int64_t)y.strides[0]); });
  return call.finish();
}

// src/level1/rot.py:9-15
extern "C" void exo_srot_stride_any(void *ctxt, int_fast32_t n, struct exo_win_1f32 x, struct exo_win_1f32 y, const float *c, const float *s) {
  exoblas::Call call("exo_srot_stride_any", exoblas::Call::kSync, ctxt);
  if (n < 0) return (void)call.fail_size("n", n);
  call.window(0, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(float), true, true);
  call.window(1, y.data, exoblas::span_lo(n, y.strides[0]), exoblas::span_hi(n, y.strides[0]), sizeof(float), true, true);
  if (!call.stage()) return (void)call.status();
  const float c_v = call.scalar(c);
  const float s_v = call.scalar(s);
  if (!call.ok()) return (void)call.status();
  call.run([&] { exoblas::ops::rot<float>(call.ctx(), (int64_t)n, call.dev<float>(0), (int64_t)x.strides[0], call.dev<float>(1), (int64_t)y.strides[0], c_v, s_v); });
  return (void)call.finish();
}
extern "C" int exoblas_srot_stride_any_async(hipStream_t stream, int_fast32_t n, struct exo_win_1f32 x, struct exo_win_1f32 y, const float *c, const float *s) {
  exoblas::Call call("exoblas_srot_stride_any_async", exoblas::Call::kAsync, stream);
  if (n < 0) return call.fail_size("n", n);
  call.window(0, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(float), true, true);
  call.window(1, y.data, exoblas::span_lo(n, y.strides[0]), exoblas::span_hi(n, y.strides[0]), sizeof(float), true, true);
  if (!call.stage()) return call.status();
  const float c_v = call.scalar(c);
  const float s_v = call.scalar(s);
  if (!call.ok()) return call.status();
  call.run([&] { exoblas::ops::rot<float>(call.ctx(), (int64_t)n, call.dev<float>(0), (int64_t)x.strides[0], call.dev<float>(1), (int64_t)y.strides[0], c_v, s_v); });
  return call.finish();
}

// src/level1/rot.py:9-15
extern "C" void exo_srot_stride_1(void *ctxt, int_fast32_t n, struct exo_win_1f32 x, struct exo_win_1f32 y, const float *c, const float *s) {
  exoblas::Call call("exo_srot_stride_1", exoblas::Call::kSync, ctxt);
  if (n < 0) return (void)call.fail_size("n", n);
  call.window(0, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(float), true, true);
  call.window(1, y.data, exoblas::span_lo(n, y.strides[0]), exoblas::span_hi(n, y.strides[0]), sizeof(float), true, true);
  if (!call.stage()) return (void)call.status();
  const float c_v = call.scalar(c);
  const float s_v = call.scalar(s);
  if (!call.ok()) return (void)call.status();
  call.run([&] { exoblas::ops::rot<float>(call.ctx(), (int64_t)n, call.dev<float>(0), (int64_t)x.strides[0], call.dev<float>(1), (int64_t)y.strides[0], c_v, s_v); });
  return (void)call.finish();
}
extern "C" int exoblas_srot_stride_1_async(hipStream_t stream, int_fast32_t n, struct exo_win_1f32 x, struct exo_win_1f32 y, const float *c, const float *s) {
  exoblas::Call call("exoblas_srot_stride_1_async", exoblas::Call::kAsync, stream);
  if (n < 0) return call.fail_size("n", n);
  call.window(0, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(float), true, true);
  call.window(1, y.data, exoblas::span_lo(n, y.strides[0]), exoblas::span_hi(n, y.strides[0]), sizeof(float), true, true);
  if (!call.stage()) return call.status();
  const float c_v = call.scalar(c);
  const float s_v = call.scalar(s);
  if (!call.ok()) return call.status();
  call.run([&] { exoblas::ops::rot<float>(call.ctx(), (int64_t)n, call.dev<float>(0), (int64_t)x.strides[0], call.dev<float>(1), (int64_t)y.strides[0], c_v, s_v); });
  return call.finish();
}

// src/level1/rot.py:9-15
extern "C" void exo_drot_stride_any(void *ctxt, int_fast32_t n, struct exo_win_1f64 x, struct exo_win_1f64 y, const double *c, const double *s) {
  exoblas::Call call("exo_drot_stride_any", exoblas::Call::kSync, ctxt);
  if (n < 0) return (void)call.fail_size("n", n);
  call.window(0, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(double), true, true);
  call.window(1, y.data, exoblas::span_lo(n, y.strides[0]), exoblas::span_hi(n, y.strides[0]), sizeof(double), true, true);
  if (!call.stage()) return (void)call.status();
  const double c_v = call.scalar(c);
  const double s_v = call.scalar(s);
  if (!call.ok()) return (void)call.status();
  call.run([&] { exoblas::ops::rot<double>(call.ctx(), (int64_t)n, call.dev<double>(0), (int64_t)x.strides[0], call.dev<double>(1), (int64_t)y.strides[0], c_v, s_v); });
  return (void)call.finish();
}
extern "C" int exoblas_drot_stride_any_async(hipStream_t stream, int_fast32_t n, struct exo_win_1f64 x, struct exo_win_1f64 y, const double *c, const double *s) {
  exoblas::Call call("exoblas_drot_stride_any_async", exoblas::Call::kAsync, stream);
  if (n < 0) return call.fail_size("n", n);
  call.window(0, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(double), true, true);
  call.window(1, y.data, exoblas::span_lo(n, y.strides[0]), exoblas::span_hi(n, y.strides[0]), sizeof(double), true, true);
  if (!call.stage()) return call.status();
  const double c_v = call.scalar(c);
  const double s_v = call.scalar(s);
  if (!call.ok()) return call.status();
  call.run([&] { exoblas::ops::rot<double>(call.ctx(), (int64_t)n, call.dev<double>(0), (int64_t)x.strides[0], call.dev<double>(1), (int64_t)y.strides[0], c_v, s_v); });
  return call.finish();
}

// src/level1/rot.py:9-15
extern "C" void exo_drot_stride_1(void *ctxt, int_fast32_t n, struct exo_win_1f64 x, struct exo_win_1f64 y, const double *c, const double *s) {
  exoblas::Call call("exo_drot_stride_1", exoblas::Call::kSync, ctxt);
  if (n < 0) return (void)call.fail_size("n", n);
  call.window(0, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(double), true, true);
  call.window(1, y.data, exoblas::span_lo(n, y.strides[0]), exoblas::span_hi(n, y.strides[0]), sizeof(double), true, true);
  if (!call.stage()) return (void)call.status();
  const double c_v = call.scalar(c);
  const double s_v = call.scalar(s);
  if (!call.ok()) return (void)call.status();
  call.run([&] { exoblas::ops::rot<double>(call.ctx(), (int64_t)n, call.dev<double>(0), (int64_t)x.strides[0], call.dev<double>(1), (int64_t)y.strides[0], c_v, s_v); });
  return (void)call.finish();
}
extern "C" int exoblas_drot_stride_1_async(hipStream_t stream, int_fast32_t n, struct exo_win_1f64 x, struct exo_win_1f64 y, const double *c, const double *s) {
  exoblas::Call call("exoblas_drot_stride_1_async", exoblas::Call::kAsync, stream);
  if (n < 0) return call.fail_size("n", n);
  call.window(0, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(double), true, true);
  call.window(1, y.data, exoblas::span_lo(n, y.strides[0]), exoblas::span_hi(n, y.strides[0]), sizeof(double), true, true);
  if (!call.stage()) return call.status();
  const double c_v = call.scalar(c);
  const double s_v = call.scalar(s);
  if (!call.ok()) return call.status();
  call.run([&] { exoblas::ops::rot<double>(call.ctx(), (int64_t)n, call.dev<double>(0), (int64_t)x.strides[0], call.dev<double>(1), (int64_t)y.strides[0], c_v, s_v); });
  return call.finish();
}

// src/level1/rotm.py:9-15
extern "C" void exo_srotm_flag_neg_one_stride_any(void *ctxt, int_fast32_t n, struct exo_win_1f32 x, struct exo_win_1f32 y, const float *H) {
  exoblas::Call call("exo_srotm_flag_neg_one_stride_any", exoblas::Call::kSync, ctxt);
  if (n < 0) return (void)call.fail_size("n", n);
  call.window(0, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(float), true, true);
  call.window(1, y.data, exoblas::span_lo(n, y.strides[0]), exoblas::span_hi(n, y.strides[0]), sizeof(float), true, true);
  if (!call.stage()) return (void)call.status();
  float H_v[4];
  call.array(H, H_v, 4);
  if (!call.ok()) return (void)call.status();
  call.run([&] { exoblas::ops::rotm_flag_neg_one<float>(call.ctx(), (int64_t)n, call.dev<float>(0), (int64_t)x.strides[0], call.dev<float>(1), (int64_t)y.strides[0], H_v); });
  return (void)call.finish();
}
extern "C" int exoblas_srotm_flag_neg_one_stride_any_async(hipStream_t stream, int_fast32_t n, struct exo_win_1f32 x, struct exo_win_1f32 y, const float *H) {
  exoblas::Call call("exoblas_srotm_flag_neg_one_stride_any_async", exoblas::Call::kAsync, stream);
  if (n < 0) return call.fail_size("n", n);
  call.window(0, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(float), true, true);
  call.window(1, y.data, exoblas::span_lo(n, y.strides[0]), exoblas::span_hi(n, y.strides[0]), sizeof(float), true, true);
  if (!call.stage()) return call.status();
  float H_v[4];
  call.array(H, H_v, 4);
  if (!call.ok()) return call.status();
  call.run([&] { exoblas::ops::rotm_flag_neg_one<float>(call.ctx(), (int64_t)n, call.dev<float>(0), (int64_t)x.strides[0], call.dev<float>(1), (int64_t)y.strides[0], H_v); });
  return call.finish();
}

// src/level1/rotm.py:9-15
extern "C" void exo_srotm_flag_neg_one_stride_1(void *ctxt, int_fast32_t n, struct exo_win_1f32 x, struct exo_win_1f32 y, const float *H) {
  exoblas::Call call("exo_srotm_flag_neg_one_stride_1", exoblas::Call::kSync, ctxt);
  if (n < 0) return (void)call.fail_size("n", n);
  call.window(0, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(float), true, true);
  call.window(1, y.data, exoblas::span_lo(n, y.strides[0]), exoblas::span_hi(n, y.strides[0]), sizeof(float), true, true);
  if (!call.stage()) return (void)call.status();
  float H_v[4];
  call.array(H, H_v, 4);
  if (!call.ok()) return (void)call.status();
  call.run([&] { exoblas::ops::rotm_flag_neg_one<float>(call.ctx(), (int64_t)n, call.dev<float>(0), (int64_t)x.strides[0], call.dev<float>(1), (int64_t)y.strides[0], H_v); });
  return (void)call.finish();
}
extern "C" int exoblas_srotm_flag_neg_one_stride_1_async(hipStream_t stream, int_fast32_t n, struct exo_win_1f32 x, struct exo_win_1f32 y, const float *H) {
  exoblas::Call call("exoblas_srotm_flag_neg_one_stride_1_async", exoblas::Call::kAsync, stream);
  if (n < 0) return call.fail_size("n", n);
  call.window(0, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(float), true, true);
  call.window(1, y.data, exoblas::span_lo(n, y.strides[0]), exoblas::span_hi(n, y.strides[0]), sizeof(float), true, true);
  if (!call.stage()) return call.status();
  float H_v[4];
  call.array(H, H_v, 4);
  if (!call.ok()) return call.status();
  call.run([&] { exoblas::ops::rotm_flag_neg_one<float>(call.ctx(), (int64_t)n, call.dev<float>(0), (int64_t)x.strides[0], call.dev<float>(1), (int64_t)y.strides[0], H_v); });
  return call.finish();
}

// src/level1/rotm.py:9-15
extern "C" void exo_drotm_flag_neg_one_stride_any(void *ctxt, int_fast32_t n, struct exo_win_1f64 x, struct exo_win_1f64 y, const double *H) {
  exoblas::Call call("exo_drotm_flag_neg_one_stride_any", exoblas::Call::kSync, ctxt);
  if (n < 0) return (void)call.fail_size("n", n);
  call.window(0, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(double), true, true);
  call.window(1, y.data, exoblas::span_lo(n, y.strides[0]), exoblas::span_hi(n, y.strides[0]), sizeof(double), true, true);
  if (!call.stage()) return (void)call.status();
  double H_v[4];
  call.array(H, H_v, 4);
  if (!call.ok()) return (void)call.status();
  call.run([&] { exoblas::ops::rotm_flag_neg_one<double>(call.ctx(), (int64_t)n, call.dev<double>(0), (int64_t)x.strides[0], call.dev<double>(1), (int64_t)y.strides[0], H_v); });
  return (void)call.finish();
}
extern "C" int exoblas_drotm_flag_neg_one_stride_any_async(hipStream_t stream, int_fast32_t n, struct exo_win_1f64 x, struct exo_win_1f64 y, const double *H) {
  exoblas::Call call("exoblas_drotm_flag_neg_one_stride_any_async", exoblas::Call::kAsync, stream);
  if (n < 0) return call.fail_size("n", n);
  call.window(0, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(double), true, true);
  call.window(1, y.data, exoblas::span_lo(n, y.strides[0]), exoblas::span_hi(n, y.strides[0]), sizeof(double), true, true);
  if (!call.stage()) return call.status();
  double H_v[4];
  call.array(H, H_v, 4);
  if (!call.ok()) return call.status();
  call.run([&] { exoblas::ops::rotm_flag_neg_one<double>(call.ctx(), (int64_t)n, call.dev<double>(0), (int64_t)x.strides[0], call.dev<double>(1), (int64_t)y.strides[0], H_v); });
  return call.finish();
}

// src/level1/rotm.py:9-15
extern "C" void exo_drotm_flag_neg_one_stride_1(void *ctxt, int_fast32_t n, struct exo_win_1f64 x, struct exo_win_1f64 y, const double *H) {
  exoblas::Call call("exo_drotm_flag_neg_one_stride_1", exoblas::Call::kSync, ctxt);
  if (n < 0) return (void)call.fail_size("n", n);
  call.window(0, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(double), true, true);
  call.window(1, y.data, exoblas::span_lo(n, y.strides[0]), exoblas::span_hi(n, y.strides[0]), sizeof(double), true, true);
  if (!call.stage()) return (void)call.status();
  double H_v[4];
  call.array(H, H_v, 4);
  if (!call.ok()) return (void)call.status();
  call.run([&] { exoblas::ops::rotm_flag_neg_one<double>(call.ctx(), (int64_t)n, call.dev<double>(0), (int64_t)x.strides[0], call.dev<double>(1), (int64_t)y.strides[0], H_v); });
  return (void)call.finish();
}
extern "C" int exoblas_drotm_flag_neg_one_stride_1_async(hipStream_t stream, int_fast32_t n, struct exo_win_1f64 x, struct exo_win_1f64 y, const double *H) {
  exoblas::Call call("exoblas_drotm_flag_neg_one_stride_1_async", exoblas::Call::kAsync, stream);
  if (n < 0) return call.fail_size("n", n);
  call.window(0, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(double), true, true);
  call.window(1, y.data, exoblas::span_lo(n, y.strides[0]), exoblas::span_hi(n, y.strides[0]), sizeof(double), true, true);
  if (!call.stage()) return call.status();
  double H_v[4];
  call.array(H, H_v, 4);
  if (!call.ok()) return call.status();
  call.run([&] { exoblas::ops::rotm_flag_neg_one<double>(call.ctx(), (int64_t)n, call.dev<double>(0), (int64_t)x.strides[0], call.dev<double>(1), (int64_t)y.strides[0], H_v); });
  return call.finish();
}

// src/level1/rotm.py:18-24
extern "C" void exo_srotm_flag_zero_stride_any(void *ctxt, int_fast32_t n, struct exo_win_1f32 x, struct exo_win_1f32 y, const float *H) {
  exoblas::Call call("exo_srotm_flag_zero_stride_any", exoblas::Call::kSync, ctxt);
  if (n < 0) return (void)call.fail_size("n", n);
  call.window(0, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(float), true, true);
  call.window(1, y.data, exoblas::span_lo(n, y.strides[0]), exoblas::span_hi(n, y.strides[0]), sizeof(float), true, true);
  if (!call.stage()) return (void)call.status();
  float H_v[4];
  call.array(H, H_v, 4);
  if (!call.ok()) return (void)call.status();
  call.run([&] { exoblas::ops::rotm_flag_zero<float>(call.ctx(), (int64_t)n, call.dev<float>(0), (int64_t)x.strides[0], call.dev<float>(1), (int64_t)y.strides[0], H_v); });
  return (void)call.finish();
}
extern "C" int exoblas_srotm_flag_zero_stride_any_async(hipStream_t stream, int_fast32_t n, struct exo_win_1f32 x, struct exo_win_1f32 y, const float *H) {
  exoblas::Call call("exoblas_srotm_flag_zero_stride_any_async", exoblas::Call::kAsync, stream);
  if (n < 0) return call.fail_size("n", n);
  call.window(0, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(float), true, true);
  call.window(1, y.data, exoblas::span_lo(n, y.strides[0]), exoblas::span_hi(n, y.strides[0]), sizeof(float), true, true);
  if (!call.stage()) return call.status();
  float H_v[4];
  call.array(H, H_v, 4);
  if (!call.ok()) return call.status();
  call.run([&] { exoblas::ops::rotm_flag_zero<float>(call.ctx(), (int64_t)n, call.dev<float>(0), (int64_t)x.strides[0], call.dev<float>(1), (int64_t)y.strides[0], H_v); });
  return call.finish();
}

// src/level1/rotm.py:18-24
extern "C" void exo_srotm_flag_zero_stride_1(void *ctxt, int_fast32_t n, struct exo_win_1f32 x, struct exo_win_1f32 y, const float *H) {
  exoblas::Call call("exo_srotm_flag_zero_stride_1", exoblas::Call::kSync, ctxt);
  if (n < 0) return (void)call.fail_size("n", n);
  call.window(0, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(float), true, true);
  call.window(1, y.data, exoblas::span_lo(n, y.strides[0]), exoblas::span_hi(n, y.strides[0]), sizeof(float), true, true);
  if (!call.stage()) return (void)call.status();
  float H_v[4];
  call.array(H, H_v, 4);
  if (!call.ok()) return (void)call.status();
  call.run([&] { exoblas::ops::rotm_flag_zero<float>(call.ctx(), (int64_t)n, call.dev<float>(0), (int64_t)x.strides[0], call.dev<float>(1), (int64_t)y.strides[0], H_v); });
  return (void)call.finish();
}
extern "C" int exoblas_srotm_flag_zero_stride_1_async(hipStream_t stream, int_fast32_t n, struct exo_win_1f32 x, struct exo_win_1f32 y, const float *H) {
  exoblas::Call call("exoblas_srotm_flag_zero_stride_1_async", exoblas::Call::kAsync, stream);
  if (n < 0) return call.fail_size("n", n);
  call.window(0, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(float), true, true);
  call.window(1, y.data, exoblas::span_lo(n, y.strides[0]), exoblas::span_hi(n, y.strides[0]), sizeof(float), true, true);
  if (!call.stage()) return call.status();
  float H_v[4];
  call.array(H, H_v, 4);
  if (!call.ok()) return call.status();
  call.run([&] { exoblas::ops::rotm_flag_zero<float>(call.ctx(), (int64_t)n, call.dev<float>(0), (int64_t)x.strides[0], call.dev<float>(1), (int64_t)y.strides[0], H_v); });
  return call.finish();
}

// src/level1/rotm.py:18-24
extern "C" void exo_drotm_flag_zero_stride_any(void *ctxt, int_fast32_t n, struct exo_win_1f64 x, struct exo_win_1f64 y, const double *H) {
  exoblas::Call call("exo_drotm_flag_zero_stride_any", exoblas::Call::kSync, ctxt);
  if (n < 0) return (void)call.fail_size("n", n);
  call.window(0, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(double), true, true);
  call.window(1, y.data, exoblas::span_lo(n, y.strides[0]), exoblas::span_hi(n, y.strides[0]), sizeof(double), true, true);
  if (!call.stage()) return (void)call.status();
  double H_v[4];
  call.array(H, H_v, 4);
  if (!call.ok()) return (void)call.status();
  call.run([&] { exoblas::ops::rotm_flag_zero<double>(call.ctx(), (int64_t)n, call.dev<double>(0), (int64_t)x.strides[0], call.dev<double>(1), (int64_t)y.strides[0], H_v); });
  return (void)call.finish();
}
extern "C" int exoblas_drotm_flag_zero_stride_any_async(hipStream_t stream, int_fast32_t n, struct exo_win_1f64 x, struct exo_win_1f64 y, const double *H) {
  exoblas::Call call("exoblas_drotm_flag_zero_stride_any_async", exoblas::Call::kAsync, stream);
  if (n < 0) return call.fail_size("n", n);
  call.window(0, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(double), true, true);
  call.window(1, y.data, exoblas::span_lo(n, y.strides[0]), exoblas::span_hi(n, y.strides[0]), sizeof(double), true, true);
  if (!call.stage()) return call.status();
  double H_v[4];
  call.array(H, H_v, 4);
  if (!call.ok()) return call.status();
  call.run([&] { exoblas::ops::rotm_flag_zero<double>(call.ctx(), (int64_t)n, call.dev<double>(0), (int64_t)x.strides[0], call.dev<double>(1), (int64_t)y.strides[0], H_v); });
  return call.finish();
}

// src/level1/rotm.py:18-24
extern "C" void exo_drotm_flag_zero_stride_1(void *ctxt, int_fast32_t n, struct exo_win_1f64 x, struct exo_win_1f64 y, const double *H) {
  exoblas::Call call("exo_drotm_flag_zero_stride_1", exoblas::Call::kSync, ctxt);
  if (n < 0) return (void)call.fail_size("n", n);
  call.window(0, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(double), true, true);
  call.window(1, y.data, exoblas::span_lo(n, y.strides[0]), exoblas::span_hi(n, y.strides[0]), sizeof(double), true, true);
  if (!call.stage()) return (void)call.status();
  double H_v[4];
  call.array(H, H_v, 4);
  if (!call.ok()) return (void)call.status();
  call.run([&] { exoblas::ops::rotm_flag_zero<double>(call.ctx(), (int64_t)n, call.dev<double>(0), (int64_t)x.strides[0], call.dev<double>(1), (int64_t)y.strides[0], H_v); });
  return (void)call.finish();
}
extern "C" int exoblas_drotm_flag_zero_stride_1_async(hipStream_t stream, int_fast32_t n, struct exo_win_1f64 x, struct exo_win_1f64 y, const double *H) {
  exoblas::Call call("exoblas_drotm_flag_zero_stride_1_async", exoblas::Call::kAsync, stream);
  if (n < 0) return call.fail_size("n", n);
  call.window(0, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(double), true, true);
  call.window(1, y.data, exoblas::span_lo(n, y.strides[0]), exoblas::span_hi(n, y.strides[0]), sizeof(double), true, true);
  if (!call.stage()) return call.status();
  double H_v[4];
  call.array(H, H_v, 4);
  if (!call.ok()) return call.status();
  call.run([&] { exoblas::ops::rotm_flag_zero<double>(call.ctx(), (int64_t)n, call.dev<double>(0), (int64_t)x.strides[0], call.dev<double>(1), (int64_t)y.strides[0], H_v); });
  return call.finish();
}

// src/level1/rotm.py:27-33
extern "C" void exo_srotm_flag_one_stride_any(void *ctxt, int_fast32_t n, struct exo_win_1f32 x, struct exo_win_1f32 y, const float *H) {
  exoblas::Call call("exo_srotm_flag_one_stride_any", exoblas::Call::kSync, ctxt);
  if (n < 0) return (void)call.fail_size("n", n);
  call.window(0, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(float), true, true);
  call.window(1, y.data, exoblas::span_lo(n, y.strides[0]), exoblas::span_hi(n, y.strides[0]), sizeof(float), true, true);
  if (!call.stage()) return (void)call.status();
  float H_v[4];
  call.array(H, H_v, 4);
  if (!call.ok()) return (void)call.status();
  call.run([&] { exoblas::ops::rotm_flag_one<float>(call.ctx(), (int64_t)n, call.dev<float>(0), (int64_t)x.strides[0], call.dev<float>(1), (int64_t)y.strides[0], H_v); });
  return (void)call.finish();
}
extern "C" int exoblas_srotm_flag_one_stride_any_async(hipStream_t stream, int_fast32_t n, struct exo_win_1f32 x, struct exo_win_1f32 y, const float *H) {
  exoblas::Call call("exoblas_srotm_flag_one_stride_any_async", exoblas::Call::kAsync, stream);
  if (n < 0) return call.fail_size("n", n);
  call.window(0, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(float), true, true);
  call.window(1, y.data, exoblas::span_lo(n, y.strides[0]), exoblas::span_hi(n, y.strides[0]), sizeof(float), true, true);
  if (!call.stage()) return call.status();
  float H_v[4];
  call.array(H, H_v, 4);
  if (!call.ok()) return call.status();
  call.run([&] { exoblas::ops::rotm_flag_one<float>(call.ctx(), (int64_t)n, call.dev<float>(0), (int64_t)x.strides[0], call.dev<float>(1), (int64_t)y.strides[0], H_v); });
  return call.finish();
}

// src/level1/rotm.py:27-33
extern "C" void exo_srotm_flag_one_stride_1(void *ctxt, int_fast32_t n, struct exo_win_1f32 x, struct exo_win_1f32 y, const float *H) {
  exoblas::Call call("exo_srotm_flag_one_stride_1", exoblas::Call::kSync, ctxt);
  if (n < 0) return (void)call.fail_size("n", n);
  call.window(0, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(float), true, true);
  call.window(1, y.data, exoblas::span_lo(n, y.strides[0]), exoblas::span_hi(n, y.strides[0]), sizeof(float), true, true);
  if (!call.stage()) return (void)call.status();
  float H_v[4];
  call.array(H, H_v, 4);
  if (!call.ok()) return (void)call.status();
  call.run([&] { exoblas::ops::rotm_flag_one<float>(call.ctx(), (int64_t)n, call.dev<float>(0), (int64_t)x.strides[0], call.dev<float>(1), (int64_t)y.strides[0], H_v); });
  return (void)call.finish();
}
extern "C" int exoblas_srotm_flag_one_stride_1_async(hipStream_t stream, int_fast32_t n, struct exo_win_1f32 x, struct exo_win_1f32 y, const float *H) {
  exoblas::Call call("exoblas_srotm_flag_one_stride_1_async", exoblas::Call::kAsync, stream);
  if (n < 0) return call.fail_size("n", n);
  call.window(0, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(float), true, true);
  call.window(1, y.data, exoblas::span_lo(n, y.strides[0]), exoblas::span_hi(n, y.strides[0]), sizeof(float), true, true);
  if (!call.stage()) return call.status();
  float H_v[4];
  call.array(H, H_v, 4);
  if (!call.ok()) return call.status();
  call.run([&] { exoblas::ops::rotm_flag_one<float>(call.ctx(), (int64_t)n, call.dev<float>(0), (int64_t)x.strides[0], call.dev<float>(1), (int64_t)y.strides[0], H_v); });
  return call.finish();
}

// src/level1/rotm.py:27-33
extern "C" void exo_drotm_flag_one_stride_any(void *ctxt, int_fast32_t n, struct exo_win_1f64 x, struct exo_win_1f64 y, const double *H) {
  exoblas::Call call("exo_drotm_flag_one_stride_any", exoblas::Call::kSync, ctxt);
  if (n < 0) return (void)call.fail_size("n", n);
  call.window(0, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(double), true, true);
  call.window(1, y.data, exoblas::span_lo(n, y.strides[0]), exoblas::span_hi(n, y.strides[0]), sizeof(double), true, true);
  if (!call.stage()) return (void)call.status();
  double H_v[4];
  call.array(H, H_v, 4);
  if (!call.ok()) return (void)call.status();
  call.run([&] { exoblas::ops::rotm_flag_one<double>(call.ctx(), (int64_t)n, call.dev<double>(0), (int64_t)x.strides[0], call.dev<double>(1), (int64_t)y.strides[0], H_v); });
  return (void)call.finish();
}
extern "C" int exoblas_drotm_flag_one_stride_any_async(hipStream_t stream, int_fast32_t n, struct exo_win_1f64 x, struct exo_win_1f64 y, const double *H) {
  exoblas::Call call("exoblas_drotm_flag_one_stride_any_async", exoblas::Call::kAsync, stream);
  if (n < 0) return call.fail_size("n", n);
  call.window(0, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(double), true, true);
  call.window(1, y.data, exoblas::span_lo(n, y.strides[0]), exoblas::span_hi(n, y.strides[0]), sizeof(double), true, true);
  if (!call.stage()) return call.status();
  double H_v[4];
  call.array(H, H_v, 4);
  if (!call.ok()) return call.status();
  call.run([&] { exoblas::ops::rotm_flag_one<double>(call.ctx(), (int64_t)n, call.dev<double>(0), (int64_t)x.strides[0], call.dev<double>(1), (int64_t)y.strides[0], H_v); });
  return call.finish();
}

// src/level1/rotm.py:27-33
extern "C" void exo_drotm_flag_one_stride_1(void *ctxt, int_fast32_t n, struct exo_win_1f64 x, struct exo_win_1f64 y, const double *H) {
  exoblas::Call call("exo_drotm_flag_one_stride_1", exoblas::Call::kSync, ctxt);
  if (n < 0) return (void)call.fail_size("n", n);
  call.window(0, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(double), true, true);
  call.window(1, y.data, exoblas::span_lo(n, y.strides[0]), exoblas::span_hi(n, y.strides[0]), sizeof(double), true, true);
  if (!call.stage()) return (void)call.status();
  double H_v[4];
  call.array(H, H_v, 4);
  if (!call.ok()) return (void)call.status();
  call.run([&] { exoblas::ops::rotm_flag_one<double>(call.ctx(), (int64_t)n, call.dev<double>(0), (int64_t)x.strides[0], call.dev<double>(1), (int64_t)y.strides[0], H_v); });
  return (void)call.finish();
}
extern "C" int exoblas_drotm_flag_one_stride_1_async(hipStream_t stream, int_fast32_t n, struct exo_win_1f64 x, struct exo_win_1f64 y, const double *H) {
  exoblas::Call call("exoblas_drotm_flag_one_stride_1_async", exoblas::Call::kAsync, stream);
  if (n < 0) return call.fail_size("n", n);
  call.window(0, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(double), true, true);
  call.window(1, y.data, exoblas::span_lo(n, y.strides[0]), exoblas::span_hi(n, y.strides[0]), sizeof(double), true, true);
  if (!call.stage()) return call.status();
  double H_v[4];
  call.array(H, H_v, 4);
  if (!call.ok()) return call.status();
  call.run([&] { exoblas::ops::rotm_flag_one<double>(call.ctx(), (int64_t)n, call.dev<double>(0), (int64_t)x.strides[0], call.dev<double>(1), (int64_t)y.strides[0], H_v); });
  return call.finish();
}

// src/level1/scal.py:9-12
extern "C" void exo_sscal_stride_any(void *ctxt, int_fast32_t n, const float *alpha, struct exo_win_1f32 x) {
  exoblas::Call call("exo_sscal_stride_any", exoblas::Call::kSync, ctxt);
  if (n < 0) return (void)call.fail_size("n", n);
  call.window(0, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(float), true, true);
  if (!call.stage()) return (void)call.status();
  const float alpha_v = call.scalar(alpha);
  if (!call.ok()) return (void)call.status();
  call.run([&] { exoblas::ops::scal<float>(call.ctx(), (int64_t)n, alpha_v, call.dev<float>(0), (int64_t)x.strides[0]); });
  return (void)call.finish();
}
extern "C" int exoblas_sscal_stride_any_async(hipStream_t stream, int_fast32_t n, const float *alpha, struct exo_win_1f32 x) {
  exoblas::Call call("exoblas_sscal_stride_any_async", exoblas::Call::kAsync, stream);
  if (n < 0) return call.fail_size("n", n);
  call.window(0, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(float), true, true);
  if (!call.stage()) return call.status();
  const float alpha_v = call.scalar(alpha);
  if (!call.ok()) return call.status();
  call.run([&] { exoblas::ops::scal<float>(call.ctx(), (int64_t)n, alpha_v, call.dev<float>(0), (int64_t)x.strides[0]); });
  return call.finish();
}

// src/level1/scal.py:9-12
extern "C" void exo_sscal_stride_1(void *ctxt, int_fast32_t n, const float *alpha, struct exo_win_1f32 x) {
  exoblas::Call call("exo_sscal_stride_1", exoblas::Call::kSync, ctxt);
  if (n < 0) return (void)call.fail_size("n", n);
  call.window(0, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(float), true, true);
  if (!call.stage()) return (void)call.status();
  const float alpha_v = call.scalar(alpha);
  if (!call.ok()) return (void)call.status();
  call.run([&] { exoblas::ops::scal<float>(call.ctx(), (int64_t)n, alpha_v, call.dev<float>(0), (int64_t)x.strides[0]); });
  return (void)call.finish();
}
extern "C" int exoblas_sscal_stride_1_async(hipStream_t stream, int_fast32_t n, const float *alpha, struct exo_win_1f32 x) {
  exoblas::Call call("exoblas_sscal_stride_1_async", exoblas::Call::kAsync, stream);
  if (n < 0) return call.fail_size("n", n);
  call.window(0, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(float), true, true);
  if (!call.stage()) return call.status();
  const float alpha_v = call.scalar(alpha);
  if (!call.ok()) return call.status();
  call.run([&] { exoblas::ops::scal<float>(call.ctx(), (int64_t)n, alpha_v, call.dev<float>(0), (int64_t)x.strides[0]); });
  return call.finish();
}

// src/level1/scal.py:9-12
extern "C" void exo_dscal_stride_any(void *ctxt, int_fast32_t n, const double *alpha, struct exo_win_1f64 x) {
  exoblas::Call call("exo_dscal_stride_any", exoblas::Call::kSync, ctxt);
  if (n < 0) return (void)call.fail_size("n", n);
  call.window(0, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(double), true, true);
  if (!call.stage()) return (void)call.status();
  const double alpha_v = call.scalar(alpha);
  if (!call.ok()) return (void)call.status();
  call.run([&] { exoblas::ops::scal<double>(call.ctx(), (int64_t)n, alpha_v, call.dev<double>(0), (int64_t)x.strides[0]); });
  return (void)call.finish();
}
extern "C" int exoblas_dscal_stride_any_async(hipStream_t stream, int_fast32_t n, const double *alpha, struct exo_win_1f64 x) {
  exoblas::Call call("exoblas_dscal_stride_any_async", exoblas::Call::kAsync, stream);
  if (n < 0) return call.fail_size("n", n);
  call.window(0, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(double), true, true);
  if (!call.stage()) return call.status();
  const double alpha_v = call.scalar(alpha);
  if (!call.ok()) return call.status();
  call.run([&] { exoblas::ops::scal<double>(call.ctx(), (int64_t)n, alpha_v, call.dev<double>(0), (int64_t)x.strides[0]); });
  return call.finish();
}

// src/level1/scal.py:9-12
extern "C" void exo_dscal_stride_1(void *ctxt, int_fast32_t n, const double *alpha, struct exo_win_1f64 x) {
  exoblas::Call call("exo_dscal_stride_1", exoblas::Call::kSync, ctxt);
  if (n < 0) return (void)call.fail_size("n", n);
  call.window(0, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(double), true, true);
  if (!call.stage()) return (void)call.status();
  const double alpha_v = call.scalar(alpha);
  if (!call.ok()) return (void)call.status();
  call.run([&] { exoblas::ops::scal<double>(call.ctx(), (int64_t)n, alpha_v, call.dev<double>(0), (int64_t)x.strides[0]); });
  return (void)call.finish();
}
extern "C" int exoblas_dscal_stride_1_async(hipStream_t stream, int_fast32_t n, const double *alpha, struct exo_win_1f64 x) {
  exoblas::Call call("exoblas_dscal_stride_1_async", exoblas::Call::kAsync, stream);
  if (n < 0) return call.fail_size("n", n);
  call.window(0, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(double), true, true);
  if (!call.stage()) return call.status();
  const double alpha_v = call.scalar(alpha);
  if (!call.ok()) return call.status();
  call.run([&] { exoblas::ops::scal<double>(call.ctx(), (int64_t)n, alpha_v, call.dev<double>(0), (int64_t)x.strides[0]); });
  return call.finish();
}

// src/level1/scal.py:15-19
extern "C" void exo_sscal_alpha_0_stride_any(void *ctxt, int_fast32_t n, struct exo_win_1f32 x) {
  exoblas::Call call("exo_sscal_alpha_0_stride_any", exoblas::Call::kSync, ctxt);
  if (n < 0) return (void)call.fail_size("n", n);
  call.window(0, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(float), true, true);
  if (!call.stage()) return (void)call.status();
  if (!call.ok()) return (void)call.status();
  call.run([&] { exoblas::ops::scal_alpha_0<float>(call.ctx(), (int64_t)n, call.dev<float>(0), (int64_t)x.strides[0]); });
  return (void)call.finish();
}
extern "C" int exoblas_sscal_alpha_0_stride_any_async(hipStream_t stream, int_fast32_t n, struct exo_win_1f32 x) {
  exoblas::Call call("exoblas_sscal_alpha_0_stride_any_async", exoblas::Call::kAsync, stream);
  if (n < 0) return call.fail_size("n", n);
  call.window(0, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(float), true, true);
  if (!call.stage()) return call.status();
  if (!call.ok()) return call.status();
  call.run([&] { exoblas::ops::scal_alpha_0<float>(call.ctx(), (int64_t)n, call.dev<float>(0), (int64_t)x.strides[0]); });
  return call.finish();
}

// src/level1/scal.py:15-19
extern "C" void exo_sscal_alpha_0_stride_1(void *ctxt, int_fast32_t n, struct exo_win_1f32 x) {
  exoblas::Call call("exo_sscal_alpha_0_stride_1", exoblas::Call::kSync, ctxt);
  if (n < 0) return (void)call.fail_size("n", n);
  call.window(0, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(float), true, true);
  if (!call.stage()) return (void)call.status();
  if (!call.ok()) return (void)call.status();
  call.run([&] { exoblas::ops::scal_alpha_0<float>(call.ctx(), (int64_t)n, call.dev<float>(0), (int64_t)x.strides[0]); });
  return (void)call.finish();
}
extern "C" int exoblas_sscal_alpha_0_stride_1_async(hipStream_t stream, int_fast32_t n, struct exo_win_1f32 x) {
  exoblas::Call call("exoblas_sscal_alpha_0_stride_1_async", exoblas::Call::kAsync, stream);
  if (n < 0) return call.fail_size("n", n);
  call.window(0, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(float), true, true);
  if (!call.stage()) return call.status();
  if (!call.ok()) return call.status();
  call.run([&] { exoblas::ops::scal_alpha_0<float>(call.ctx(), (int64_t)n, call.dev<float>(0), (int64_t)x.strides[0]); });
  return call.finish();
}

// src/level1/scal.py:15-19
extern "C" void exo_dscal_alpha_0_stride_any(void *ctxt, int_fast32_t n, struct exo_win_1f64 x) {
  exoblas::Call call("exo_dscal_alpha_0_stride_any", exoblas::Call::kSync, ctxt);
  if (n < 0) return (void)call.fail_size("n", n);
  call.window(0, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(double), true, true);
  if (!call.stage()) return (void)call.status();
  if (!call.ok()) return (void)call.status();
  call.run([&] { exoblas::ops::scal_alpha_0<double>(call.ctx(), (int64_t)n, call.dev<double>(0), (int64_t)x.strides[0]); });
  return (void)call.finish();
}
extern "C" int exoblas_dscal_alpha_0_stride_any_async(hipStream_t stream, int_fast32_t n, struct exo_win_1f64 x) {
  exoblas::Call call("exoblas_dscal_alpha_0_stride_any_async", exoblas::Call::kAsync, stream);
  if (n < 0) return call.fail_size("n", n);
  call.window(0, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(double), true, true);
  if (!call.stage()) return call.status();
  if (!call.ok()) return call.status();
  call.run([&] { exoblas::ops::scal_alpha_0<double>(call.ctx(), (int64_t)n, call.dev<double>(0), (int64_t)x.strides[0]); });
  return call.finish();
}

// src/level1/scal.py:15-19
extern "C" void exo_dscal_alpha_0_stride_1(void *ctxt, int_fast32_t n, struct exo_win_1f64 x) {
  exoblas::Call call("exo_dscal_alpha_0_stride_1", exoblas::Call::kSync, ctxt);
  if (n < 0) return (void)call.fail_size("n", n);
  call.window(0, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(double), true, true);
  if (!call.stage()) return (void)call.status();
  if (!call.ok()) return (void)call.status();
  call.run([&] { exoblas::ops::scal_alpha_0<double>(call.ctx(), (int64_t)n, call.dev<double>(0), (int64_t)x.strides[0]); });
  return (void)call.finish();
}
extern "C" int exoblas_dscal_alpha_0_stride_1_async(hipStream_t stream, int_fast32_t n, struct exo_win_1f64 x) {
  exoblas::Call call("exoblas_dscal_alpha_0_stride_1_async", exoblas::Call::kAsync, stream);
  if (n < 0) return call.fail_size("n", n);
  call.window(0, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(double), true, true);
  if (!call.stage()) return call.status();
  if (!call.ok()) return call.status();
  call.run([&] { exoblas::ops::scal_alpha_0<double>(call.ctx(), (int64_t)n, call.dev<double>(0), (int64_t)x.strides[0]); });
  return call.finish();
}

// src/level2/gemv.py:10-22
extern "C" void exo_sgemv_rm_stride_any(void *ctxt, int_fast32_t TransA, int_fast32_t M, int_fast32_t N, const float *alpha, const float *beta, struct exo_win_2f32c A, struct exo_win_2f32c AT, struct exo_win_1f32c x, struct exo_win_1f32 y) {
  exoblas::Call call("exo_sgemv_rm_stride_any", exoblas::Call::kSync, ctxt);
  if (!(TransA == 111 || TransA == 112 || TransA == 113)) return (void)call.fail_enum("TransA", TransA);
  if (M < 0) return (void)call.fail_size("M", M);
  if (N < 0) return (void)call.fail_size("N", N);
  if ((TransA == 111) && A.strides[1] != 1) return (void)call.fail_stride("A", A.strides[1]);
  if ((TransA == 111) && M > 1 && A.strides[0] < N) return (void)call.fail_ld("A", A.strides[0]);
  if ((TransA != 111) && AT.strides[1] != 1) return (void)call.fail_stride("AT", AT.strides[1]);
  if ((TransA != 111) && N > 1 && AT.strides[0] < M) return (void)call.fail_ld("AT", AT.strides[0]);
  call.window(0, A.data, 0, exoblas::span2_hi(M, N, A.strides[0]), sizeof(float), false, TransA == 111);
  call.window(1, AT.data, 0, exoblas::span2_hi(N, M, AT.strides[0]), sizeof(float), false, TransA != 111);
  call.window(2, x.data, exoblas::span_lo(N, x.strides[0]), exoblas::span_hi(N, x.strides[0]), sizeof(float), false, true);
  call.window(3, y.data, exoblas::span_lo(M, y.strides[0]), exoblas::span_hi(M, y.strides[0]), sizeof(float), true, true);
  if (!call.stage()) return (void)call.status();
  const float alpha_v = call.scalar(alpha);
  const float beta_v = call.scalar(beta);
  if (!call.ok()) return (void)call.status();
  call.run([&] { exoblas::ops::gemv_rm<float>(call.ctx(), (int64_t)TransA, (int64_t)M, (int64_t)N, alpha_v, beta_v, call.dev<const float>(0), (int64_t)A.strides[0], call.dev<const float>(1), (int64_t)AT.strides[0], call.dev<const float>(2), (int64_t)x.strides[0], call.dev<float>(3), (int64_t)y.strides[0]); });
  return (void)call.finish();
}
extern "C" int exoblas_sgemv_rm_stride_any_async(hipStream_t stream, int_fast32_t TransA, int_fast32_t M, int_fast32_t N, const float *alpha, const float *beta, struct exo_win_2f32c A, struct exo_win_2f32c AT, struct exo_win_1f32c x, struct exo_win_1f32 y) {
  exoblas::Call call("exoblas_sgemv_rm_stride_any_async", exoblas::Call::kAsync, stream);
  if (!(TransA == 111 || TransA == 112 || TransA == 113)) return call.fail_enum("TransA", TransA);
  if (M < 0) return call.fail_size("M", M);
  if (N < 0) return call.fail_size("N", N);
  if ((TransA == 111) && A.strides[1] != 1) return call.fail_stride("A", A.strides[1]);
  if ((TransA == 111) && M > 1 && A.strides[0] < N) return call.fail_ld("A", A.strides[0]);
  if ((TransA != 111) && AT.strides[1] != 1) return call.fail_stride("AT", AT.strides[1]);
  if ((TransA != 111) && N > 1 && AT.strides[0] < M) return call.fail_ld("AT", AT.strides[0]);
  call.window(0, A.data, 0, exoblas::span2_hi(M, N, A.strides[0]), sizeof(float), false, TransA == 111);
  call.window(1, AT.data, 0, exoblas::span2_hi(N, M, AT.strides[0]), sizeof(float), false, TransA != 111);
  call.window(2, x.data, exoblas::span_lo(N, x.strides[0]), exoblas::span_hi(N, x.strides[0]), sizeof(float), false, true);
  call.window(3, y.data, exoblas::span_lo(M, y.strides[0]), exoblas::span_hi(M, y.strides[0]), sizeof(float), true, true);
  if (!call.stage()) return call.status();
  const float alpha_v = call.scalar(alpha);
  const float beta_v = call.scalar(beta);
  if (!call.ok()) return call.status();
  call.run([&] { exoblas::ops::gemv_rm<float>(call.ctx(), (int64_t)TransA, (int64_t)M, (int64_t)N, alpha_v, beta_v, call.dev<const float>(0), (int64_t)A.strides[0], call.dev<const float>(1), (int64_t)AT.strides[0], call.dev<const float>(2), (int64_t)x.strides[0], call.dev<float>(3), (int64_t)y.strides[0]); });
  return call.finish();
}

// src/level2/gemv.py:10-22
extern "C" void exo_sgemv_rm_stride_1(void *ctxt, int_fast32_t TransA, int_fast32_t M, int_fast32_t N, const float *alpha, const float *beta, struct exo_win_2f32c A, struct exo_win_2f32c AT, struct exo_win_1f32c x, struct exo_win_1f32 y) {
  exoblas::Call call("exo_sgemv_rm_stride_1", exoblas::Call::kSync, ctxt);
  if (!(TransA == 111 || TransA == 112 || TransA == 113)) return (void)call.fail_enum("TransA", TransA);
  if (M < 0) return (void)call.fail_size("M", M);
  if (N < 0) return (void)call.fail_size("N", N);
  if ((TransA == 111) && A.strides[1] != 1) return (void)call.fail_stride("A", A.strides[1]);
  if ((TransA == 111) && M > 1 && A.strides[0] < N) return (void)call.fail_ld("A", A.strides[0]);
  if ((TransA != 111) && AT.strides[1] != 1) return (void)call.fail_stride("AT", AT.strides[1]);
  if ((TransA != 111) && N > 1 && AT.strides[0] < M) return (void)call.fail_ld("AT", AT.strides[0]);
  call.window(0, A.data, 0, exoblas::span2_hi(M, N, A.strides[0]), sizeof(float), false, TransA == 111);
  call.window(1, AT.data, 0, exoblas::span2_hi(N, M, AT.strides[0]), sizeof(float), false, TransA != 111);
  call.window(2, x.data, exoblas::span_lo(N, x.strides[0]), exoblas::span_hi(N, x.strides[0]), sizeof(float), false, true);
  call.window(3, y.data, exoblas::span_lo(M, y.strides[0]), exoblas::span_hi(M, y.strides[0]), sizeof(float), true, true);
  if (!call.stage()) return (void)call.status();
  const float alpha_v = call.scalar(alpha);
  const float beta_v = call.scalar(beta);
  if (!call.ok()) return (void)call.status();
  call.run([&] { exoblas::ops::gemv_rm<float>(call.ctx(), (int64_t)TransA, (int64_t)M, (int64_t)N, alpha_v, beta_v, call.dev<const float>(0), (int64_t)A.strides[0], call.dev<const float>(1), (int64_t)AT.strides[0], call.dev<const float>(2), (int64_t)x.strides[0], call.dev<float>(3), (int64_t)y.strides[0]); });
  return (void)call.finish();
}
extern "C" int exoblas_sgemv_rm_stride_1_async(hipStream_t stream, int_fast32_t TransA, int_fast32_t M, int_fast32_t N, const float *alpha, const float *beta, struct exo_win_2f32c A, struct exo_win_2f32c AT, struct exo_win_1f32c x, struct exo_win_1f32 y) {
  exoblas::Call call("exoblas_sgemv_rm_stride_1_async", exoblas::Call::kAsync, stream);
  if (!(TransA == 111 || TransA == 112 || TransA == 113)) return call.fail_enum("TransA", TransA);
  if (M < 0) return call.fail_size("M", M);
  if (N < 0) return call.fail_size("N", N);
  if ((TransA == 111) && A.strides[1] != 1) return call.fail_stride("A", A.strides[1]);
  if ((TransA == 111) && M > 1 && A.strides[0] < N) return call.fail_ld("A", A.strides[0]);
  if ((TransA != 111) && AT.strides[1] != 1) return call.fail_stride("AT", AT.strides[1]);
  if ((TransA != 111) && N > 1 && AT.strides[0] < M) return call.fail_ld("AT", AT.strides[0]);
  call.window(0, A.data, 0, exoblas::span2_hi(M, N, A.strides[0]), sizeof(float), false, TransA == 111);
  call.window(1, AT.data, 0, exoblas::span2_hi(N, M, AT.strides[0]), sizeof(float), false, TransA != 111);
  call.window(2, x.data, exoblas::span_lo(N, x.strides[0]), exoblas::span_hi(N, x.strides[0]), sizeof(float), false, true);
  call.window(3, y.data, exoblas::span_lo(M, y.strides[0]), exoblas::span_hi(M, y.strides[0]), sizeof(float), true, true);
  if (!call.stage()) return call.status();
  const float alpha_v = call.scalar(alpha);
  const float beta_v = call.scalar(beta);
  if (!call.ok()) return call.status();
  call.run([&] { exoblas::ops::gemv_rm<float>(call.ctx(), (int64_t)TransA, (int64_t)M, (int64_t)N, alpha_v, beta_v, call.dev<const float>(0), (int64_t)A.strides[0], call.dev<const float>(1), (int64_t)AT.strides[0], call.dev<const float>(2), (int64_t)x.strides[0], call.dev<float>(3), (int64_t)y.strides[0]); });
  return call.finish();
}

// src/level2/gemv.py:10-22
extern "C" void exo_dgemv_rm_stride_any(void *ctxt, int_fast32_t TransA, int_fast32_t M, int_fast32_t N, const double *alpha, const double *beta, struct exo_win_2f64c A, struct exo_win_2f64c AT, struct exo_win_1f64c x, struct exo_win_1f64 y) {
  exoblas::Call call("exo_dgemv_rm_stride_any", exoblas::Call::kSync, ctxt);
  if (!(TransA == 111 || TransA == 112 || TransA == 113)) return (void)call.fail_enum("TransA", TransA);
  if (M < 0) return (void)call.fail_size("M", M);
  if (N < 0) return (void)call.fail_size("N", N);
  if ((TransA == 111) && A.strides[1] != 1) return (void)call.fail_stride("A", A.strides[1]);
  if ((TransA == 111) && M > 1 && A.strides[0] < N) return (void)call.fail_ld("A", A.strides[0]);
  if ((TransA != 111) && AT.strides[1] != 1) return (void)call.fail_stride("AT", AT.strides[1]);
  if ((TransA != 111) && N > 1 && AT.strides[0] < M) return (void)call.fail_ld("AT", AT.strides[0]);
  call.window(0, A.data, 0, exoblas::span2_hi(M, N, A.strides[0]), sizeof(double), false, TransA == 111);
  call.window(1, AT.data, 0, exoblas::span2_hi(N, M, AT.strides[0]), sizeof(double), false, TransA != 111);
  call.window(2, x.data, exoblas::span_lo(N, x.strides[0]), exoblas::span_hi(N, x.strides[0]), sizeof(double), false, true);
  call.window(3, y.data, exoblas::span_lo(M, y.strides[0]), exoblas::span_hi(M, y.strides[0]), sizeof(double), true, true);
  if (!call.stage()) return (void)call.status();
  const double alpha_v = call.scalar(alpha);
  const double beta_v = call.scalar(beta);
  if (!call.ok()) return (void)call.status();
  call.run([&] { exoblas::ops::gemv_rm<double>(call.ctx(), (int64_t)TransA, (int64_t)M, (int64_t)N, alpha_v, beta_v, call.dev<const double>(0), (int64_t)A.strides[0], call.dev<const double>(1), (int64_t)AT.strides[0], call.dev<const double>(2), (int64_t)x.strides[0], call.dev<double>(3), (int64_t)y.strides[0]); });
  return (void)call.finish();
}
extern "C" int exoblas_dgemv_rm_stride_any_async(hipStream_t stream, int_fast32_t TransA, int_fast32_t M, int_fast32_t N, const double *alpha, const double *beta, struct exo_win_2f64c A, struct exo_win_2f64c AT, struct exo_win_1f64c x, struct exo_win_1f64 y) {
  exoblas::Call call("exoblas_dgemv_rm_stride_any_async", exoblas::Call::kAsync, stream);
  if (!(TransA == 111 || TransA == 112 || TransA == 113)) return call.fail_enum("TransA", TransA);
  if (M < 0) return call.fail_size("M", M);
  if (N < 0) return call.fail_size("N", N);
  if ((TransA == 111) && A.strides[1] != 1) return call.fail_stride("A", A.strides[1]);
  if ((TransA == 111) && M > 1 && A.strides[0] < N) return call.fail_ld("A", A.strides[0]);
  if ((TransA != 111) && AT.strides[1] != 1) return call.fail_stride("AT", AT.strides[1]);
  if ((TransA != 111) && N > 1 && AT.strides[0] < M) return call.fail_ld("AT", AT.strides[0]);
  call.window(0, A.data, 0, exoblas::span2_hi(M, N, A.strides[0]), sizeof(double), false, TransA == 111);
  call.window(1, AT.data, 0, exoblas::span2_hi(N, M, AT.strides[0]), sizeof(double), false, TransA != 111);
  call.window(2, x.data, exoblas::span_lo(N, x.strides[0]), exoblas::span_hi(N, x.strides[0]), sizeof(double), false, true);
  call.window(3, y.data, exoblas::span_lo(M, y.strides[0]), exoblas::span_hi(M, y.strides[0]), sizeof(double), true, true);
  if (!call.stage()) return call.status();
  const double alpha_v = call.scalar(alpha);
  const double beta_v = call.scalar(beta);
  if (!call.ok()) return call.status();
  call.run([&] { exoblas::ops::gemv_rm<double>(call.ctx(), (int64_t)TransA, (int64_t)M, (int64_t)N, alpha_v, beta_v, call.dev<const double>(0), (int64_t)A.strides[0], call.dev<const double>(1), (int64_t)AT.strides[0], call.dev<const double>(2), (int64_t)x.strides[0], call.dev<double>(3), (int64_t)y.strides[0]); });
  return call.finish();
}

// src/level2/gemv.py:10-22
extern "C" void exo_dgemv_rm_stride_1(void *ctxt, int_fast32_t TransA, int_fast32_t M, int_fast32_t N, const double *alpha, const double *beta, struct exo_win_2f64c A, struct exo_win_2f64c AT, struct exo_win_1f64c x, struct exo_win_1f64 y) {
  exoblas::Call call("exo_dgemv_rm_stride_1", exoblas::Call::kSync, ctxt);
  if (!(TransA == 111 || TransA == 112 || TransA == 113)) return (void)call.fail_enum("TransA", TransA);
  if (M < 0) return (void)call.fail_size("M", M);
  if (N < 0) return (void)call.fail_size("N", N);
  if ((TransA == 111) && A.strides[1] != 1) return (void)call.fail_stride("A", A.strides[1]);
  if ((TransA == 111) && M > 1 && A.strides[0] < N) return (void)call.fail_ld("A", A.strides[0]);
  if ((TransA != 111) && AT.strides[1] != 1) return (void)call.fail_stride("AT", AT.strides[1]);
  if ((TransA != 111) && N > 1 && AT.strides[0] < M) return (void)call.fail_ld("AT", AT.strides[0]);
  call.window(0, A.data, 0, exoblas::span2_hi(M, N, A.strides[0]), sizeof(double), false, TransA == 111);
  call.window(1, AT.data, 0, exoblas::span2_hi(N, M, AT.strides[0]), sizeof(double), false, TransA != 111);
  call.window(2, x.data, exoblas::span_lo(N, x.strides[0]), exoblas::span_hi(N, x.strides[0]), sizeof(double), false, true);
  call.window(3, y.data, exoblas::span_lo(M, y.strides[0]), exoblas::span_hi(M, y.strides[0]), sizeof(double), true, true);
  if (!call.stage()) return (void)call.status();
  const double alpha_v = call.scalar(alpha);
  const double beta_v = call.scalar(beta);
  if (!call.ok()) return (void)call.status();
  call.run([&] { exoblas::ops::gemv_rm<double>(call.ctx(), (int64_t)TransA, (int64_t)M, (int64_t)N, alpha_v, beta_v, call.dev<const double>(0), (int64_t)A.strides[0], call.dev<const double>(1), (int64_t)AT.strides[0], call.dev<const double>(2), (int64_t)x.strides[0], call.dev<double>(3), (int64_t)y.strides[0]); });
  return (void)call.finish();
}
extern "C" int exoblas_dgemv_rm_stride_1_async(hipStream_t stream, int_fast32_t TransA, int_fast32_t M, int_fast32_t N, const double *alpha, const double *beta, struct exo_win_2f64c A, struct exo_win_2f64c AT, struct exo_win_1f64c x, struct exo_win_1f64 y) {
  exoblas::Call call("exoblas_dgemv_rm_stride_1_async", exoblas::Call::kAsync, stream);
  if (!(TransA == 111 || TransA == 112 || TransA == 113)) return call.fail_enum("TransA", TransA);
  if (M < 0) return call.fail_size("M", M);
  if (N < 0) return call.fail_size("N", N);
  if ((TransA == 111) && A.strides[1] != 1) return call.fail_stride("A", A.strides[1]);
  if ((TransA == 111) && M > 1 && A.strides[0] < N) return call.fail_ld("A", A.strides[0]);
  if ((TransA != 111) && AT.strides[1] != 1) return call.fail_stride("AT", AT.strides[1]);
  if ((TransA != 111) && N > 1 && AT.strides[0] < M) return call.fail_ld("AT", AT.strides[0]);
  call.window(0, A.data, 0, exoblas::span2_hi(M, N, A.strides[0]), sizeof(double), false, TransA == 111);
  call.window(1, AT.data, 0, exoblas::span2_hi(N, M, AT.strides[0]), sizeof(double), false, TransA != 111);
  call.window(2, x.data, exoblas::span_lo(N, x.strides[0]), exoblas::span_hi(N, x.strides[0]), sizeof(double), false, true);
  call.window(3, y.data, exoblas::span_lo(M, y.strides[0]), exoblas::span_hi(M, y.strides[0]), sizeof(double), true, true);
  if (!call.stage()) return call.status();
  const double alpha_v = call.scalar(alpha);
  const double beta_v = call.scalar(beta);
  if (!call.ok()) return call.status();
  call.run([&] { exoblas::ops::gemv_rm<double>(call.ctx(), (int64_t)TransA, (int64_t)M, (int64_t)N, alpha_v, beta_v, call.dev<const double>(0), (int64_t)A.strides[0], call.dev<const double>(1), (int64_t)AT.strides[0], call.dev<const double>(2), (int64_t)x.strides[0], call.dev<double>(3), (int64_t)y.strides[0]); });
  return call.finish();
}

// src/level2/ger.py:9-15
extern "C" void exo_sger_rm_stride_any(void *ctxt, int_fast32_t m, int_fast32_t n, const float *alpha, struct exo_win_1f32c x, struct exo_win_1f32c y, struct exo_win_2f32 A) {
  exoblas::Call call("exo_sger_rm_stride_any", exoblas::Call::kSync, ctxt);
  if (m < 0) return (void)call.fail_size("m", m);
  if (n < 0) return (void)call.fail_size("n", n);
  if ((true) && A.strides[1] != 1) return (void)call.fail_stride("A", A.strides[1]);
  if ((true) && m > 1 && A.strides[0] < n) return (void)call.fail_ld("A", A.strides[0]);
  call.window(0, x.data, exoblas::span_lo(m, x.strides[0]), exoblas::span_hi(m, x.strides[0]), sizeof(float), false, true);
  call.window(1, y.data, exoblas::span_lo(n, y.strides[0]), exoblas::span_hi(n, y.strides[0]), sizeof(float), false, true);
  call.window(2, A.data, 0, exoblas::span2_hi(m, n, A.strides[0]), sizeof(float), true, true);
  if (!call.stage()) return (void)call.status();
  const float alpha_v = call.scalar(alpha);
  if (!call.ok()) return (void)call.status();
  call.run([&] { exoblas::ops::ger_rm<float>(call.ctx(), (int64_t)m, (int64_t)n, alpha_v, call.dev<const float>(0), (int64_t)x.strides[0], call.dev<const float>(1), (int64_t)y.strides[0], call.dev<float>(2), (int64_t)A.strides[0]); });
  return (void)call.finish();
}
extern "C" int exoblas_sger_rm_stride_any_async(hipStream_t stream, int_fast32_t m, int_fast32_t n, const float *alpha, struct exo_win_1f32c x, struct exo_win_1f32c y, struct exo_win_2f32 A) {
  exoblas::Call call("exoblas_sger_rm_stride_any_async", exoblas::Call::kAsync, stream);
  if (m < 0) return call.fail_size("m", m);
  if (n < 0) return call.fail_size("n", n);
  if ((true) && A.strides[1] != 1) return call.fail_stride("A", A.strides[1]);
  if ((true) && m > 1 && A.strides[0] < n) return call.fail_ld("A", A.strides[0]);
  call.window(0, x.data, exoblas::span_lo(m, x.strides[0]), exoblas::span_hi(m, x.strides[0]), sizeof(float), false, true);
  call.window(1, y.data, exoblas::span_lo(n, y.strides[0]), exoblas::span_hi(n, y.strides[0]), sizeof(float), false, true);
  call.window(2, A.data, 0, exoblas::span2_hi(m, n, A.strides[0]), sizeof(float), true, true);
  if (!call.stage()) return call.status();
  const float alpha_v = call.scalar(alpha);
  if (!call.ok()) return call.status();
  call.run([&] { exoblas::ops::ger_rm<float>(call.ctx(), (int64_t)m, (int64_t)n, alpha_v, call.dev<const float>(0), (int64_t)x.strides[0], call.dev<const float>(1), (int64_t)y.strides[0], call.dev<float>(2), (int64_t)A.strides[0]); });
  return call.finish();
}

// src/level2/ger.py:9-15
extern "C" void exo_sger_rm_stride_1(void *ctxt, int_fast32_t m, int_fast32_t n, const float *alpha, struct exo_win_1f32c x, struct exo_win_1f32c y, struct exo_win_2f32 A) {
  exoblas::Call call("exo_sger_rm_stride_1", exoblas::Call::kSync, ctxt);
  if (m < 0) return (void)call.fail_size("m", m);
  if (n < 0) return (void)call.fail_size("n", n);
  if ((true) && A.strides[1] != 1) return (void)call.fail_stride("A", A.strides[1]);
  if ((true) && m > 1 && A.strides[0] < n) return (void)call.fail_ld("A", A.strides[0]);
  call.window(0, x.data, exoblas::span_lo(m, x.strides[0]), exoblas::span_hi(m, x.strides[0]), sizeof(float), false, true);
  call.window(1, y.data, exoblas::span_lo(n, y.strides[0]), exoblas::span_hi(n, y.strides[0]), sizeof(float), false, true);
  call.window(2, A.data, 0, exoblas::span2_hi(m, n, A.strides[0]), sizeof(float), true, true);
  if (!call.stage()) return (void)call.status();
  const float alpha_v = call.scalar(alpha);
  if (!call.ok()) return (void)call.status();
  call.run([&] { exoblas::ops::ger_rm<float>(call.ctx(), (int64_t)m, (int64_t)n, alpha_v, call.dev<const float>(0), (int64_t)x.strides[0], call.dev<const float>(1), (int64_t)y.strides[0], call.dev<float>(2), (int64_t)A.strides[0]); });
  return (void)call.finish();
}
extern "C" int exoblas_sger_rm_stride_1_async(hipStream_t stream, int_fast32_t m, int_fast32_t n, const float *alpha, struct exo_win_1f32c x, struct exo_win_1f32c y, struct exo_win_2f32 A) {
  exoblas::Call call("exoblas_sger_rm_stride_1_async", exoblas::Call::kAsync, stream);
  if (m < 0) return call.fail_size("m", m);
  if (n < 0) return call.fail_size("n", n);
  if ((true) && A.strides[1] != 1) return call.fail_stride("A", A.strides[1]);
  if ((true) && m > 1 && A.strides[0] < n) return call.fail_ld("A", A.strides[0]);
  call.window(0, x.data, exoblas::span_lo(m, x.strides[0]), exoblas::span_hi(m, x.strides[0]), sizeof(float), false, true);
  call.window(1, y.data, exoblas::span_lo(n, y.strides[0]), exoblas::span_hi(n, y.strides[0]), sizeof(float), false, true);
  call.window(2, A.data, 0, exoblas::span2_hi(m, n, A.strides[0]), sizeof(float), true, true);
  if (!call.stage()) return call.status();
  const float alpha_v = call.scalar(alpha);
  if (!call.ok()) return call.status();
  call.run([&] { exoblas::ops::ger_rm<float>(call.ctx(), (int64_t)m, (int64_t)n, alpha_v, call.dev<const float>(0), (int64_t)x.strides[0], call.dev<const float>(1), (int64_t)y.strides[0], call.dev<float>(2), (int64_t)A.strides[0]); });
  return call.finish();
}

// src/level2/ger.py:9-15
extern "C" void exo_dger_rm_stride_any(void *ctxt, int_fast32_t m, int_fast32_t n, const double *alpha, struct exo_win_1f64c x, struct exo_win_1f64c y, struct exo_win_2f64 A) {
  exoblas::Call call("exo_dger_rm_stride_any", exoblas::Call::kSync, ctxt);
  if (m < 0) return (void)call.fail_size("m", m);
  if (n < 0) return (void)call.fail_size("n", n);
  if ((true) && A.strides[1] != 1) return (void)call.fail_stride("A", A.strides[1]);
  if ((true) && m > 1 && A.strides[0] < n) return (void)call.fail_ld("A", A.strides[0]);
  call.window(0, x.data, exoblas::span_lo(m, x.strides[0]), exoblas::span_hi(m, x.strides[0]), sizeof(double), false, true);
  call.window(1, y.data, exoblas::span_lo(n, y.strides[0]), exoblas::span_hi(n, y.strides[0]), sizeof(double), false, true);
  call.window(2, A.data, 0, exoblas::span2_hi(m, n, A.strides[0]), sizeof(double), true, true);
  if (!call.stage()) return (void)call.status();
  const double alpha_v = call.scalar(alpha);
  if (!call.ok()) return (void)call.status();
  call.run([&] { exoblas::ops::ger_rm<double>(call.ctx(), (int64_t)m, (int64_t)n, alpha_v, call.dev<const double>(0), (int64_t)x.strides[0], call.dev<const double>(1), (int64_t)y.strides[0], call.dev<double>(2), (int64_t)A.strides[0]); });
  return (void)call.finish();
}
extern "C" int exoblas_dger_rm_stride_any_async(hipStream_t stream, int_fast32_t m, int_fast32_t n, const double *alpha, struct exo_win_1f64c x, struct exo_win_1f64c y, struct exo_win_2f64 A) {
  exoblas::Call call("exoblas_dger_rm_stride_any_async", exoblas::Call::kAsync, stream);
  if (m < 0) return call.fail_size("m", m);
  if (n < 0) return call.fail_size("n", n);
  if ((true) && A.strides[1] != 1) return call.fail_stride("A", A.strides[1]);
  if ((true) && m > 1 && A.strides[0] < n) return call.fail_ld("A", A.strides[0]);
  call.window(0, x.data, exoblas::span_lo(m, x.strides[0]), exoblas::span_hi(m, x.strides[0]), sizeof(double), false, true);
  call.window(1, y.data, exoblas::span_lo(n, y.strides[0]), exoblas::span_hi(n, y.strides[0]), sizeof(double), false, true);
  call.window(2, A.data, 0, exoblas::span2_hi(m, n, A.strides[0]), sizeof(double), true, true);
  if (!call.stage()) return call.status();
  const double alpha_v = call.scalar(alpha);
  if (!call.ok()) return call.status();
  call.run([&] { exoblas::ops::ger_rm<double>(call.ctx(), (int64_t)m, (int64_t)n, alpha_v, call.dev<const double>(0), (int64_t)x.strides[0], call.dev<const double>(1), (int64_t)y.strides[0], call.dev<double>(2), (int64_t)A.strides[0]); });
  return call.finish();
}

// src/level2/ger.py:9-15
extern "C" void exo_dger_rm_stride_1(void *ctxt, int_fast32_t m, int_fast32_t n, const double *alpha, struct exo_win_1f64c x, struct exo_win_1f64c y, struct exo_win_2f64 A) {
  exoblas::Call call("exo_dger_rm_stride_1", exoblas::Call::kSync, ctxt);
  if (m < 0) return (void)call.fail_size("m", m);
  if (n < 0) return (void)call.fail_size("n", n);
  if ((true) && A.strides[1] != 1) return (void)call.fail_stride("A", A.strides[1]);
  if ((true) && m > 1 && A.strides[0] < n) return (void)call.fail_ld("A", A.strides[0]);
  call.window(0, x.data, exoblas::span_lo(m, x.strides[0]), exoblas::span_hi(m, x.strides[0]), sizeof(double), false, true);
  call.window(1, y.data, exoblas::span_lo(n, y.strides[0]), exoblas::span_hi(n, y.strides[0]), sizeof(double), false, true);
  call.window(2, A.data, 0, exoblas::span2_hi(m, n, A.strides[0]), sizeof(double), true, true);
  if (!call.stage()) return (void)call.status();
  const double alpha_v = call.scalar(alpha);
  if (!call.ok()) return (void)call.status();
  call.run([&] { exoblas::ops::ger_rm<double>(call.ctx(), (int64_t)m, (int64_t)n, alpha_v, call.dev<const double>(0), (int64_t)x.strides[0], call.dev<const double>(1), (int64_t)y.strides[0], call.dev<double>(2), (int64_t)A.strides[0]); });
  return (void)call.finish();
}
extern "C" int exoblas_dger_rm_stride_1_async(hipStream_t stream, int_fast32_t m, int_fast32_t n, const double *alpha, struct exo_win_1f64c x, struct exo_win_1f64c y, struct exo_win_2f64 A) {
  exoblas::Call call("exoblas_dger_rm_stride_1_async", exoblas::Call::kAsync, stream);
  if (m < 0) return call.fail_size("m", m);
  if (n < 0) return call.fail_size("n", n);
  if ((true) && A.strides[1] != 1) return call.fail_stride("A", A.strides[1]);
  if ((true) && m > 1 && A.strides[0] < n) return call.fail_ld("A", A.strides[0]);
  call.window(0, x.data, exoblas::span_lo(m, x.strides[0]), exoblas::span_hi(m, x.strides[0]), sizeof(double), false, true);
  call.window(1, y.data, exoblas::span_lo(n, y.strides[0]), exoblas::span_hi(n, y.strides[0]), sizeof(double), false, true);
  call.window(2, A.data, 0, exoblas::span2_hi(m, n, A.strides[0]), sizeof(double), true, true);
  if (!call.stage()) return call.status();
  const double alpha_v = call.scalar(alpha);
  if (!call.ok()) return call.status();
  call.run([&] { exoblas::ops::ger_rm<double>(call.ctx(), (int64_t)m, (int64_t)n, alpha_v, call.dev<const double>(0), (int64_t)x.strides[0], call.dev<const double>(1), (int64_t)y.strides[0], call.dev<double>(2), (int64_t)A.strides[0]); });
  return call.finish();
}

// src/level2/symv.py:10-35
extern "C" void exo_ssymv_rm_stride_any(void *ctxt, int_fast32_t Uplo, int_fast32_t n, const float *alpha, struct exo_win_2f32c A, struct exo_win_1f32c x, const float *beta, struct exo_win_1f32 y) {
  exoblas::Call call("exo_ssymv_rm_stride_any", exoblas::Call::kSync, ctxt);
  if (!(Uplo == 121 || Uplo == 122)) return (void)call.fail_enum("Uplo", Uplo);
  if (n < 0) return (void)call.fail_size("n", n);
  if ((true) && A.strides[1] != 1) return (void)call.fail_stride("A", A.strides[1]);
  if ((true) && n > 1 && A.strides[0] < n) return (void)call.fail_ld("A", A.strides[0]);
  call.window(0, A.data, 0, exoblas::span2_hi(n, n, A.strides[0]), sizeof(float), false, true);
  call.window(1, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(float), false, true);
  call.window(2, y.data, exoblas::span_lo(n, y.strides[0]), exoblas::span_hi(n, y.strides[0]), sizeof(float), true, true);
  if (!call.stage()) return (void)call.status();
  const float alpha_v = call.scalar(alpha);
  const float beta_v = call.scalar(beta);
  if (!call.ok()) return (void)call.status();
  call.run([&] { exoblas::ops::symv_rm<float>(call.ctx(), (int64_t)Uplo, (int64_t)n, alpha_v, call.dev<const float>(0), (int64_t)A.strides[0], call.dev<const float>(1), (int64_t)x.strides[0], beta_v, call.dev<float>(2), (int64_t)y.strides[0]); });
  return (void)call.finish();
}
extern "C" int exoblas_ssymv_rm_stride_any_async(hipStream_t stream, int_fast32_t Uplo, int_fast32_t n, const float *alpha, struct exo_win_2f32c A, struct exo_win_1f32c x, const float *beta, struct exo_win_1f32 y) {
  exoblas::Call call("exoblas_ssymv_rm_stride_any_async", exoblas::Call::kAsync, stream);
  if (!(Uplo == 121 || Uplo == 122)) return call.fail_enum("Uplo", Uplo);
  if (n < 0) return call.fail_size("n", n);
  if ((true) && A.strides[1] != 1) return call.fail_stride("A", A.strides[1]);
  if ((true) && n > 1 && A.strides[0] < n) return call.fail_ld("A", A.strides[0]);
  call.window(0, A.data, 0, exoblas::span2_hi(n, n, A.strides[0]), sizeof(float), false, true);
  call.window(1, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(float), false, true);
  call.window(2, y.data, exoblas::span_lo(n, y.strides[0]), exoblas::span_hi(n, y.strides[0]), sizeof(float), true, true);
  if (!call.stage()) return call.status();
  const float alpha_v = call.scalar(alpha);
  const float beta_v = call.scalar(beta);
  if (!call.ok()) return call.status();
  call.run([&] { exoblas::ops::symv_rm<float>(call.ctx(), (int64_t)Uplo, (int64_t)n, alpha_v, call.dev<const float>(0), (int64_t)A.strides[0], call.dev<const float>(1), (int64_t)x.strides[0], beta_v, call.dev<float>(2), (int64_t)y.strides[0]); });
  return call.finish();
}

// src/level2/symv.py:10-35
extern "C" void exo_ssymv_rm_stride_1(void *ctxt, int_fast32_t Uplo, int_fast32_t n, const float *alpha, struct exo_win_2f32c A, struct exo_win_1f32c x, const float *beta, struct exo_win_1f32 y) {
  exoblas::Call call("exo_ssymv_rm_stride_1", exoblas::Call::kSync, ctxt);
  if (!(Uplo == 121 || Uplo == 122)) return (void)call.fail_enum("Uplo", Uplo);
  if (n < 0) return (void)call.fail_size("n", n);
  if ((true) && A.strides[1] != 1) return (void)call.fail_stride("A", A.strides[1]);
  if ((true) && n > 1 && A.strides[0] < n) return (void)call.fail_ld("A", A.strides[0]);
  call.window(0, A.data, 0, exoblas::span2_hi(n, n, A.strides[0]), sizeof(float), false, true);
  call.window(1, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(float), false, true);
  call.window(2, y.data, exoblas::span_lo(n, y.strides[0]), exoblas::span_hi(n, y.strides[0]), sizeof(float), true, true);
  if (!call.stage()) return (void)call.status();
  const float alpha_v = call.scalar(alpha);
  const float beta_v = call.scalar(beta);
  if (!call.ok()) return (void)call.status();
  call.run([&] { exoblas::ops::symv_rm<float>(call.ctx(), (int64_t)Uplo, (int64_t)n, alpha_v, call.dev<const float>(0), (int64_t)A.strides[0], call.dev<const float>(1), (int64_t)x.strides[0], beta_v, call.dev<float>(2), (int64_t)y.strides[0]); });
  return (void)call.finish();
}
extern "C" int exoblas_ssymv_rm_stride_1_async(hipStream_t stream, int_fast32_t Uplo, int_fast32_t n, const float *alpha, struct exo_win_2f32c A, struct exo_win_1f32c x, const float *beta, struct exo_win_1f32 y) {
  exoblas::Call call("exoblas_ssymv_rm_stride_1_async", exoblas::Call::kAsync, stream);
  if (!(Uplo == 121 || Uplo == 122)) return call.fail_enum("Uplo", Uplo);
  if (n < 0) return call.fail_size("n", n);
  if ((true) && A.strides[1] != 1) return call.fail_stride("A", A.strides[1]);
  if ((true) && n > 1 && A.strides[0] < n) return call.fail_ld("A", A.strides[0]);
  call.window(0, A.data, 0, exoblas::span2_hi(n, n, A.strides[0]), sizeof(float), false, true);
  call.window(1, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(float), false, true);
  call.window(2, y.data, exoblas::span_lo(n, y.strides[0]), exoblas::span_hi(n, y.strides[0]), sizeof(float), true, true);
  if (!call.stage()) return call.status();
  const float alpha_v = call.scalar(alpha);
  const float beta_v = call.scalar(beta);
  if (!call.ok()) return call.status();
  call.run([&] { exoblas::ops::symv_rm<float>(call.ctx(), (int64_t)Uplo, (int64_t)n, alpha_v, call.dev<const float>(0), (int64_t)A.strides[0], call.dev<const float>(1), (int64_t)x.strides[0], beta_v, call.dev<float>(2), (int64_t)y.strides[0]); });
  return call.finish();
}

// src/level2/symv.py:10-35
extern "C" void exo_dsymv_rm_stride_any(void *ctxt, int_fast32_t Uplo, int_fast32_t n, const double *alpha, struct exo_win_2f64c A, struct exo_win_1f64c x, const double *beta, struct exo_win_1f64 y) {
  exoblas::Call call("exo_dsymv_rm_stride_any", exoblas::Call::kSync, ctxt);
  if (!(Uplo == 121 || Uplo == 122)) return (void)call.fail_enum("Uplo", Uplo);
  if (n < 0) return (void)call.fail_size("n", n);
  if ((true) && A.strides[1] != 1) return (void)call.fail_stride("A", A.strides[1]);
  if ((true) && n > 1 && A.strides[0] < n) return (void)call.fail_ld("A", A.strides[0]);
  call.window(0, A.data, 0, exoblas::span2_hi(n, n, A.strides[0]), sizeof(double), false, true);
  call.window(1, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(double), false, true);
  call.window(2, y.data, exoblas::span_lo(n, y.strides[0]), exoblas::span_hi(n, y.strides[0]), sizeof(double), true, true);
  if (!call.stage()) return (void)call.status();
  const double alpha_v = call.scalar(alpha);
  const double beta_v = call.scalar(beta);
  if (!call.ok()) return (void)call.status();
  call.run([&] { exoblas::ops::symv_rm<double>(call.ctx(), (int64_t)Uplo, (int64_t)n, alpha_v, call.dev<const double>(0), (int64_t)A.strides[0], call.dev<const double>(1), (int64_t)x.strides[0], beta_v, call.dev<double>(2), (int64_t)y.strides[0]); });
  return (void)call.finish();
}
extern "C" int exoblas_dsymv_rm_stride_any_async(hipStream_t stream, int_fast32_t Uplo, int_fast32_t n, const double *alpha, struct exo_win_2f64c A, struct exo_win_1f64c x, const double *beta, struct exo_win_1f64 y) {
  exoblas::Call call("exoblas_dsymv_rm_stride_any_async", exoblas::Call::kAsync, stream);
  if (!(Uplo == 121 || Uplo == 122)) return call.fail_enum("Uplo", Uplo);
  if (n < 0) return call.fail_size("n", n);
  if ((true) && A.strides[1] != 1) return call.fail_stride("A", A.strides[1]);
  if ((true) && n > 1 && A.strides[0] < n) return call.fail_ld("A", A.strides[0]);
  call.window(0, A.data, 0, exoblas::span2_hi(n, n, A.strides[0]), sizeof(double), false, true);
  call.window(1, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(double), false, true);
  call.window(2, y.data, exoblas::span_lo(n, y.strides[0]), exoblas::span_hi(n, y.strides[0]), sizeof(double), true, true);
  if (!call.stage()) return call.status();
  const double alpha_v = call.scalar(alpha);
  const double beta_v = call.scalar(beta);
  if (!call.ok()) return call.status();
  call.run([&] { exoblas::ops::symv_rm<double>(call.ctx(), (int64_t)Uplo, (int64_t)n, alpha_v, call.dev<const double>(0), (int64_t)A.strides[0], call.dev<const double>(1), (int64_t)x.strides[0], beta_v, call.dev<double>(2), (int64_t)y.strides[0]); });
  return call.finish();
}

// src/level2/symv.py:10-35
extern "C" void exo_dsymv_rm_stride_1(void *ctxt, int_fast32_t Uplo, int_fast32_t n, const double *alpha, struct exo_win_2f64c A, struct exo_win_1f64c x, const double *beta, struct exo_win_1f64 y) {
  exoblas::Call call("exo_dsymv_rm_stride_1", exoblas::Call::kSync, ctxt);
  if (!(Uplo == 121 || Uplo == 122)) return (void)call.fail_enum("Uplo", Uplo);
  if (n < 0) return (void)call.fail_size("n", n);
  if ((true) && A.strides[1] != 1) return (void)call.fail_stride("A", A.strides[1]);
  if ((true) && n > 1 && A.strides[0] < n) return (void)call.fail_ld("A", A.strides[0]);
  call.window(0, A.data, 0, exoblas::span2_hi(n, n, A.strides[0]), sizeof(double), false, true);
  call.window(1, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(double), false, true);
  call.window(2, y.data, exoblas::span_lo(n, y.strides[0]), exoblas::span_hi(n, y.strides[0]), sizeof(double), true, true);
  if (!call.stage()) return (void)call.status();
  const double alpha_v = call.scalar(alpha);
  const double beta_v = call.scalar(beta);
  if (!call.ok()) return (void)call.status();
  call.run([&] { exoblas::ops::symv_rm<double>(call.ctx(), (int64_t)Uplo, (int64_t)n, alpha_v, call.dev<const double>(0), (int64_t)A.strides[0], call.dev<const double>(1), (int64_t)x.strides[0], beta_v, call.dev<double>(2), (int64_t)y.strides[0]); });
  return (void)call.finish();
}
extern "C" int exoblas_dsymv_rm_stride_1_async(hipStream_t stream, int_fast32_t Uplo, int_fast32_t n, const double *alpha, struct exo_win_2f64c A, struct exo_win_1f64c x, const double *beta, struct exo_win_1f64 y) {
  exoblas::Call call("exoblas_dsymv_rm_stride_1_async", exoblas::Call::kAsync, stream);
  if (!(Uplo == 121 || Uplo == 122)) return call.fail_enum("Uplo", Uplo);
  if (n < 0) return call.fail_size("n", n);
  if ((true) && A.strides[1] != 1) return call.fail_stride("A", A.strides[1]);
  if ((true) && n > 1 && A.strides[0] < n) return call.fail_ld("A", A.strides[0]);
  call.window(0, A.data, 0, exoblas::span2_hi(n, n, A.strides[0]), sizeof(double), false, true);
  call.window(1, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(double), false, true);
  call.window(2, y.data, exoblas::span_lo(n, y.strides[0]), exoblas::span_hi(n, y.strides[0]), sizeof(double), true, true);
  if (!call.stage()) return call.status();
  const double alpha_v = call.scalar(alpha);
  const double beta_v = call.scalar(beta);
  if (!call.ok()) return call.status();
  call.run([&] { exoblas::ops::symv_rm<double>(call.ctx(), (int64_t)Uplo, (int64_t)n, alpha_v, call.dev<const double>(0), (int64_t)A.strides[0], call.dev<const double>(1), (int64_t)x.strides[0], beta_v, call.dev<double>(2), (int64_t)y.strides[0]); });
  return call.finish();
}

// src/level2/syr.py:10-19
extern "C" void exo_ssyr_rm_stride_any(void *ctxt, int_fast32_t Uplo, int_fast32_t n, const float *alpha, struct exo_win_1f32c x, struct exo_win_1f32c x_copy, struct exo_win_2f32 A) {
  exoblas::Call call("exo_ssyr_rm_stride_any", exoblas::Call::kSync, ctxt);
  if (!(Uplo == 121 || Uplo == 122)) return (void)call.fail_enum("Uplo", Uplo);
  if (n < 0) return (void)call.fail_size("n", n);
  if ((true) && A.strides[1] != 1) return (void)call.fail_stride("A", A.strides[1]);
  if ((true) && n > 1 && A.strides[0] < n) return (void)call.fail_ld("A", A.strides[0]);
  call.window(0, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(float), false, true);
  call.window(1, x_copy.data, exoblas::span_lo(n, x_copy.strides[0]), exoblas::span_hi(n, x_copy.strides[0]), sizeof(float), false, true);
  call.window(2, A.data, 0, exoblas::span2_hi(n, n, A.strides[0]), sizeof(float), true, true);
  if (!call.stage()) return (void)call.status();
  const float alpha_v = call.scalar(alpha);
  if (!call.ok()) return (void)call.status();
  call.run([&] { exoblas::ops::syr_rm<float>(call.ctx(), (int64_t)Uplo, (int64_t)n, alpha_v, call.dev<const float>(0), (int64_t)x.strides[0], call.dev<const float>(1), (int64_t)x_copy.strides[0], call.dev<float>(2), (int64_t)A.strides[0]); });
  return (void)call.finish();
}
extern "C" int exoblas_ssyr_rm_stride_any_async(hipStream_t stream, int_fast32_t Uplo, int_fast32_t n, const float *alpha, struct exo_win_1f32c x, struct exo_win_1f32c x_copy, struct exo_win_2f32 A) {
  exoblas::Call call("exoblas_ssyr_rm_stride_any_async", exoblas::Call::kAsync, stream);
  if (!(Uplo == 121 || Uplo == 122)) return call.fail_enum("Uplo", Uplo);
  if (n < 0) return call.fail_size("n", n);
  if ((true) && A.strides[1] != 1) return call.fail_stride("A", A.strides[1]);
  if ((true) && n > 1 && A.strides[0] < n) return call.fail_ld("A", A.strides[0]);
  call.window(0, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(float), false, true);
  call.window(1, x_copy.data, exoblas::span_lo(n, x_copy.strides[0]), exoblas::span_hi(n, x_copy.strides[0]), sizeof(float), false, true);
  call.window(2, A.data, 0, exoblas::span2_hi(n, n, A.strides[0]), sizeof(float), true, true);
  if (!call.stage()) return call.status();
  const float alpha_v = call.scalar(alpha);
  if (!call.ok()) return call.status();
  call.run([&] { exoblas::ops::syr_rm<float>(call.ctx(), (int64_t)Uplo, (int64_t)n, alpha_v, call.dev<const float>(0), (int64_t)x.strides[0], call.dev<const float>(1), (int64_t)x_copy.strides[0], call.dev<float>(2), (int64_t)A.strides[0]); });
  return call.finish();
}

// src/level2/syr.py:10-19
extern "C" void exo_ssyr_rm_stride_1(void *ctxt, int_fast32_t Uplo, int_fast32_t n, const float *alpha, struct exo_win_1f32c x, struct exo_win_1f32c x_copy, struct exo_win_2f32 A) {
  exoblas::Call call("exo_ssyr_rm_stride_1", exoblas::Call::kSync, ctxt);
  if (!(Uplo == 121 || Uplo == 122)) return (void)call.fail_enum("Uplo", Uplo);
  if (n < 0) return (void)call.fail_size("n", n);
  if ((true) && A.strides[1] != 1) return (void)call.fail_stride("A", A.strides[1]);
  if ((true) && n > 1 && A.strides[0] < n) return (void)call.fail_ld("A", A.strides[0]);
  call.window(0, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(float), false, true);
  call.window(1, x_copy.data, exoblas::span_lo(n, x_copy.strides[0]), exoblas::span_hi(n, x_copy.strides[0]), sizeof(float), false, true);
  call.window(2, A.data, 0, exoblas::span2_hi(n, n, A.strides[0]), sizeof(float), true, true);
  if (!call.stage()) return (void)call.status();
  const float alpha_v = call.scalar(alpha);
  if (!call.ok()) return (void)call.status();
  call.run([&] { exoblas::ops::syr_rm<float>(call.ctx(), (int64_t)Uplo, (int64_t)n, alpha_v, call.dev<const float>(0), (int64_t)x.strides[0], call.dev<const float>(1), (int64_t)x_copy.strides[0], call.dev<float>(2), (int64_t)A.strides[0]); });
  return (void)call.finish();
}
extern "C" int exoblas_ssyr_rm_stride_1_async(hipStream_t stream, int_fast32_t Uplo, int_fast32_t n, const float *alpha, struct exo_win_1f32c x, struct exo_win_1f32c x_copy, struct exo_win_2f32 A) {
  exoblas::Call call("exoblas_ssyr_rm_stride_1_async", exoblas::Call::kAsync, stream);
  if (!(Uplo == 121 || Uplo == 122)) return call.fail_enum("Uplo", Uplo);
  if (n < 0) return call.fail_size("n", n);
  if ((true) && A.strides[1] != 1) return call.fail_stride("A", A.strides[1]);
  if ((true) && n > 1 && A.strides[0] < n) return call.fail_ld("A", A.strides[0]);
  call.window(0, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(float), false, true);
  call.window(1, x_copy.data, exoblas::span_lo(n, x_copy.strides[0]), exoblas::span_hi(n, x_copy.strides[0]), sizeof(float), false, true);
  call.window(2, A.data, 0, exoblas::span2_hi(n, n, A.strides[0]), sizeof(float), true, true);
  if (!call.stage()) return call.status();
  const float alpha_v = call.scalar(alpha);
  if (!call.ok()) return call.status();
  call.run([&] { exoblas::ops::syr_rm<float>(call.ctx(), (int64_t)Uplo, (int64_t)n, alpha_v, call.dev<const float>(0), (int64_t)x.strides[0], call.dev<const float>(1), (int64_t)x_copy.strides[0], call.dev<float>(2), (int64_t)A.strides[0]); });
  return call.finish();
}

// src/level2/syr.py:10-19
extern "C" void exo_dsyr_rm_stride_any(void *ctxt, int_fast32_t Uplo, int_fast32_t n, const double *alpha, struct exo_win_1f64c x, struct exo_win_1f64c x_copy, struct exo_win_2f64 A) {
  exoblas::Call call("exo_dsyr_rm_stride_any", exoblas::Call::kSync, ctxt);
  if (!(Uplo == 121 || Uplo == 122)) return (void)call.fail_enum("Uplo", Uplo);
  if (n < 0) return (void)call.fail_size("n", n);
  if ((true) && A.strides[1] != 1) return (void)call.fail_stride("A", A.strides[1]);
  if ((true) && n > 1 && A.strides[0] < n) return (void)call.fail_ld("A", A.strides[0]);
  call.window(0, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(double), false, true);
  call.window(1, x_copy.data, exoblas::span_lo(n, x_copy.strides[0]), exoblas::span_hi(n, x_copy.strides[0]), sizeof(double), false, true);
  call.window(2, A.data, 0, exoblas::span2_hi(n, n, A.strides[0]), sizeof(double), true, true);
  if (!call.stage()) return (void)call.status();
  const double alpha_v = call.scalar(alpha);
  if (!call.ok()) return (void)call.status();
  call.run([&] { exoblas::ops::syr_rm<double>(call.ctx(), (int64_t)Uplo, (int64_t)n, alpha_v, call.dev<const double>(0), (int64_t)x.strides[0], call.dev<const double>(1), (int64_t)x_copy.strides[0], call.dev<double>(2), (int64_t)A.strides[0]); });
  return (void)call.finish();
}
extern "C" int exoblas_dsyr_rm_stride_any_async(hipStream_t stream, int_fast32_t Uplo, int_fast32_t n, const double *alpha, struct exo_win_1f64c x, struct exo_win_1f64c x_copy, struct exo_win_2f64 A) {
  exoblas::Call call("exoblas_dsyr_rm_stride_any_async", exoblas::Call::kAsync, stream);
  if (!(Uplo == 121 || Uplo == 122)) return call.fail_enum("Uplo", Uplo);
  if (n < 0) return call.fail_size("n", n);
  if ((true) && A.strides[1] != 1) return call.fail_stride("A", A.strides[1]);
  if ((true) && n > 1 && A.strides[0] < n) return call.fail_ld("A", A.strides[0]);
  call.window(0, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(double), false, true);
  call.window(1, x_copy.data, exoblas::span_lo(n, x_copy.strides[0]), exoblas::span_hi(n, x_copy.strides[0]), sizeof(double), false, true);
  call.window(2, A.data, 0, exoblas::span2_hi(n, n, A.strides[0]), sizeof(double), true, true);
  if (!call.stage()) return call.status();
  const double alpha_v = call.scalar(alpha);
  if (!call.ok()) return call.status();
  call.run([&] { exoblas::ops::syr_rm<double>(call.ctx(), (int64_t)Uplo, (int64_t)n, alpha_v, call.dev<const double>(0), (int64_t)x.strides[0], call.dev<const double>(1), (int64_t)x_copy.strides[0], call.dev<double>(2), (int64_t)A.strides[0]); });
  return call.finish();
}

// src/level2/syr.py:10-19
extern "C" void exo_dsyr_rm_stride_1(void *ctxt, int_fast32_t Uplo, int_fast32_t n, const double *alpha, struct exo_win_1f64c x, struct exo_win_1f64c x_copy, struct exo_win_2f64 A) {
  exoblas::Call call("exo_dsyr_rm_stride_1", exoblas::Call::kSync, ctxt);
  if (!(Uplo == 121 || Uplo == 122)) return (void)call.fail_enum("Uplo", Uplo);
  if (n < 0) return (void)call.fail_size("n", n);
  if ((true) && A.strides[1] != 1) return (void)call.fail_stride("A", A.strides[1]);
  if ((true) && n > 1 && A.strides[0] < n) return (void)call.fail_ld("A", A.strides[0]);
  call.window(0, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(double), false, true);
  call.window(1, x_copy.data, exoblas::span_lo(n, x_copy.strides[0]), exoblas::span_hi(n, x_copy.strides[0]), sizeof(double), false, true);
  call.window(2, A.data, 0, exoblas::span2_hi(n, n, A.strides[0]), sizeof(double), true, true);
  if (!call.stage()) return (void)call.status();
  const double alpha_v = call.scalar(alpha);
  if (!call.ok()) return (void)call.status();
  call.run([&] { exoblas::ops::syr_rm<double>(call.ctx(), (int64_t)Uplo, (int64_t)n, alpha_v, call.dev<const double>(0), (int64_t)x.strides[0], call.dev<const double>(1), (int64_t)x_copy.strides[0], call.dev<double>(2), (int64_t)A.strides[0]); });
  return (void)call.finish();
}
extern "C" int exoblas_dsyr_rm_stride_1_async(hipStream_t stream, int_fast32_t Uplo, int_fast32_t n, const double *alpha, struct exo_win_1f64c x, struct exo_win_1f64c x_copy, struct exo_win_2f64 A) {
  exoblas::Call call("exoblas_dsyr_rm_stride_1_async", exoblas::Call::kAsync, stream);
  if (!(Uplo == 121 || Uplo == 122)) return call.fail_enum("Uplo", Uplo);
  if (n < 0) return call.fail_size("n", n);
  if ((true) && A.strides[1] != 1) return call.fail_stride("A", A.strides[1]);
  if ((true) && n > 1 && A.strides[0] < n) return call.fail_ld("A", A.strides[0]);
  call.window(0, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(double), false, true);
  call.window(1, x_copy.data, exoblas::span_lo(n, x_copy.strides[0]), exoblas::span_hi(n, x_copy.strides[0]), sizeof(double), false, true);
  call.window(2, A.data, 0, exoblas::span2_hi(n, n, A.strides[0]), sizeof(double), true, true);
  if (!call.stage()) return call.status();
  const double alpha_v = call.scalar(alpha);
  if (!call.ok()) return call.status();
  call.run([&] { exoblas::ops::syr_rm<double>(call.ctx(), (int64_t)Uplo, (int64_t)n, alpha_v, call.dev<const double>(0), (int64_t)x.strides[0], call.dev<const double>(1), (int64_t)x_copy.strides[0], call.dev<double>(2), (int64_t)A.strides[0]); });
  return call.finish();
}

// src/level2/syr2.py:9-27
extern "C" void exo_ssyr2_rm_stride_any(void *ctxt, int_fast32_t Uplo, int_fast32_t n, const float *alpha, struct exo_win_1f32c x, struct exo_win_1f32c x_copy, struct exo_win_1f32c y, struct exo_win_1f32c y_copy, struct exo_win_2f32 A) {
  exoblas::Call call("exo_ssyr2_rm_stride_any", exoblas::Call::kSync, ctxt);
  if (!(Uplo == 121 || Uplo == 122)) return (void)call.fail_enum("Uplo", Uplo);
  if (n < 0) return (void)call.fail_size("n", n);
  if ((true) && A.strides[1] != 1) return (void)call.fail_stride("A", A.strides[1]);
  if ((true) && n > 1 && A.strides[0] < n) return (void)call.fail_ld("A", A.strides[0]);
  call.window(0, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(float), false, true);
  call.window(1, x_copy.data, exoblas::span_lo(n, x_copy.strides[0]), exoblas::span_hi(n, x_copy.strides[0]), sizeof(float), false, true);
  call.window(2, y.data, exoblas::span_lo(n, y.strides[0]), exoblas::span_hi(n, y.strides[0]), sizeof(float), false, true);
  call.window(3, y_copy.data, exoblas::span_lo(n, y_copy.strides[0]), exoblas::span_hi(n, y_copy.strides[0]), sizeof(float), false, true);
  call.window(4, A.data, 0, exoblas::span2_hi(n, n, A.strides[0]), sizeof(float), true, true);
  if (!call.stage()) return (void)call.status();
  const float alpha_v = call.scalar(alpha);
  if (!call.ok()) return (void)call.status();
  call.run([&] { exoblas::ops::syr2_rm<float>(call.ctx(), (int64_t)Uplo, (int64_t)n, alpha_v, call.dev<const float>(0), (int64_t)x.strides[0], call.dev<const float>(1), (int64_t)x_copy.strides[0], call.dev<const float>(2), (int64_t)y.strides[0], call.dev<const float>(3), (int64_t)y_copy.strides[0], call.dev<float>(4), (int64_t)A.strides[0]); });
  return (void)call.finish();
}
extern "C" int exoblas_ssyr2_rm_stride_any_async(hipStream_t stream, int_fast32_t Uplo, int_fast32_t n, const float *alpha, struct exo_win_1f32c x, struct exo_win_1f32c x_copy, struct exo_win_1f32c y, struct exo_win_1f32c y_copy, struct exo_win_2f32 A) {
  exoblas::Call call("exoblas_ssyr2_rm_stride_any_async", exoblas::Call::kAsync, stream);
  if (!(Uplo == 121 || Uplo == 122)) return call.fail_enum("Uplo", Uplo);
  if (n < 0) return call.fail_size("n", n);
  if ((true) && A.strides[1] != 1) return call.fail_stride("A", A.strides[1]);
  if ((true) && n > 1 && A.strides[0] < n) return call.fail_ld("A", A.strides[0]);
  call.window(0, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(float), false, true);
  call.window(1, x_copy.data, exoblas::span_lo(n, x_copy.strides[0]), exoblas::span_hi(n, x_copy.strides[0]), sizeof(float), false, true);
  call.window(2, y.data, exoblas::span_lo(n, y.strides[0]), exoblas::span_hi(n, y.strides[0]), sizeof(float), false, true);
  call.window(3, y_copy.data, exoblas::span_lo(n, y_copy.strides[0]), exoblas::span_hi(n, y_copy.strides[0]), sizeof(float), false, true);
  call.window(4, A.data, 0, exoblas::span2_hi(n, n, A.strides[0]), sizeof(float), true, true);
  if (!call.stage()) return call.status();
  const float alpha_v = call.scalar(alpha);
  if (!call.ok()) return call.status();
  call.run([&] { exoblas::ops::syr2_rm<float>(call.ctx(), (int64_t)Uplo, (int64_t)n, alpha_v, call.dev<const float>(0), (int64_t)x.strides[0], call.dev<const float>(1), (int64_t)x_copy.strides[0], call.dev<const float>(2), (int64_t)y.strides[0], call.dev<const float>(3), (int64_t)y_copy.strides[0], call.dev<float>(4), (int64_t)A.strides[0]); });
  return call.finish();
}

// src/level2/syr2.py:9-27
extern "C" void exo_ssyr2_rm_stride_1(void *ctxt, int_fast32_t Uplo, int_fast32_t n, const float *alpha, struct exo_win_1f32c x, struct exo_win_1f32c x_copy, struct exo_win_1f32c y, struct exo_win_1f32c y_copy, struct exo_win_2f32 A) {
  exoblas::Call call("exo_ssyr2_rm_stride_1", exoblas::Call::kSync, ctxt);
  if (!(Uplo == 121 || Uplo == 122)) return (void)call.fail_enum("Uplo", Uplo);
  if (n < 0) return (void)call.fail_size("n", n);
  if ((true) && A.strides[1] != 1) return (void)call.fail_stride("A", A.strides[1]);
  if ((true) && n > 1 && A.strides[0] < n) return (void)call.fail_ld("A", A.strides[0]);
  call.window(0, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(float), false, true);
  call.window(1, x_copy.data, exoblas::span_lo(n, x_copy.strides[0]), exoblas::span_hi(n, x_copy.strides[0]), sizeof(float), false, true);
  call.window(2, y.data, exoblas::span_lo(n, y.strides[0]), exoblas::span_hi(n, y.strides[0]), sizeof(float), false, true);
  call.window(3, y_copy.data, exoblas::span_lo(n, y_copy.strides[0]), exoblas::span_hi(n, y_copy.strides[0]), sizeof(float), false, true);
  call.window(4, A.data, 0, exoblas::span2_hi(n, n, A.strides[0]), sizeof(float), true, true);
  if (!call.stage()) return (void)call.status();
  const float alpha_v = call.scalar(alpha);
  if (!call.ok()) return (void)call.status();
  call.run([&] { exoblas::ops::syr2_rm<float>(call.ctx(), (int64_t)Uplo, (int64_t)n, alpha_v, call.dev<const float>(0), (int64_t)x.strides[0], call.dev<const float>(1), (int64_t)x_copy.strides[0], call.dev<const float>(2), (int64_t)y.strides[0], call.dev<const float>(3), (int64_t)y_copy.strides[0], call.dev<float>(4), (int64_t)A.strides[0]); });
  return (void)call.finish();
}
extern "C" int exoblas_ssyr2_rm_stride_1_async(hipStream_t stream, int_fast32_t Uplo, int_fast32_t n, const float *alpha, struct exo_win_1f32c x, struct exo_win_1f32c x_copy, struct exo_win_1f32c y, struct exo_win_1f32c y_copy, struct exo_win_2f32 A) {
  exoblas::Call call("exoblas_ssyr2_rm_stride_1_async", exoblas::Call::kAsync, stream);
  if (!(Uplo == 121 || Uplo == 122)) return call.fail_enum("Uplo", Uplo);
  if (n < 0) return call.fail_size("n", n);
  if ((true) && A.strides[1] != 1) return call.fail_stride("A", A.strides[1]);
  if ((true) && n > 1 && A.strides[0] < n) return call.fail_ld("A", A.strides[0]);
  call.window(0, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(float), false, true);
  call.window(1, x_copy.data, exoblas::span_lo(n, x_copy.strides[0]), exoblas::span_hi(n, x_copy.strides[0]), sizeof(float), false, true);
  call.window(2, y.data, exoblas::span_lo(n, y.strides[0]), exoblas::span_hi(n, y.strides[0]), sizeof(float), false, true);
  call.window(3, y_copy.data, exoblas::span_lo(n, y_copy.strides[0]), exoblas::span_hi(n, y_copy.strides[0]), sizeof(float), false, true);
  call.window(4, A.data, 0, exoblas::span2_hi(n, n, A.strides[0]), sizeof(float), true, true);
  if (!call.stage()) return call.status();
  const float alpha_v = call.scalar(alpha);
  if (!call.ok()) return call.status();
  call.run([&] { exoblas::ops::syr2_rm<float>(call.ctx(), (int64_t)Uplo, (int64_t)n, alpha_v, call.dev<const float>(0), (int64_t)x.strides[0], call.dev<const float>(1), (int64_t)x_copy.strides[0], call.dev<const float>(2), (int64_t)y.strides[0], call.dev<const float>(3), (int64_t)y_copy.strides[0], call.dev<float>(4), (int64_t)A.strides[0]); });
  return call.finish();
}

// src/level2/syr2.py:9-27
extern "C" void exo_dsyr2_rm_stride_any(void *ctxt, int_fast32_t Uplo, int_fast32_t n, const double *alpha, struct exo_win_1f64c x, struct exo_win_1f64c x_copy, struct exo_win_1f64c y, struct exo_win_1f64c y_copy, struct exo_win_2f64 A) {
  exoblas::Call call("exo_dsyr2_rm_stride_any", exoblas::Call::kSync, ctxt);
  if (!(Uplo == 121 || Uplo == 122)) return (void)call.fail_enum("Uplo", Uplo);
  if (n < 0) return (void)call.fail_size("n", n);
  if ((true) && A.strides[1] != 1) return (void)call.fail_stride("A", A.strides[1]);
  if ((true) && n > 1 && A.strides[0] < n) return (void)call.fail_ld("A", A.strides[0]);
  call.window(0, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(double), false, true);
  call.window(1, x_copy.data, exoblas::span_lo(n, x_copy.strides[0]), exoblas::span_hi(n, x_copy.strides[0]), sizeof(double), false, true);
  call.window(2, y.data, exoblas::span_lo(n, y.strides[0]), exoblas::span_hi(n, y.strides[0]), sizeof(double), false, true);
  call.window(3, y_copy.data, exoblas::span_lo(n, y_copy.strides[0]), exoblas::span_hi(n, y_copy.strides[0]), sizeof(double), false, true);
  call.window(4, A.data, 0, exoblas::span2_hi(n, n, A.strides[0]), sizeof(double), true, true);
  if (!call.stage()) return (void)call.status();
  const double alpha_v = call.scalar(alpha);
  if (!call.ok()) return (void)call.status();
  call.run([&] { exoblas::ops::syr2_rm<double>(call.ctx(), (int64_t)Uplo, (int64_t)n, alpha_v, call.dev<const double>(0), (int64_t)x.strides[0], call.dev<const double>(1), (int64_t)x_copy.strides[0], call.dev<const double>(2), (int64_t)y.strides[0], call.dev<const double>(3), (int64_t)y_copy.strides[0], call.dev<double>(4), (int64_t)A.strides[0]); });
  return (void)call.finish();
}
extern "C" int exoblas_dsyr2_rm_stride_any_async(hipStream_t stream, int_fast32_t Uplo, int_fast32_t n, const double *alpha, struct exo_win_1f64c x, struct exo_win_1f64c x_copy, struct exo_win_1f64c y, struct exo_win_1f64c y_copy, struct exo_win_2f64 A) {
  exoblas::Call call("exoblas_dsyr2_rm_stride_any_async", exoblas::Call::kAsync, stream);
  if (!(Uplo == 121 || Uplo == 122)) return call.fail_enum("Uplo", Uplo);
  if (n < 0) return call.fail_size("n", n);
  if ((true) && A.strides[1] != 1) return call.fail_stride("A", A.strides[1]);
  if ((true) && n > 1 && A.strides[0] < n) return call.fail_ld("A", A.strides[0]);
  call.window(0, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(double), false, true);
  call.window(1, x_copy.data, exoblas::span_lo(n, x_copy.strides[0]), exoblas::span_hi(n, x_copy.strides[0]), sizeof(double), false, true);
  call.window(2, y.data, exoblas::span_lo(n, y.strides[0]), exoblas::span_hi(n, y.strides[0]), sizeof(double), false, true);
  call.window(3, y_copy.data, exoblas::span_lo(n, y_copy.strides[0]), exoblas::span_hi(n, y_copy.strides[0]), sizeof(double), false, true);
  call.window(4, A.data, 0, exoblas::span2_hi(n, n, A.strides[0]), sizeof(double), true, true);
  if (!call.stage()) return call.status();
  const double alpha_v = call.scalar(alpha);
  if (!call.ok()) return call.status();
  call.run([&] { exoblas::ops::syr2_rm<double>(call.ctx(), (int64_t)Uplo, (int64_t)n, alpha_v, call.dev<const double>(0), (int64_t)x.strides[0], call.dev<const double>(1), (int64_t)x_copy.strides[0], call.dev<const double>(2), (int64_t)y.strides[0], call.dev<const double>(3), (int64_t)y_copy.strides[0], call.dev<double>(4), (int64_t)A.strides[0]); });
  return call.finish();
}

// src/level2/syr2.py:9-27
extern "C" void exo_dsyr2_rm_stride_1(void *ctxt, int_fast32_t Uplo, int_fast32_t n, const double *alpha, struct exo_win_1f64c x, struct exo_win_1f64c x_copy, struct exo_win_1f64c y, struct exo_win_1f64c y_copy, struct exo_win_2f64 A) {
  exoblas::Call call("exo_dsyr2_rm_stride_1", exoblas::Call::kSync, ctxt);
  if (!(Uplo == 121 || Uplo == 122)) return (void)call.fail_enum("Uplo", Uplo);
  if (n < 0) return (void)call.fail_size("n", n);
  if ((true) && A.strides[1] != 1) return (void)call.fail_stride("A", A.strides[1]);
  if ((true) && n > 1 && A.strides[0] < n) return (void)call.fail_ld("A", A.strides[0]);
  call.window(0, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(double), false, true);
  call.window(1, x_copy.data, exoblas::span_lo(n, x_copy.strides[0]), exoblas::span_hi(n, x_copy.strides[0]), sizeof(double), false, true);
  call.window(2, y.data, exoblas::span_lo(n, y.strides[0]), exoblas::span_hi(n, y.strides[0]), sizeof(double), false, true);
  call.window(3, y_copy.data, exoblas::span_lo(n, y_copy.strides[0]), exoblas::span_hi(n, y_copy.strides[0]), sizeof(double), false, true);
  call.window(4, A.data, 0, exoblas::span2_hi(n, n, A.strides[0]), sizeof(double), true, true);
  if (!call.stage()) return (void)call.status();
  const double alpha_v = call.scalar(alpha);
  if (!call.ok()) return (void)call.status();
  call.run([&] { exoblas::ops::syr2_rm<double>(call.ctx(), (int64_t)Uplo, (int64_t)n, alpha_v, call.dev<const double>(0), (int64_t)x.strides[0], call.dev<const double>(1), (int64_t)x_copy.strides[0], call.dev<const double>(2), (int64_t)y.strides[0], call.dev<const double>(3), (int64_t)y_copy.strides[0], call.dev<double>(4), (int64_t)A.strides[0]); });
  return (void)call.finish();
}
extern "C" int exoblas_dsyr2_rm_stride_1_async(hipStream_t stream, int_fast32_t Uplo, int_fast32_t n, const double *alpha, struct exo_win_1f64c x, struct exo_win_1f64c x_copy, struct exo_win_1f64c y, struct exo_win_1f64c y_copy, struct exo_win_2f64 A) {
  exoblas::Call call("exoblas_dsyr2_rm_stride_1_async", exoblas::Call::kAsync, stream);
  if (!(Uplo == 121 || Uplo == 122)) return call.fail_enum("Uplo", Uplo);
  if (n < 0) return call.fail_size("n", n);
  if ((true) && A.strides[1] != 1) return call.fail_stride("A", A.strides[1]);
  if ((true) && n > 1 && A.strides[0] < n) return call.fail_ld("A", A.strides[0]);
  call.window(0, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(double), false, true);
  call.window(1, x_copy.data, exoblas::span_lo(n, x_copy.strides[0]), exoblas::span_hi(n, x_copy.strides[0]), sizeof(double), false, true);
  call.window(2, y.data, exoblas::span_lo(n, y.strides[0]), exoblas::span_hi(n, y.strides[0]), sizeof(double), false, true);
  call.window(3, y_copy.data, exoblas::span_lo(n, y_copy.strides[0]), exoblas::span_hi(n, y_copy.strides[0]), sizeof(double), false, true);
  call.window(4, A.data, 0, exoblas::span2_hi(n, n, A.strides[0]), sizeof(double), true, true);
  if (!call.stage()) return call.status();
  const double alpha_v = call.scalar(alpha);
  if (!call.ok()) return call.status();
  call.run([&] { exoblas::ops::syr2_rm<double>(call.ctx(), (int64_t)Uplo, (int64_t)n, alpha_v, call.dev<const double>(0), (int64_t)x.strides[0], call.dev<const double>(1), (int64_t)x_copy.strides[0], call.dev<const double>(2), (int64_t)y.strides[0], call.dev<const double>(3), (int64_t)y_copy.strides[0], call.dev<double>(4), (int64_t)A.strides[0]); });
  return call.finish();
}

// src/level2/trmv.py:9-43
extern "C" void exo_strmv_rm_stride_any(void *ctxt, int_fast32_t Uplo, int_fast32_t TransA, int_fast32_t Diag, int_fast32_t n, struct exo_win_1f32 x, struct exo_win_2f32c A) {
  exoblas::Call call("exo_strmv_rm_stride_any", exoblas::Call::kSync, ctxt);
  if (!(Uplo == 121 || Uplo == 122)) return (void)call.fail_enum("Uplo", Uplo);
  if (!(TransA == 111 || TransA == 112 || TransA == 113)) return (void)call.fail_enum("TransA", TransA);
  if (!(Diag == 131 || Diag == 132)) return (void)call.fail_enum("Diag", Diag);
  if (n < 0) return (void)call.fail_size("n", n);
  if ((true) && A.strides[1] != 1) return (void)call.fail_stride("A", A.strides[1]);
  if ((true) && n > 1 && A.strides[0] < n) return (void)call.fail_ld("A", A.strides[0]);
  call.window(0, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(float), true, true);
  call.window(1, A.data, 0, exoblas::span2_hi(n, n, A.strides[0]), sizeof(float), false, true);
  if (!call.stage()) return (void)call.status();
  if (!call.ok()) return (void)call.status();
  call.run([&] { exoblas::ops::trmv_rm<float>(call.ctx(), (int64_t)Uplo, (int64_t)TransA, (int64_t)Diag, (int64_t)n, call.dev<float>(0), (int64_t)x.strides[0], call.dev<const float>(1), (int64_t)A.strides[0]); });
  return (void)call.finish();
}
extern "C" int exoblas_strmv_rm_stride_any_async(hipStream_t stream, int_fast32_t Uplo, int_fast32_t TransA, int_fast32_t Diag, int_fast32_t n, struct exo_win_1f32 x, struct exo_win_2f32c A) {
  exoblas::Call call("exoblas_strmv_rm_stride_any_async", exoblas::Call::kAsync, stream);
  if (!(Uplo == 121 || Uplo == 122)) return call.fail_enum("Uplo", Uplo);
  if (!(TransA == 111 || TransA == 112 || TransA == 113)) return call.fail_enum("TransA", TransA);
  if (!(Diag == 131 || Diag == 132)) return call.fail_enum("Diag", Diag);
  if (n < 0) return call.fail_size("n", n);
  if ((true) && A.strides[1] != 1) return call.fail_stride("A", A.strides[1]);
  if ((true) && n > 1 && A.strides[0] < n) return call.fail_ld("A", A.strides[0]);
  call.window(0, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(float), true, true);
  call.window(1, A.data, 0, exoblas::span2_hi(n, n, A.strides[0]), sizeof(float), false, true);
  if (!call.stage()) return call.status();
  if (!call.ok()) return call.status();
  call.run([&] { exoblas::ops::trmv_rm<float>(call.ctx(), (int64_t)Uplo, (int64_t)TransA, (int64_t)Diag, (int64_t)n, call.dev<float>(0), (int64_t)x.strides[0], call.dev<const float>(1), (int64_t)A.strides[0]); });
  return call.finish();
}

// src/level2/trmv.py:9-43
extern "C" void exo_strmv_rm_stride_1(void *ctxt, int_fast32_t Uplo, int_fast32_t TransA, int_fast32_t Diag, int_fast32_t n, struct exo_win_1f32 x, struct exo_win_2f32c A) {
  exoblas::Call call("exo_strmv_rm_stride_1", exoblas::Call::kSync, ctxt);
  if (!(Uplo == 121 || Uplo == 122)) return (void)call.fail_enum("Uplo", Uplo);
  if (!(TransA == 111 || TransA == 112 || TransA == 113)) return (void)call.fail_enum("TransA", TransA);
  if (!(Diag == 131 || Diag == 132)) return (void)call.fail_enum("Diag", Diag);
  if (n < 0) return (void)call.fail_size("n", n);
  if ((true) && A.strides[1] != 1) return (void)call.fail_stride("A", A.strides[1]);
  if ((true) && n > 1 && A.strides[0] < n) return (void)call.fail_ld("A", A.strides[0]);
  call.window(0, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(float), true, true);
  call.window(1, A.data, 0, exoblas::span2_hi(n, n, A.strides[0]), sizeof(float), false, true);
  if (!call.stage()) return (void)call.status();
  if (!call.ok()) return (void)call.status();
  call.run([&] { exoblas::ops::trmv_rm<float>(call.ctx(), (int64_t)Uplo, (int64_t)TransA, (int64_t)Diag, (int64_t)n, call.dev<float>(0), (int64_t)x.strides[0], call.dev<const float>(1), (int64_t)A.strides[0]); });
  return (void)call.finish();
}
extern "C" int exoblas_strmv_rm_stride_1_async(hipStream_t stream, int_fast32_t Uplo, int_fast32_t TransA, int_fast32_t Diag, int_fast32_t n, struct exo_win_1f32 x, struct exo_win_2f32c A) {
  exoblas::Call call("exoblas_strmv_rm_stride_1_async", exoblas::Call::kAsync, stream);
  if (!(Uplo == 121 || Uplo == 122)) return call.fail_enum("Uplo", Uplo);
  if (!(TransA == 111 || TransA == 112 || TransA == 113)) return call.fail_enum("TransA", TransA);
  if (!(Diag == 131 || Diag == 132)) return call.fail_enum("Diag", Diag);
  if (n < 0) return call.fail_size("n", n);
  if ((true) && A.strides[1] != 1) return call.fail_stride("A", A.strides[1]);
  if ((true) && n > 1 && A.strides[0] < n) return call.fail_ld("A", A.strides[0]);
  call.window(0, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(float), true, true);
  call.window(1, A.data, 0, exoblas::span2_hi(n, n, A.strides[0]), sizeof(float), false, true);
  if (!call.stage()) return call.status();
  if (!call.ok()) return call.status();
  call.run([&] { exoblas::ops::trmv_rm<float>(call.ctx(), (int64_t)Uplo, (int64_t)TransA, (int64_t)Diag, (int64_t)n, call.dev<float>(0), (int64_t)x.strides[0], call.dev<const float>(1), (int64_t)A.strides[0]); });
  return call.finish();
}

// src/level2/trmv.py:9-43
extern "C" void exo_dtrmv_rm_stride_any(void *ctxt, int_fast32_t Uplo, int_fast32_t TransA, int_fast32_t Diag, int_fast32_t n, struct exo_win_1f64 x, struct exo_win_2f64c A) {
  exoblas::Call call("exo_dtrmv_rm_stride_any", exoblas::Call::kSync, ctxt);
  if (!(Uplo == 121 || Uplo == 122)) return (void)call.fail_enum("Uplo", Uplo);
  if (!(TransA == 111 || TransA == 112 || TransA == 113)) return (void)call.fail_enum("TransA", TransA);
  if (!(Diag == 131 || Diag == 132)) return (void)call.fail_enum("Diag", Diag);
  if (n < 0) return (void)call.fail_size("n", n);
  if ((true) && A.strides[1] != 1) return (void)call.fail_stride("A", A.strides[1]);
  if ((true) && n > 1 && A.strides[0] < n) return (void)call.fail_ld("A", A.strides[0]);
  call.window(0, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(double), true, true);
  call.window(1, A.data, 0, exoblas::span2_hi(n, n, A.strides[0]), sizeof(double), false, true);
  if (!call.stage()) return (void)call.status();
  if (!call.ok()) return (void)call.status();
  call.run([&] { exoblas::ops::trmv_rm<double>(call.ctx(), (int64_t)Uplo, (int64_t)TransA, (int64_t)Diag, (int64_t)n, call.dev<double>(0), (int64_t)x.strides[0], call.dev<const double>(1), (int64_t)A.strides[0]); });
  return (void)call.finish();
}
extern "C" int exoblas_dtrmv_rm_stride_any_async(hipStream_t stream, int_fast32_t Uplo, int_fast32_t TransA, int_fast32_t Diag, int_fast32_t n, struct exo_win_1f64 x, struct exo_win_2f64c A) {
  exoblas::Call call("exoblas_dtrmv_rm_stride_any_async", exoblas::Call::kAsync, stream);
  if (!(Uplo == 121 || Uplo == 122)) return call.fail_enum("Uplo", Uplo);
  if (!(TransA == 111 || TransA == 112 || TransA == 113)) return call.fail_enum("TransA", TransA);
  if (!(Diag == 131 || Diag == 132)) return call.fail_enum("Diag", Diag);
  if (n < 0) return call.fail_size("n", n);
  if ((true) && A.strides[1] != 1) return call.fail_stride("A", A.strides[1]);
  if ((true) && n > 1 && A.strides[0] < n) return call.fail_ld("A", A.strides[0]);
  call.window(0, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(double), true, true);
  call.window(1, A.data, 0, exoblas::span2_hi(n, n, A.strides[0]), sizeof(double), false, true);
  if (!call.stage()) return call.status();
  if (!call.ok()) return call.status();
  call.run([&] { exoblas::ops::trmv_rm<double>(call.ctx(), (int64_t)Uplo, (int64_t)TransA, (int64_t)Diag, (int64_t)n, call.dev<double>(0), (int64_t)x.strides[0], call.dev<const double>(1), (int64_t)A.strides[0]); });
  return call.finish();
}

// src/level2/trmv.py:9-43
extern "C" void exo_dtrmv_rm_stride_1(void *ctxt, int_fast32_t Uplo, int_fast32_t TransA, int_fast32_t Diag, int_fast32_t n, struct exo_win_1f64 x, struct exo_win_2f64c A) {
  exoblas::Call call("exo_dtrmv_rm_stride_1", exoblas::Call::kSync, ctxt);
  if (!(Uplo == 121 || Uplo == 122)) return (void)call.fail_enum("Uplo", Uplo);
  if (!(TransA == 111 || TransA == 112 || TransA == 113)) return (void)call.fail_enum("TransA", TransA);
  if (!(Diag == 131 || Diag == 132)) return (void)call.fail_enum("Diag", Diag);
  if (n < 0) return (void)call.fail_size("n", n);
  if ((true) && A.strides[1] != 1) return (void)call.fail_stride("A", A.strides[1]);
  if ((true) && n > 1 && A.strides[0] < n) return (void)call.fail_ld("A", A.strides[0]);
  call.window(0, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(double), true, true);
  call.window(1, A.data, 0, exoblas::span2_hi(n, n, A.strides[0]), sizeof(double), false, true);
  if (!call.stage()) return (void)call.status();
  if (!call.ok()) return (void)call.status();
  call.run([&] { exoblas::ops::trmv_rm<double>(call.ctx(), (int64_t)Uplo, (int64_t)TransA, (int64_t)Diag, (int64_t)n, call.dev<double>(0), (int64_t)x.strides[0], call.dev<const double>(1), (int64_t)A.strides[0]); });
  return (void)call.finish();
}
extern "C" int exoblas_dtrmv_rm_stride_1_async(hipStream_t stream, int_fast32_t Uplo, int_fast32_t TransA, int_fast32_t Diag, int_fast32_t n, struct exo_win_1f64 x, struct exo_win_2f64c A) {
  exoblas::Call call("exoblas_dtrmv_rm_stride_1_async", exoblas::Call::kAsync, stream);
  if (!(Uplo == 121 || Uplo == 122)) return call.fail_enum("Uplo", Uplo);
  if (!(TransA == 111 || TransA == 112 || TransA == 113)) return call.fail_enum("TransA", TransA);
  if (!(Diag == 131 || Diag == 132)) return call.fail_enum("Diag", Diag);
  if (n < 0) return call.fail_size("n", n);
  if ((true) && A.strides[1] != 1) return call.fail_stride("A", A.strides[1]);
  if ((true) && n > 1 && A.strides[0] < n) return call.fail_ld("A", A.strides[0]);
  call.window(0, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(double), true, true);
  call.window(1, A.data, 0, exoblas::span2_hi(n, n, A.strides[0]), sizeof(double), false, true);
  if (!call.stage()) return call.status();
  if (!call.ok()) return call.status();
  call.run([&] { exoblas::ops::trmv_rm<double>(call.ctx(), (int64_t)Uplo, (int64_t)TransA, (int64_t)Diag, (int64_t)n, call.dev<double>(0), (int64_t)x.strides[0], call.dev<const double>(1), (int64_t)A.strides[0]); });
  return call.finish();
}

// src/level2/trsv.py:10-52
extern "C" void exo_strsv_rm_stride_any(void *ctxt, int_fast32_t Uplo, int_fast32_t TransA, int_fast32_t Diag, int_fast32_t n, struct exo_win_1f32 x, struct exo_win_2f32c A) {
  exoblas::Call call("exo_strsv_rm_stride_any", exoblas::Call::kSync, ctxt);
  if (!(Uplo == 121 || Uplo == 122)) return (void)call.fail_enum("Uplo", Uplo);
  if (!(TransA == 111 || TransA == 112 || TransA == 113)) return (void)call.fail_enum("TransA", TransA);
  if (!(Diag == 131 || Diag == 132)) return (void)call.fail_enum("Diag", Diag);
  if (n < 0) return (void)call.fail_size("n", n);
  if ((true) && A.strides[1] != 1) return (void)call.fail_stride("A", A.strides[1]);
  if ((true) && n > 1 && A.strides[0] < n) return (void)call.fail_ld("A", A.strides[0]);
  call.window(0, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(float), true, true);
  call.window(1, A.data, 0, exoblas::span2_hi(n, n, A.strides[0]), sizeof(float), false, true);
  if (!call.stage()) return (void)call.status();
  if (!call.ok()) return (void)call.status();
  call.run([&] { exoblas::ops::trsv_rm<float>(call.ctx(), (int64_t)Uplo, (int64_t)TransA, (int64_t)Diag, (int64_t)n, call.dev<float>(0), (int64_t)x.strides[0], call.dev<const float>(1), (int64_t)A.strides[0]); });
  return (void)call.finish();
}
extern "C" int exoblas_strsv_rm_stride_any_async(hipStream_t stream, int_fast32_t Uplo, int_fast32_t TransA, int_fast32_t Diag, int_fast32_t n, struct exo_win_1f32 x, struct exo_win_2f32c A) {
  exoblas::Call call("exoblas_strsv_rm_stride_any_async", exoblas::Call::kAsync, stream);
  if (!(Uplo == 121 || Uplo == 122)) return call.fail_enum("Uplo", Uplo);
  if (!(TransA == 111 || TransA == 112 || TransA == 113)) return call.fail_enum("TransA", TransA);
  if (!(Diag == 131 || Diag == 132)) return call.fail_enum("Diag", Diag);
  if (n < 0) return call.fail_size("n", n);
  if ((true) && A.strides[1] != 1) return call.fail_stride("A", A.strides[1]);
  if ((true) && n > 1 && A.strides[0] < n) return call.fail_ld("A", A.strides[0]);
  call.window(0, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(float), true, true);
  call.window(1, A.data, 0, exoblas::span2_hi(n, n, A.strides[0]), sizeof(float), false, true);
  if (!call.stage()) return call.status();
  if (!call.ok()) return call.status();
  call.run([&] { exoblas::ops::trsv_rm<float>(call.ctx(), (int64_t)Uplo, (int64_t)TransA, (int64_t)Diag, (int64_t)n, call.dev<float>(0), (int64_t)x.strides[0], call.dev<const float>(1), (int64_t)A.strides[0]); });
  return call.finish();
}

// src/level2/trsv.py:10-52
extern "C" void exo_strsv_rm_stride_1(void *ctxt, int_fast32_t Uplo, int_fast32_t TransA, int_fast32_t Diag, int_fast32_t n, struct exo_win_1f32 x, struct exo_win_2f32c A) {
  exoblas::Call call("exo_strsv_rm_stride_1", exoblas::Call::kSync, ctxt);
  if (!(Uplo == 121 || Uplo == 122)) return (void)call.fail_enum("Uplo", Uplo);
  if (!(TransA == 111 || TransA == 112 || TransA == 113)) return (void)call.fail_enum("TransA", TransA);
  if (!(Diag == 131 || Diag == 132)) return (void)call.fail_enum("Diag", Diag);
  if (n < 0) return (void)call.fail_size("n", n);
  if ((true) && A.strides[1] != 1) return (void)call.fail_stride("A", A.strides[1]);
  if ((true) && n > 1 && A.strides[0] < n) return (void)call.fail_ld("A", A.strides[0]);
  call.window(0, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(float), true, true);
  call.window(1, A.data, 0, exoblas::span2_hi(n, n, A.strides[0]), sizeof(float), false, true);
  if (!call.stage()) return (void)call.status();
  if (!call.ok()) return (void)call.status();
  call.run([&] { exoblas::ops::trsv_rm<float>(call.ctx(), (int64_t)Uplo, (int64_t)TransA, (int64_t)Diag, (int64_t)n, call.dev<float>(0), (int64_t)x.strides[0], call.dev<const float>(1), (int64_t)A.strides[0]); });
  return (void)call.finish();
}
extern "C" int exoblas_strsv_rm_stride_1_async(hipStream_t stream, int_fast32_t Uplo, int_fast32_t TransA, int_fast32_t Diag, int_fast32_t n, struct exo_win_1f32 x, struct exo_win_2f32c A) {
  exoblas::Call call("exoblas_strsv_rm_stride_1_async", exoblas::Call::kAsync, stream);
  if (!(Uplo == 121 || Uplo == 122)) return call.fail_enum("Uplo", Uplo);
  if (!(TransA == 111 || TransA == 112 || TransA == 113)) return call.fail_enum("TransA", TransA);
  if (!(Diag == 131 || Diag == 132)) return call.fail_enum("Diag", Diag);
  if (n < 0) return call.fail_size("n", n);
  if ((true) && A.strides[1] != 1) return call.fail_stride("A", A.strides[1]);
  if ((true) && n > 1 && A.strides[0] < n) return call.fail_ld("A", A.strides[0]);
  call.window(0, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(float), true, true);
  call.window(1, A.data, 0, exoblas::span2_hi(n, n, A.strides[0]), sizeof(float), false, true);
  if (!call.stage()) return call.status();
  if (!call.ok()) return call.status();
  call.run([&] { exoblas::ops::trsv_rm<float>(call.ctx(), (int64_t)Uplo, (int64_t)TransA, (int64_t)Diag, (int64_t)n, call.dev<float>(0), (int64_t)x.strides[0], call.dev<const float>(1), (int64_t)A.strides[0]); });
  return call.finish();
}

// src/level2/trsv.py:10-52
extern "C" void exo_dtrsv_rm_stride_any(void *ctxt, int_fast32_t Uplo, int_fast32_t TransA, int_fast32_t Diag, int_fast32_t n, struct exo_win_1f64 x, struct exo_win_2f64c A) {
  exoblas::Call call("exo_dtrsv_rm_stride_any", exoblas::Call::kSync, ctxt);
  if (!(Uplo == 121 || Uplo == 122)) return (void)call.fail_enum("Uplo", Uplo);
  if (!(TransA == 111 || TransA == 112 || TransA == 113)) return (void)call.fail_enum("TransA", TransA);
  if (!(Diag == 131 || Diag == 132)) return (void)call.fail_enum("Diag", Diag);
  if (n < 0) return (void)call.fail_size("n", n);
  if ((true) && A.strides[1] != 1) return (void)call.fail_stride("A", A.strides[1]);
  if ((true) && n > 1 && A.strides[0] < n) return (void)call.fail_ld("A", A.strides[0]);
  call.window(0, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(double), true, true);
  call.window(1, A.data, 0, exoblas::span2_hi(n, n, A.strides[0]), sizeof(double), false, true);
  if (!call.stage()) return (void)call.status();
  if (!call.ok()) return (void)call.status();
  call.run([&] { exoblas::ops::trsv_rm<double>(call.ctx(), (int64_t)Uplo, (int64_t)TransA, (int64_t)Diag, (int64_t)n, call.dev<double>(0), (int64_t)x.strides[0], call.dev<const double>(1), (int64_t)A.strides[0]); });
  return (void)call.finish();
}
extern "C" int exoblas_dtrsv_rm_stride_any_async(hipStream_t stream, int_fast32_t Uplo, int_fast32_t TransA, int_fast32_t Diag, int_fast32_t n, struct exo_win_1f64 x, struct exo_win_2f64c A) {
  exoblas::Call call("exoblas_dtrsv_rm_stride_any_async", exoblas::Call::kAsync, stream);
  if (!(Uplo == 121 || Uplo == 122)) return call.fail_enum("Uplo", Uplo);
  if (!(TransA == 111 || TransA == 112 || TransA == 113)) return call.fail_enum("TransA", TransA);
  if (!(Diag == 131 || Diag == 132)) return call.fail_enum("Diag", Diag);
  if (n < 0) return call.fail_size("n", n);
  if ((true) && A.strides[1] != 1) return call.fail_stride("A", A.strides[1]);
  if ((true) && n > 1 && A.strides[0] < n) return call.fail_ld("A", A.strides[0]);
  call.window(0, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(double), true, true);
  call.window(1, A.data, 0, exoblas::span2_hi(n, n, A.strides[0]), sizeof(double), false, true);
  if (!call.stage()) return call.status();
  if (!call.ok()) return call.status();
  call.run([&] { exoblas::ops::trsv_rm<double>(call.ctx(), (int64_t)Uplo, (int64_t)TransA, (int64_t)Diag, (int64_t)n, call.dev<double>(0), (int64_t)x.strides[0], call.dev<const double>(1), (int64_t)A.strides[0]); });
  return call.finish();
}

// src/level2/trsv.py:10-52
extern "C" void exo_dtrsv_rm_stride_1(void *ctxt, int_fast32_t Uplo, int_fast32_t TransA, int_fast32_t Diag, int_fast32_t n, struct exo_win_1f64 x, struct exo_win_2f64c A) {
  exoblas::Call call("exo_dtrsv_rm_stride_1", exoblas::Call::kSync, ctxt);
  if (!(Uplo == 121 || Uplo == 122)) return (void)call.fail_enum("Uplo", Uplo);
  if (!(TransA == 111 || TransA == 112 || TransA == 113)) return (void)call.fail_enum("TransA", TransA);
  if (!(Diag == 131 || Diag == 132)) return (void)call.fail_enum("Diag", Diag);
  if (n < 0) return (void)call.fail_size("n", n);
  if ((true) && A.strides[1] != 1) return (void)call.fail_stride("A", A.strides[1]);
  if ((true) && n > 1 && A.strides[0] < n) return (void)call.fail_ld("A", A.strides[0]);
  call.window(0, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(double), true, true);
  call.window(1, A.data, 0, exoblas::span2_hi(n, n, A.strides[0]), sizeof(double), false, true);
  if (!call.stage()) return (void)call.status();
  if (!call.ok()) return (void)call.status();
  call.run([&] { exoblas::ops::trsv_rm<double>(call.ctx(), (int64_t)Uplo, (int64_t)TransA, (int64_t)Diag, (int64_t)n, call.dev<double>(0), (int64_t)x.strides[0], call.dev<const double>(1), (int64_t)A.strides[0]); });
  return (void)call.finish();
}
extern "C" int exoblas_dtrsv_rm_stride_1_async(hipStream_t stream, int_fast32_t Uplo, int_fast32_t TransA, int_fast32_t Diag, int_fast32_t n, struct exo_win_1f64 x, struct exo_win_2f64c A) {
  exoblas::Call call("exoblas_dtrsv_rm_stride_1_async", exoblas::Call::kAsync, stream);
  if (!(Uplo == 121 || Uplo == 122)) return call.fail_enum("Uplo", Uplo);
  if (!(TransA == 111 || TransA == 112 || TransA == 113)) return call.fail_enum("TransA", TransA);
  if (!(Diag == 131 || Diag == 132)) return call.fail_enum("Diag", Diag);
  if (n < 0) return call.fail_size("n", n);
  if ((true) && A.strides[1] != 1) return call.fail_stride("A", A.strides[1]);
  if ((true) && n > 1 && A.strides[0] < n) return call.fail_ld("A", A.strides[0]);
  call.window(0, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(double), true, true);
  call.window(1, A.data, 0, exoblas::span2_hi(n, n, A.strides[0]), sizeof(double), false, true);
  if (!call.stage()) return call.status();
  if (!call.ok()) return call.status();
  call.run([&] { exoblas::ops::trsv_rm<double>(call.ctx(), (int64_t)Uplo, (int64_t)TransA, (int64_t)Diag, (int64_t)n, call.dev<double>(0), (int64_t)x.strides[0], call.dev<const double>(1), (int64_t)A.strides[0]); });
  return call.finish();
}

// src/level2/gbmv.py:13-52
extern "C" void exo_sgbmv_row_major_NonTrans_stride_any(void *ctxt, int_fast32_t m, int_fast32_t n, int_fast32_t kl, int_fast32_t ku, const float *alpha, const float *beta, struct exo_win_2f32c a, struct exo_win_1f32c x, struct exo_win_1f32 y) {
  exoblas::Call call("exo_sgbmv_row_major_NonTrans_stride_any", exoblas::Call::kSync, ctxt);
  if (m < 0) return (void)call.fail_size("m", m);
  if (n < 0) return (void)call.fail_size("n", n);
  if (kl < 0) return (void)call.fail_size("kl", kl);
  if (ku < 0) return (void)call.fail_size("ku", ku);
  if ((true) && a.strides[1] != 1) return (void)call.fail_stride("a", a.strides[1]);
  if ((true) && m > 1 && a.strides[0] < kl + ku + 1) return (void)call.fail_ld("a", a.strides[0]);
  call.window(0, a.data, 0, exoblas::span2_hi(m, kl + ku + 1, a.strides[0]), sizeof(float), false, true);
  call.window(1, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(float), false, true);
  call.window(2, y.data, exoblas::span_lo(m, y.strides[0]), exoblas::span_hi(m, y.strides[0]), sizeof(float), true, true);
  if (!call.stage()) return (void)call.status();
  const float alpha_v = call.scalar(alpha);
  const float beta_v = call.scalar(beta);
  if (!call.ok()) return (void)call.status();
  call.run([&] { exoblas::ops::gbmv_row_major_NonTrans<float>(call.ctx(), (int64_t)m, (int64_t)n, (int64_t)kl, (int64_t)ku, alpha_v, beta_v, call.dev<const float>(0), (int64_t)a.strides[0], call.dev<const float>(1), (int64_t)x.strides[0], call.dev<float>(2), (int64_t)y.strides[0]); });
  return (void)call.finish();
}
extern "C" int exoblas_sgbmv_row_major_NonTrans_stride_any_async(hipStream_t stream, int_fast32_t m, int_fast32_t n, int_fast32_t kl, int_fast32_t ku, const float *alpha, const float *beta, struct exo_win_2f32c a, struct exo_win_1f32c x, struct exo_win_1f32 y) {
  exoblas::Call call("exoblas_sgbmv_row_major_NonTrans_stride_any_async", exoblas::Call::kAsync, stream);
  if (m < 0) return call.fail_size("m", m);
  if (n < 0) return call.fail_size("n", n);
  if (kl < 0) return call.fail_size("kl", kl);
  if (ku < 0) return call.fail_size("ku", ku);
  if ((true) && a.strides[1] != 1) return call.fail_stride("a", a.strides[1]);
  if ((true) && m > 1 && a.strides[0] < kl + ku + 1) return call.fail_ld("a", a.strides[0]);
  call.window(0, a.data, 0, exoblas::span2_hi(m, kl + ku + 1, a.strides[0]), sizeof(float), false, true);
  call.window(1, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(float), false, true);
  call.window(2, y.data, exoblas::span_lo(m, y.strides[0]), exoblas::span_hi(m, y.strides[0]), sizeof(float), true, true);
  if (!call.stage()) return call.status();
  const float alpha_v = call.scalar(alpha);
  const float beta_v = call.scalar(beta);
  if (!call.ok()) return call.status();
  call.run([&] { exoblas::ops::gbmv_row_major_NonTrans<float>(call.ctx(), (int64_t)m, (int64_t)n, (int64_t)kl, (int64_t)ku, alpha_v, beta_v, call.dev<const float>(0), (int64_t)a.strides[0], call.dev<const float>(1), (int64_t)x.strides[0], call.dev<float>(2), (int64_t)y.strides[0]); });
  return call.finish();
}

// src/level2/gbmv.py:13-52
extern "C" void exo_sgbmv_row_major_NonTrans_stride_1(void *ctxt, int_fast32_t m, int_fast32_t n, int_fast32_t kl, int_fast32_t ku, const float *alpha, const float *beta, struct exo_win_2f32c a, struct exo_win_1f32c x, struct exo_win_1f32 y) {
  exoblas::Call call("exo_sgbmv_row_major_NonTrans_stride_1", exoblas::Call::kSync, ctxt);
  if (m < 0) return (void)call.fail_size("m", m);
  if (n < 0) return (void)call.fail_size("n", n);
  if (kl < 0) return (void)call.fail_size("kl", kl);
  if (ku < 0) return (void)call.fail_size("ku", ku);
  if ((true) && a.strides[1] != 1) return (void)call.fail_stride("a", a.strides[1]);
  if ((true) && m > 1 && a.strides[0] < kl + ku + 1) return (void)call.fail_ld("a", a.strides[0]);
  call.window(0, a.data, 0, exoblas::span2_hi(m, kl + ku + 1, a.strides[0]), sizeof(float), false, true);
  call.window(1, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(float), false, true);
  call.window(2, y.data, exoblas::span_lo(m, y.strides[0]), exoblas::span_hi(m, y.strides[0]), sizeof(float), true, true);
  if (!call.stage()) return (void)call.status();
  const float alpha_v = call.scalar(alpha);
  const float beta_v = call.scalar(beta);
  if (!call.ok()) return (void)call.status();
  call.run([&] { exoblas::ops::gbmv_row_major_NonTrans<float>(call.ctx(), (int64_t)m, (int64_t)n, (int64_t)kl, (int64_t)ku, alpha_v, beta_v, call.dev<const float>(0), (int64_t)a.strides[0], call.dev<const float>(1), (int64_t)x.strides[0], call.dev<float>(2), (int64_t)y.strides[0]); });
  return (void)call.finish();
}
extern "C" int exoblas_sgbmv_row_major_NonTrans_stride_1_async(hipStream_t stream, int_fast32_t m, int_fast32_t n, int_fast32_t kl, int_fast32_t ku, const float *alpha, const float *beta, struct exo_win_2f32c a, struct exo_win_1f32c x, struct exo_win_1f32 y) {
  exoblas::Call call("exoblas_sgbmv_row_major_NonTrans_stride_1_async", exoblas::Call::kAsync, stream);
  if (m < 0) return call.fail_size("m", m);
  if (n < 0) return call.fail_size("n", n);
  if (kl < 0) return call.fail_size("kl", kl);
  if (ku < 0) return call.fail_size("ku", ku);
  if ((true) && a.strides[1] != 1) return call.fail_stride("a", a.strides[1]);
  if ((true) && m > 1 && a.strides[0] < kl + ku + 1) return call.fail_ld("a", a.strides[0]);
  call.window(0, a.data, 0, exoblas::span2_hi(m, kl + ku + 1, a.strides[0]), sizeof(float), false, true);
  call.window(1, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(float), false, true);
  call.window(2, y.data, exoblas::span_lo(m, y.strides[0]), exoblas::span_hi(m, y.strides[0]), sizeof(float), true, true);
  if (!call.stage()) return call.status();
  const float alpha_v = call.scalar(alpha);
  const float beta_v = call.scalar(beta);
  if (!call.ok()) return call.status();
  call.run([&] { exoblas::ops::gbmv_row_major_NonTrans<float>(call.ctx(), (int64_t)m, (int64_t)n, (int64_t)kl, (int64_t)ku, alpha_v, beta_v, call.dev<const float>(0), (int64_t)a.strides[0], call.dev<const float>(1), (int64_t)x.strides[0], call.dev<float>(2), (int64_t)y.strides[0]); });
  return call.finish();
}

// src/level2/gbmv.py:13-52
extern "C" void exo_dgbmv_row_major_NonTrans_stride_any(void *ctxt, int_fast32_t m, int_fast32_t n, int_fast32_t kl, int_fast32_t ku, const double *alpha, const double *beta, struct exo_win_2f64c a, struct exo_win_1f64c x, struct exo_win_1f64 y) {
  exoblas::Call call("exo_dgbmv_row_major_NonTrans_stride_any", exoblas::Call::kSync, ctxt);
  if (m < 0) return (void)call.fail_size("m", m);
  if (n < 0) return (void)call.fail_size("n", n);
  if (kl < 0) return (void)call.fail_size("kl", kl);
  if (ku < 0) return (void)call.fail_size("ku", ku);
  if ((true) && a.strides[1] != 1) return (void)call.fail_stride("a", a.strides[1]);
  if ((true) && m > 1 && a.strides[0] < kl + ku + 1) return (void)call.fail_ld("a", a.strides[0]);
  call.window(0, a.data, 0, exoblas::span2_hi(m, kl + ku + 1, a.strides[0]), sizeof(double), false, true);
  call.window(1, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(double), false, true);
  call.window(2, y.data, exoblas::span_lo(m, y.strides[0]), exoblas::span_hi(m, y.strides[0]), sizeof(double), true, true);
  if (!call.stage()) return (void)call.status();
  const double alpha_v = call.scalar(alpha);
  const double beta_v = call.scalar(beta);
  if (!call.ok()) return (void)call.status();
  call.run([&] { exoblas::ops::gbmv_row_major_NonTrans<double>(call.ctx(), (int64_t)m, (int64_t)n, (int64_t)kl, (int64_t)ku, alpha_v, beta_v, call.dev<const double>(0), (int64_t)a.strides[0], call.dev<const double>(1), (int64_t)x.strides[0], call.dev<double>(2), (int64_t)y.strides[0]); });
  return (void)call.finish();
}
extern "C" int exoblas_dgbmv_row_major_NonTrans_stride_any_async(hipStream_t stream, int_fast32_t m, int_fast32_t n, int_fast32_t kl, int_fast32_t ku, const double *alpha, const double *beta, struct exo_win_2f64c a, struct exo_win_1f64c x, struct exo_win_1f64 y) {
  exoblas::Call call("exoblas_dgbmv_row_major_NonTrans_stride_any_async", exoblas::Call::kAsync, stream);
  if (m < 0) return call.fail_size("m", m);
  if (n < 0) return call.fail_size("n", n);
  if (kl < 0) return call.fail_size("kl", kl);
  if (ku < 0) return call.fail_size("ku", ku);
  if ((true) && a.strides[1] != 1) return call.fail_stride("a", a.strides[1]);
  if ((true) && m > 1 && a.strides[0] < kl + ku + 1) return call.fail_ld("a", a.strides[0]);
  call.window(0, a.data, 0, exoblas::span2_hi(m, kl + ku + 1, a.strides[0]), sizeof(double), false, true);
  call.window(1, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(double), false, true);
  call.window(2, y.data, exoblas::span_lo(m, y.strides[0]), exoblas::span_hi(m, y.strides[0]), sizeof(double), true, true);
  if (!call.stage()) return call.status();
  const double alpha_v = call.scalar(alpha);
  const double beta_v = call.scalar(beta);
  if (!call.ok()) return call.status();
  call.run([&] { exoblas::ops::gbmv_row_major_NonTrans<double>(call.ctx(), (int64_t)m, (int64_t)n, (int64_t)kl, (int64_t)ku, alpha_v, beta_v, call.dev<const double>(0), (int64_t)a.strides[0], call.dev<const double>(1), (int64_t)x.strides[0], call.dev<double>(2), (int64_t)y.strides[0]); });
  return call.finish();
}

// src/level2/gbmv.py:13-52
extern "C" void exo_dgbmv_row_major_NonTrans_stride_1(void *ctxt, int_fast32_t m, int_fast32_t n, int_fast32_t kl, int_fast32_t ku, const double *alpha, const double *beta, struct exo_win_2f64c a, struct exo_win_1f64c x, struct exo_win_1f64 y) {
  exoblas::Call call("exo_dgbmv_row_major_NonTrans_stride_1", exoblas::Call::kSync, ctxt);
  if (m < 0) return (void)call.fail_size("m", m);
  if (n < 0) return (void)call.fail_size("n", n);
  if (kl < 0) return (void)call.fail_size("kl", kl);
  if (ku < 0) return (void)call.fail_size("ku", ku);
  if ((true) && a.strides[1] != 1) return (void)call.fail_stride("a", a.strides[1]);
  if ((true) && m > 1 && a.strides[0] < kl + ku + 1) return (void)call.fail_ld("a", a.strides[0]);
  call.window(0, a.data, 0, exoblas::span2_hi(m, kl + ku + 1, a.strides[0]), sizeof(double), false, true);
  call.window(1, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(double), false, true);
  call.window(2, y.data, exoblas::span_lo(m, y.strides[0]), exoblas::span_hi(m, y.strides[0]), sizeof(double), true, true);
  if (!call.stage()) return (void)call.status();
  const double alpha_v = call.scalar(alpha);
  const double beta_v = call.scalar(beta);
  if (!call.ok()) return (void)call.status();
  call.run([&] { exoblas::ops::gbmv_row_major_NonTrans<double>(call.ctx(), (int64_t)m, (int64_t)n, (int64_t)kl, (int64_t)ku, alpha_v, beta_v, call.dev<const double>(0), (int64_t)a.strides[0], call.dev<const double>(1), (int64_t)x.strides[0], call.dev<double>(2), (int64_t)y.strides[0]); });
  return (void)call.finish();
}
extern "C" int exoblas_dgbmv_row_major_NonTrans_stride_1_async(hipStream_t stream, int_fast32_t m, int_fast32_t n, int_fast32_t kl, int_fast32_t ku, const double *alpha, const double *beta, struct exo_win_2f64c a, struct exo_win_1f64c x, struct exo_win_1f64 y) {
  exoblas::Call call("exoblas_dgbmv_row_major_NonTrans_stride_1_async", exoblas::Call::kAsync, stream);
  if (m < 0) return call.fail_size("m", m);
  if (n < 0) return call.fail_size("n", n);
  if (kl < 0) return call.fail_size("kl", kl);
  if (ku < 0) return call.fail_size("ku", ku);
  if ((true) && a.strides[1] != 1) return call.fail_stride("a", a.strides[1]);
  if ((true) && m > 1 && a.strides[0] < kl + ku + 1) return call.fail_ld("a", a.strides[0]);
  call.window(0, a.data, 0, exoblas::span2_hi(m, kl + ku + 1, a.strides[0]), sizeof(double), false, true);
  call.window(1, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(double), false, true);
  call.window(2, y.data, exoblas::span_lo(m, y.strides[0]), exoblas::span_hi(m, y.strides[0]), sizeof(double), true, true);
  if (!call.stage()) return call.status();
  const double alpha_v = call.scalar(alpha);
  const double beta_v = call.scalar(beta);
  if (!call.ok()) return call.status();
  call.run([&] { exoblas::ops::gbmv_row_major_NonTrans<double>(call.ctx(), (int64_t)m, (int64_t)n, (int64_t)kl, (int64_t)ku, alpha_v, beta_v, call.dev<const double>(0), (int64_t)a.strides[0], call.dev<const double>(1), (int64_t)x.strides[0], call.dev<double>(2), (int64_t)y.strides[0]); });
  return call.finish();
}

// src/level2/sbmv.py:12-16
extern "C" void exo_ssbmv_scal_y_stride_any(void *ctxt, int_fast32_t n, const float *beta, struct exo_win_1f32 y) {
  exoblas::Call call("exo_ssbmv_scal_y_stride_any", exoblas::Call::kSync, ctxt);
  if (n < 0) return (void)call.fail_size("n", n);
  call.window(0, y.data, exoblas::span_lo(n, y.strides[0]), exoblas::span_hi(n, y.strides[0]), sizeof(float), true, true);
  if (!call.stage()) return (void)call.status();
  const float beta_v = call.scalar(beta);
  if (!call.ok()) return (void)call.status();
  call.run([&] { exoblas::ops::sbmv_scal_y<float>(call.ctx(), (int64_t)n, beta_v, call.dev<float>(0), (int64_t)y.strides[0]); });
  return (void)call.finish();
}
extern "C" int exoblas_ssbmv_scal_y_stride_any_async(hipStream_t stream, int_fast32_t n, const float *beta, struct exo_win_1f32 y) {
  exoblas::Call call("exoblas_ssbmv_scal_y_stride_any_async", exoblas::Call::kAsync, stream);
  if (n < 0) return call.fail_size("n", n);
  call.window(0, y.data, exoblas::span_lo(n, y.strides[0]), exoblas::span_hi(n, y.strides[0]), sizeof(float), true, true);
  if (!call.stage()) return call.status();
  const float beta_v = call.scalar(beta);
  if (!call.ok()) return call.status();
  call.run([&] { exoblas::ops::sbmv_scal_y<float>(call.ctx(), (int64_t)n, beta_v, call.dev<float>(0), (int64_t)y.strides[0]); });
  return call.finish();
}

// src/level2/sbmv.py:12-16
extern "C" void exo_ssbmv_scal_y_stride_1(void *ctxt, int_fast32_t n, const float *beta, struct exo_win_1f32 y) {
  exoblas::Call call("exo_ssbmv_scal_y_stride_1", exoblas::Call::kSync, ctxt);
  if (n < 0) return (void)call.fail_size("n", n);
  call.window(0, y.data, exoblas::span_lo(n, y.strides[0]), exoblas::span_hi(n, y.strides[0]), sizeof(float), true, true);
  if (!call.stage()) return (void)call.status();
  const float beta_v = call.scalar(beta);
  if (!call.ok()) return (void)call.status();
  call.run([&] { exoblas::ops::sbmv_scal_y<float>(call.ctx(), (int64_t)n, beta_v, call.dev<float>(0), (int64_t)y.strides[0]); });
  return (void)call.finish();
}
extern "C" int exoblas_ssbmv_scal_y_stride_1_async(hipStream_t stream, int_fast32_t n, const float *beta, struct exo_win_1f32 y) {
  exoblas::Call call("exoblas_ssbmv_scal_y_stride_1_async", exoblas::Call::kAsync, stream);
  if (n < 0) return call.fail_size("n", n);
  call.window(0, y.data, exoblas::span_lo(n, y.strides[0]), exoblas::span_hi(n, y.strides[0]), sizeof(float), true, true);
  if (!call.stage()) return call.status();
  const float beta_v = call.scalar(beta);
  if (!call.ok()) return call.status();
  call.run([&] { exoblas::ops::sbmv_scal_y<float>(call.ctx(), (int64_t)n, beta_v, call.dev<float>(0), (int64_t)y.strides[0]); });
  return call.finish();
}

// src/level2/sbmv.py:12-16
extern "C" void exo_dsbmv_scal_y_stride_any(void *ctxt, int_fast32_t n, const double *beta, struct exo_win_1f64 y) {
  exoblas::Call call("exo_dsbmv_scal_y_stride_any", exoblas::Call::kSync, ctxt);
  if (n < 0) return (void)call.fail_size("n", n);
  call.window(0, y.data, exoblas::span_lo(n, y.strides[0]), exoblas::span_hi(n, y.strides[0]), sizeof(double), true, true);
  if (!call.stage()) return (void)call.status();
  const double beta_v = call.scalar(beta);
  if (!call.ok()) return (void)call.status();
  call.run([&] { exoblas::ops::sbmv_scal_y<double>(call.ctx(), (int64_t)n, beta_v, call.dev<double>(0), (int64_t)y.strides[0]); });
  return (void)call.finish();
}
extern "C" int exoblas_dsbmv_scal_y_stride_any_async(hipStream_t stream, int_fast32_t n, const double *beta, struct exo_win_1f64 y) {
  exoblas::Call call("exoblas_dsbmv_scal_y_stride_any_async", exoblas::Call::kAsync, stream);
  if (n < 0) return call.fail_size("n", n);
  call.window(0, y.data, exoblas::span_lo(n, y.strides[0]), exoblas::span_hi(n, y.strides[0]), sizeof(double), true, true);
  if (!call.stage()) return call.status();
  const double beta_v = call.scalar(beta);
  if (!call.ok()) return call.status();
  call.run([&] { exoblas::ops::sbmv_scal_y<double>(call.ctx(), (int64_t)n, beta_v, call.dev<double>(0), (int64_t)y.strides[0]); });
  return call.finish();
}

// src/level2/sbmv.py:12-16
extern "C" void exo_dsbmv_scal_y_stride_1(void *ctxt, int_fast32_t n, const double *beta, struct exo_win_1f64 y) {
  exoblas::Call call("exo_dsbmv_scal_y_stride_1", exoblas::Call::kSync, ctxt);
  if (n < 0) return (void)call.fail_size("n", n);
  call.window(0, y.data, exoblas::span_lo(n, y.strides[0]), exoblas::span_hi(n, y.strides[0]), sizeof(double), true, true);
  if (!call.stage()) return (void)call.status();
  const double beta_v = call.scalar(beta);
  if (!call.ok()) return (void)call.status();
  call.run([&] { exoblas::ops::sbmv_scal_y<double>(call.ctx(), (int64_t)n, beta_v, call.dev<double>(0), (int64_t)y.strides[0]); });
  return (void)call.finish();
}
extern "C" int exoblas_dsbmv_scal_y_stride_1_async(hipStream_t stream, int_fast32_t n, const double *beta, struct exo_win_1f64 y) {
  exoblas::Call call("exoblas_dsbmv_scal_y_stride_1_async", exoblas::Call::kAsync, stream);
  if (n < 0) return call.fail_size("n", n);
  call.window(0, y.data, exoblas::span_lo(n, y.strides[0]), exoblas::span_hi(n, y.strides[0]), sizeof(double), true, true);
  if (!call.stage()) return call.status();
  const double beta_v = call.scalar(beta);
  if (!call.ok()) return call.status();
  call.run([&] { exoblas::ops::sbmv_scal_y<double>(call.ctx(), (int64_t)n, beta_v, call.dev<double>(0), (int64_t)y.strides[0]); });
  return call.finish();
}

// src/level2/sbmv.py:19-33
extern "C" void exo_ssbmv_row_major_Upper_stride_any(void *ctxt, int_fast32_t n, int_fast32_t k, const float *alpha, struct exo_win_2f32c A, struct exo_win_1f32c x, struct exo_win_1f32 y) {
  exoblas::Call call("exo_ssbmv_row_major_Upper_stride_any", exoblas::Call::kSync, ctxt);
  if (n < 0) return (void)call.fail_size("n", n);
  if (k < 0) return (void)call.fail_size("k", k);
  if ((true) && A.strides[1] != 1) return (void)call.fail_stride("A", A.strides[1]);
  if ((true) && n > 1 && A.strides[0] < k + 1) return (void)call.fail_ld("A", A.strides[0]);
  call.window(0, A.data, 0, exoblas::span2_hi(n, k + 1, A.strides[0]), sizeof(float), false, true);
  call.window(1, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(float), false, true);
  call.window(2, y.data, exoblas::span_lo(n, y.strides[0]), exoblas::span_hi(n, y.strides[0]), sizeof(float), true, true);
  if (!call.stage()) return (void)call.status();
  const float alpha_v = call.scalar(alpha);
  if (!call.ok()) return (void)call.status();
  call.run([&] { exoblas::ops::sbmv_row_major_Upper<float>(call.ctx(), (int64_t)n, (int64_t)k, alpha_v, call.dev<const float>(0), (int64_t)A.strides[0], call.dev<const float>(1), (int64_t)x.strides[0], call.dev<float>(2), (int64_t)y.strides[0]); });
  return (void)call.finish();
}
extern "C" int exoblas_ssbmv_row_major_Upper_stride_any_async(hipStream_t stream, int_fast32_t n, int_fast32_t k, const float *alpha, struct exo_win_2f32c A, struct exo_win_1f32c x, struct exo_win_1f32 y) {
  exoblas::Call call("exoblas_ssbmv_row_major_Upper_stride_any_async", exoblas::Call::kAsync, stream);
  if (n < 0) return call.fail_size("n", n);
  if (k < 0) return call.fail_size("k", k);
  if ((true) && A.strides[1] != 1) return call.fail_stride("A", A.strides[1]);
  if ((true) && n > 1 && A.strides[0] < k + 1) return call.fail_ld("A", A.strides[0]);
  call.window(0, A.data, 0, exoblas::span2_hi(n, k + 1, A.strides[0]), sizeof(float), false, true);
  call.window(1, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(float), false, true);
  call.window(2, y.data, exoblas::span_lo(n, y.strides[0]), exoblas::span_hi(n, y.strides[0]), sizeof(float), true, true);
  if (!call.stage()) return call.status();
  const float alpha_v = call.scalar(alpha);
  if (!call.ok()) return call.status();
  call.run([&] { exoblas::ops::sbmv_row_major_Upper<float>(call.ctx(), (int64_t)n, (int64_t)k, alpha_v, call.dev<const float>(0), (int64_t)A.strides[0], call.dev<const float>(1), (int64_t)x.strides[0], call.dev<float>(2), (int64_t)y.strides[0]); });
  return call.finish();
}

// src/level2/sbmv.py:19-33
extern "C" void exo_ssbmv_row_major_Upper_stride_1(void *ctxt, int_fast32_t n, int_fast32_t k, const float *alpha, struct exo_win_2f32c A, struct exo_win_1f32c x, struct exo_win_1f32 y) {
  exoblas::Call call("exo_ssbmv_row_major_Upper_stride_1", exoblas::Call::kSync, ctxt);
  if (n < 0) return (void)call.fail_size("n", n);
  if (k < 0) return (void)call.fail_size("k", k);
  if ((true) && A.strides[1] != 1) return (void)call.fail_stride("A", A.strides[1]);
  if ((true) && n > 1 && A.strides[0] < k + 1) return (void)call.fail_ld("A", A.strides[0]);
  call.window(0, A.data, 0, exoblas::span2_hi(n, k + 1, A.strides[0]), sizeof(float), false, true);
  call.window(1, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(float), false, true);
  call.window(2, y.data, exoblas::span_lo(n, y.strides[0]), exoblas::span_hi(n, y.strides[0]), sizeof(float), true, true);
  if (!call.stage()) return (void)call.status();
  const float alpha_v = call.scalar(alpha);
  if (!call.ok()) return (void)call.status();
  call.run([&] { exoblas::ops::sbmv_row_major_Upper<float>(call.ctx(), (int64_t)n, (int64_t)k, alpha_v, call.dev<const float>(0), (int64_t)A.strides[0], call.dev<const float>(1), (int64_t)x.strides[0], call.dev<float>(2), (int64_t)y.strides[0]); });
  return (void)call.finish();
}
extern "C" int exoblas_ssbmv_row_major_Upper_stride_1_async(hipStream_t stream, int_fast32_t n, int_fast32_t k, const float *alpha, struct exo_win_2f32c A, struct exo_win_1f32c x, struct exo_win_1f32 y) {
  exoblas::Call call("exoblas_ssbmv_row_major_Upper_stride_1_async", exoblas::Call::kAsync, stream);
  if (n < 0) return call.fail_size("n", n);
  if (k < 0) return call.fail_size("k", k);
  if ((true) && A.strides[1] != 1) return call.fail_stride("A", A.strides[1]);
  if ((true) && n > 1 && A.strides[0] < k + 1) return call.fail_ld("A", A.strides[0]);
  call.window(0, A.data, 0, exoblas::span2_hi(n, k + 1, A.strides[0]), sizeof(float), false, true);
  call.window(1, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(float), false, true);
  call.window(2, y.data, exoblas::span_lo(n, y.strides[0]), exoblas::span_hi(n, y.strides[0]), sizeof(float), true, true);
  if (!call.stage()) return call.status();
  const float alpha_v = call.scalar(alpha);
  if (!call.ok()) return call.status();
  call.run([&] { exoblas::ops::sbmv_row_major_Upper<float>(call.ctx(), (int64_t)n, (int64_t)k, alpha_v, call.dev<const float>(0), (int64_t)A.strides[0], call.dev<const float>(1), (int64_t)x.strides[0], call.dev<float>(2), (int64_t)y.strides[0]); });
  return call.finish();
}

// src/level2/sbmv.py:19-33
extern "C" void exo_dsbmv_row_major_Upper_stride_any(void *ctxt, int_fast32_t n, int_fast32_t k, const double *alpha, struct exo_win_2f64c A, struct exo_win_1f64c x, struct exo_win_1f64 y) {
  exoblas::Call call("exo_dsbmv_row_major_Upper_stride_any", exoblas::Call::kSync, ctxt);
  if (n < 0) return (void)call.fail_size("n", n);
  if (k < 0) return (void)call.fail_size("k", k);
  if ((true) && A.strides[1] != 1) return (void)call.fail_stride("A", A.strides[1]);
  if ((true) && n > 1 && A.strides[0] < k + 1) return (void)call.fail_ld("A", A.strides[0]);
  call.window(0, A.data, 0, exoblas::span2_hi(n, k + 1, A.strides[0]), sizeof(double), false, true);
  call.window(1, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(double), false, true);
  call.window(2, y.data, exoblas::span_lo(n, y.strides[0]), exoblas::span_hi(n, y.strides[0]), sizeof(double), true, true);
  if (!call.stage()) return (void)call.status();
  const double alpha_v = call.scalar(alpha);
  if (!call.ok()) return (void)call.status();
  call.run([&] { exoblas::ops::sbmv_row_major_Upper<double>(call.ctx(), (int64_t)n, (int64_t)k, alpha_v, call.dev<const double>(0), (int64_t)A.strides[0], call.dev<const double>(1), (int64_t)x.strides[0], call.dev<double>(2), (int64_t)y.strides[0]); });
  return (void)call.finish();
}
extern "C" int exoblas_dsbmv_row_major_Upper_stride_any_async(hipStream_t stream, int_fast32_t n, int_fast32_t k, const double *alpha, struct exo_win_2f64c A, struct exo_win_1f64c x, struct exo_win_1f64 y) {
  exoblas::Call call("exoblas_dsbmv_row_major_Upper_stride_any_async", exoblas::Call::kAsync, stream);
  if (n < 0) return call.fail_size("n", n);
  if (k < 0) return call.fail_size("k", k);
  if ((true) && A.strides[1] != 1) return call.fail_stride("A", A.strides[1]);
  if ((true) && n > 1 && A.strides[0] < k + 1) return call.fail_ld("A", A.strides[0]);
  call.window(0, A.data, 0, exoblas::span2_hi(n, k + 1, A.strides[0]), sizeof(double), false, true);
  call.window(1, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(double), false, true);
  call.window(2, y.data, exoblas::span_lo(n, y.strides[0]), exoblas::span_hi(n, y.strides[0]), sizeof(double), true, true);
  if (!call.stage()) return call.status();
  const double alpha_v = call.scalar(alpha);
  if (!call.ok()) return call.status();
  call.run([&] { exoblas::ops::sbmv_row_major_Upper<double>(call.ctx(), (int64_t)n, (int64_t)k, alpha_v, call.dev<const double>(0), (int64_t)A.strides[0], call.dev<const double>(1), (int64_t)x.strides[0], call.dev<double>(2), (int64_t)y.strides[0]); });
  return call.finish();
}

// src/level2/sbmv.py:19-33
extern "C" void exo_dsbmv_row_major_Upper_stride_1(void *ctxt, int_fast32_t n, int_fast32_t k, const double *alpha, struct exo_win_2f64c A, struct exo_win_1f64c x, struct exo_win_1f64 y) {
  exoblas::Call call("exo_dsbmv_row_major_Upper_stride_1", exoblas::Call::kSync, ctxt);
  if (n < 0) return (void)call.fail_size("n", n);
  if (k < 0) return (void)call.fail_size("k", k);
  if ((true) && A.strides[1] != 1) return (void)call.fail_stride("A", A.strides[1]);
  if ((true) && n > 1 && A.strides[0] < k + 1) return (void)call.fail_ld("A", A.strides[0]);
  call.window(0, A.data, 0, exoblas::span2_hi(n, k + 1, A.strides[0]), sizeof(double), false, true);
  call.window(1, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(double), false, true);
  call.window(2, y.data, exoblas::span_lo(n, y.strides[0]), exoblas::span_hi(n, y.strides[0]), sizeof(double), true, true);
  if (!call.stage()) return (void)call.status();
  const double alpha_v = call.scalar(alpha);
  if (!call.ok()) return (void)call.status();
  call.run([&] { exoblas::ops::sbmv_row_major_Upper<double>(call.ctx(), (int64_t)n, (int64_t)k, alpha_v, call.dev<const double>(0), (int64_t)A.strides[0], call.dev<const double>(1), (int64_t)x.strides[0], call.dev<double>(2), (int64_t)y.strides[0]); });
  return (void)call.finish();
}
extern "C" int exoblas_dsbmv_row_major_Upper_stride_1_async(hipStream_t stream, int_fast32_t n, int_fast32_t k, const double *alpha, struct exo_win_2f64c A, struct exo_win_1f64c x, struct exo_win_1f64 y) {
  exoblas::Call call("exoblas_dsbmv_row_major_Upper_stride_1_async", exoblas::Call::kAsync, stream);
  if (n < 0) return call.fail_size("n", n);
  if (k < 0) return call.fail_size("k", k);
  if ((true) && A.strides[1] != 1) return call.fail_stride("A", A.strides[1]);
  if ((true) && n > 1 && A.strides[0] < k + 1) return call.fail_ld("A", A.strides[0]);
  call.window(0, A.data, 0, exoblas::span2_hi(n, k + 1, A.strides[0]), sizeof(double), false, true);
  call.window(1, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(double), false, true);
  call.window(2, y.data, exoblas::span_lo(n, y.strides[0]), exoblas::span_hi(n, y.strides[0]), sizeof(double), true, true);
  if (!call.stage()) return call.status();
  const double alpha_v = call.scalar(alpha);
  if (!call.ok()) return call.status();
  call.run([&] { exoblas::ops::sbmv_row_major_Upper<double>(call.ctx(), (int64_t)n, (int64_t)k, alpha_v, call.dev<const double>(0), (int64_t)A.strides[0], call.dev<const double>(1), (int64_t)x.strides[0], call.dev<double>(2), (int64_t)y.strides[0]); });
  return call.finish();
}

// src/level2/sbmv.py:36-50
extern "C" void exo_ssbmv_row_major_Lower_stride_any(void *ctxt, int_fast32_t n, int_fast32_t k, const float *alpha, struct exo_win_2f32c A, struct exo_win_1f32c x, struct exo_win_1f32 y) {
  exoblas::Call call("exo_ssbmv_row_major_Lower_stride_any", exoblas::Call::kSync, ctxt);
  if (n < 0) return (void)call.fail_size("n", n);
  if (k < 0) return (void)call.fail_size("k", k);
  if ((true) && A.strides[1] != 1) return (void)call.fail_stride("A", A.strides[1]);
  if ((true) && n > 1 && A.strides[0] < k + 1) return (void)call.fail_ld("A", A.strides[0]);
  call.window(0, A.data, 0, exoblas::span2_hi(n, k + 1, A.strides[0]), sizeof(float), false, true);
  call.window(1, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(float), false, true);
  call.window(2, y.data, exoblas::span_lo(n, y.strides[0]), exoblas::span_hi(n, y.strides[0]), sizeof(float), true, true);
  if (!call.stage()) return (void)call.status();
  const float alpha_v = call.scalar(alpha);
  if (!call.ok()) return (void)call.status();
  call.run([&] { exoblas::ops::sbmv_row_major_Lower<float>(call.ctx(), (int64_t)n, (int64_t)k, alpha_v, call.dev<const float>(0), (int64_t)A.strides[0], call.dev<const float>(1), (int64_t)x.strides[0], call.dev<float>(2), (int64_t)y.strides[0]); });
  return (void)call.finish();
}
extern "C" int exoblas_ssbmv_row_major_Lower_stride_any_async(hipStream_t stream, int_fast32_t n, int_fast32_t k, const float *alpha, struct exo_win_2f32c A, struct exo_win_1f32c x, struct exo_win_1f32 y) {
  exoblas::Call call("exoblas_ssbmv_row_major_Lower_stride_any_async", exoblas::Call::kAsync, stream);
  if (n < 0) return call.fail_size("n", n);
  if (k < 0) return call.fail_size("k", k);
  if ((true) && A.strides[1] != 1) return call.fail_stride("A", A.strides[1]);
  if ((true) && n > 1 && A.strides[0] < k + 1) return call.fail_ld("A", A.strides[0]);
  call.window(0, A.data, 0, exoblas::span2_hi(n, k + 1, A.strides[0]), sizeof(float), false, true);
  call.window(1, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(float), false, true);
  call.window(2, y.data, exoblas::span_lo(n, y.strides[0]), exoblas::span_hi(n, y.strides[0]), sizeof(float), true, true);
  if (!call.stage()) return call.status();
  const float alpha_v = call.scalar(alpha);
  if (!call.ok()) return call.status();
  call.run([&] { exoblas::ops::sbmv_row_major_Lower<float>(call.ctx(), (int64_t)n, (int64_t)k, alpha_v, call.dev<const float>(0), (int64_t)A.strides[0], call.dev<const float>(1), (int64_t)x.strides[0], call.dev<float>(2), (int64_t)y.strides[0]); });
  return call.finish();
}

// src/level2/sbmv.py:36-50
extern "C" void exo_ssbmv_row_major_Lower_stride_1(void *ctxt, int_fast32_t n, int_fast32_t k, const float *alpha, struct exo_win_2f32c A, struct exo_win_1f32c x, struct exo_win_1f32 y) {
  exoblas::Call call("exo_ssbmv_row_major_Lower_stride_1", exoblas::Call::kSync, ctxt);
  if (n < 0) return (void)call.fail_size("n", n);
  if (k < 0) return (void)call.fail_size("k", k);
  if ((true) && A.strides[1] != 1) return (void)call.fail_stride("A", A.strides[1]);
  if ((true) && n > 1 && A.strides[0] < k + 1) return (void)call.fail_ld("A", A.strides[0]);
  call.window(0, A.data, 0, exoblas::span2_hi(n, k + 1, A.strides[0]), sizeof(float), false, true);
  call.window(1, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(float), false, true);
  call.window(2, y.data, exoblas::span_lo(n, y.strides[0]), exoblas::span_hi(n, y.strides[0]), sizeof(float), true, true);
  if (!call.stage()) return (void)call.status();
  const float alpha_v = call.scalar(alpha);
  if (!call.ok()) return (void)call.status();
  call.run([&] { exoblas::ops::sbmv_row_major_Lower<float>(call.ctx(), (int64_t)n, (int64_t)k, alpha_v, call.dev<const float>(0), (int64_t)A.strides[0], call.dev<const float>(1), (int64_t)x.strides[0], call.dev<float>(2), (int64_t)y.strides[0]); });
  return (void)call.finish();
}
extern "C" int exoblas_ssbmv_row_major_Lower_stride_1_async(hipStream_t stream, int_fast32_t n, int_fast32_t k, const float *alpha, struct exo_win_2f32c A, struct exo_win_1f32c x, struct exo_win_1f32 y) {
  exoblas::Call call("exoblas_ssbmv_row_major_Lower_stride_1_async", exoblas::Call::kAsync, stream);
  if (n < 0) return call.fail_size("n", n);
  if (k < 0) return call.fail_size("k", k);
  if ((true) && A.strides[1] != 1) return call.fail_stride("A", A.strides[1]);
  if ((true) && n > 1 && A.strides[0] < k + 1) return call.fail_ld("A", A.strides[0]);
  call.window(0, A.data, 0, exoblas::span2_hi(n, k + 1, A.strides[0]), sizeof(float), false, true);
  call.window(1, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(float), false, true);
  call.window(2, y.data, exoblas::span_lo(n, y.strides[0]), exoblas::span_hi(n, y.strides[0]), sizeof(float), true, true);
  if (!call.stage()) return call.status();
  const float alpha_v = call.scalar(alpha);
  if (!call.ok()) return call.status();
  call.run([&] { exoblas::ops::sbmv_row_major_Lower<float>(call.ctx(), (int64_t)n, (int64_t)k, alpha_v, call.dev<const float>(0), (int64_t)A.strides[0], call.dev<const float>(1), (int64_t)x.strides[0], call.dev<float>(2), (int64_t)y.strides[0]); });
  return call.finish();
}

// src/level2/sbmv.py:36-50
extern "C" void exo_dsbmv_row_major_Lower_stride_any(void *ctxt, int_fast32_t n, int_fast32_t k, const double *alpha, struct exo_win_2f64c A, struct exo_win_1f64c x, struct exo_win_1f64 y) {
  exoblas::Call call("exo_dsbmv_row_major_Lower_stride_any", exoblas::Call::kSync, ctxt);
  if (n < 0) return (void)call.fail_size("n", n);
  if (k < 0) return (void)call.fail_size("k", k);
  if ((true) && A.strides[1] != 1) return (void)call.fail_stride("A", A.strides[1]);
  if ((true) && n > 1 && A.strides[0] < k + 1) return (void)call.fail_ld("A", A.strides[0]);
  call.window(0, A.data, 0, exoblas::span2_hi(n, k + 1, A.strides[0]), sizeof(double), false, true);
  call.window(1, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(double), false, true);
  call.window(2, y.data, exoblas::span_lo(n, y.strides[0]), exoblas::span_hi(n, y.strides[0]), sizeof(double), true, true);
  if (!call.stage()) return (void)call.status();
  const double alpha_v = call.scalar(alpha);
  if (!call.ok()) return (void)call.status();
  call.run([&] { exoblas::ops::sbmv_row_major_Lower<double>(call.ctx(), (int64_t)n, (int64_t)k, alpha_v, call.dev<const double>(0), (int64_t)A.strides[0], call.dev<const double>(1), (int64_t)x.strides[0], call.dev<double>(2), (int64_t)y.strides[0]); });
  return (void)call.finish();
}
extern "C" int exoblas_dsbmv_row_major_Lower_stride_any_async(hipStream_t stream, int_fast32_t n, int_fast32_t k, const double *alpha, struct exo_win_2f64c A, struct exo_win_1f64c x, struct exo_win_1f64 y) {
  exoblas::Call call("exoblas_dsbmv_row_major_Lower_stride_any_async", exoblas::Call::kAsync, stream);
  if (n < 0) return call.fail_size("n", n);
  if (k < 0) return call.fail_size("k", k);
  if ((true) && A.strides[1] != 1) return call.fail_stride("A", A.strides[1]);
  if ((true) && n > 1 && A.strides[0] < k + 1) return call.fail_ld("A", A.strides[0]);
  call.window(0, A.data, 0, exoblas::span2_hi(n, k + 1, A.strides[0]), sizeof(double), false, true);
  call.window(1, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(double), false, true);
  call.window(2, y.data, exoblas::span_lo(n, y.strides[0]), exoblas::span_hi(n, y.strides[0]), sizeof(double), true, true);
  if (!call.stage()) return call.status();
  const double alpha_v = call.scalar(alpha);
  if (!call.ok()) return call.status();
  call.run([&] { exoblas::ops::sbmv_row_major_Lower<double>(call.ctx(), (int64_t)n, (int64_t)k, alpha_v, call.dev<const double>(0), (int64_t)A.strides[0], call.dev<const double>(1), (int64_t)x.strides[0], call.dev<double>(2), (int64_t)y.strides[0]); });
  return call.finish();
}

// src/level2/sbmv.py:36-50
extern "C" void exo_dsbmv_row_major_Lower_stride_1(void *ctxt, int_fast32_t n, int_fast32_t k, const double *alpha, struct exo_win_2f64c A, struct exo_win_1f64c x, struct exo_win_1f64 y) {
  exoblas::Call call("exo_dsbmv_row_major_Lower_stride_1", exoblas::Call::kSync, ctxt);
  if (n < 0) return (void)call.fail_size("n", n);
  if (k < 0) return (void)call.fail_size("k", k);
  if ((true) && A.strides[1] != 1) return (void)call.fail_stride("A", A.strides[1]);
  if ((true) && n > 1 && A.strides[0] < k + 1) return (void)call.fail_ld("A", A.strides[0]);
  call.window(0, A.data, 0, exoblas::span2_hi(n, k + 1, A.strides[0]), sizeof(double), false, true);
  call.window(1, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(double), false, true);
  call.window(2, y.data, exoblas::span_lo(n, y.strides[0]), exoblas::span_hi(n, y.strides[0]), sizeof(double), true, true);
  if (!call.stage()) return (void)call.status();
  const double alpha_v = call.scalar(alpha);
  if (!call.ok()) return (void)call.status();
  call.run([&] { exoblas::ops::sbmv_row_major_Lower<double>(call.ctx(), (int64_t)n, (int64_t)k, alpha_v, call.dev<const double>(0), (int64_t)A.strides[0], call.dev<const double>(1), (int64_t)x.strides[0], call.dev<double>(2), (int64_t)y.strides[0]); });
  return (void)call.finish();
}
extern "C" int exoblas_dsbmv_row_major_Lower_stride_1_async(hipStream_t stream, int_fast32_t n, int_fast32_t k, const double *alpha, struct exo_win_2f64c A, struct exo_win_1f64c x, struct exo_win_1f64 y) {
  exoblas::Call call("exoblas_dsbmv_row_major_Lower_stride_1_async", exoblas::Call::kAsync, stream);
  if (n < 0) return call.fail_size("n", n);
  if (k < 0) return call.fail_size("k", k);
  if ((true) && A.strides[1] != 1) return call.fail_stride("A", A.strides[1]);
  if ((true) && n > 1 && A.strides[0] < k + 1) return call.fail_ld("A", A.strides[0]);
  call.window(0, A.data, 0, exoblas::span2_hi(n, k + 1, A.strides[0]), sizeof(double), false, true);
  call.window(1, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(double), false, true);
  call.window(2, y.data, exoblas::span_lo(n, y.strides[0]), exoblas::span_hi(n, y.strides[0]), sizeof(double), true, true);
  if (!call.stage()) return call.status();
  const double alpha_v = call.scalar(alpha);
  if (!call.ok()) return call.status();
  call.run([&] { exoblas::ops::sbmv_row_major_Lower<double>(call.ctx(), (int64_t)n, (int64_t)k, alpha_v, call.dev<const double>(0), (int64_t)A.strides[0], call.dev<const double>(1), (int64_t)x.strides[0], call.dev<double>(2), (int64_t)y.strides[0]); });
  return call.finish();
}

// src/level2/tbmv.py:12-43
extern "C" void exo_stbmv_row_major_Upper_NonTrans_stride_any(void *ctxt, int_fast32_t n, int_fast32_t k, struct exo_win_1f32 x, struct exo_win_2f32c A, int_fast32_t Diag) {
  exoblas::Call call("exo_stbmv_row_major_Upper_NonTrans_stride_any", exoblas::Call::kSync, ctxt);
  if (n < 0) return (void)call.fail_size("n", n);
  if (k < 0) return (void)call.fail_size("k", k);
  if (Diag < 0) return (void)call.fail_size("Diag", Diag);
  if ((true) && A.strides[1] != 1) return (void)call.fail_stride("A", A.strides[1]);
  if ((true) && n > 1 && A.strides[0] < k + 1) return (void)call.fail_ld("A", A.strides[0]);
  call.window(0, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(float), true, true);
  call.window(1, A.data, 0, exoblas::span2_hi(n, k + 1, A.strides[0]), sizeof(float), false, true);
  if (!call.stage()) return (void)call.status();
  if (!call.ok()) return (void)call.status();
  call.run([&] { exoblas::ops::tbmv_row_major_Upper_NonTrans<float>(call.ctx(), (int64_t)n, (int64_t)k, call.dev<float>(0), (int64_t)x.strides[0], call.dev<const float>(1), (int64_t)A.strides[0], (int64_t)Diag); });
  return (void)call.finish();
}
extern "C" int exoblas_stbmv_row_major_Upper_NonTrans_stride_any_async(hipStream_t stream, int_fast32_t n, int_fast32_t k, struct exo_win_1f32 x, struct exo_win_2f32c A, int_fast32_t Diag) {
  exoblas::Call call("exoblas_stbmv_row_major_Upper_NonTrans_stride_any_async", exoblas::Call::kAsync, stream);
  if (n < 0) return call.fail_size("n", n);
  if (k < 0) return call.fail_size("k", k);
  if (Diag < 0) return call.fail_size("Diag", Diag);
  if ((true) && A.strides[1] != 1) return call.fail_stride("A", A.strides[1]);
  if ((true) && n > 1 && A.strides[0] < k + 1) return call.fail_ld("A", A.strides[0]);
  call.window(0, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(float), true, true);
  call.window(1, A.data, 0, exoblas::span2_hi(n, k + 1, A.strides[0]), sizeof(float), false, true);
  if (!call.stage()) return call.status();
  if (!call.ok()) return call.status();
  call.run([&] { exoblas::ops::tbmv_row_major_Upper_NonTrans<float>(call.ctx(), (int64_t)n, (int64_t)k, call.dev<float>(0), (int64_t)x.strides[0], call.dev<const float>(1), (int64_t)A.strides[0], (int64_t)Diag); });
  return call.finish();
}

// src/level2/tbmv.py:12-43
extern "C" void exo_stbmv_row_major_Upper_NonTrans_stride_1(void *ctxt, int_fast32_t n, int_fast32_t k, struct exo_win_1f32 x, struct exo_win_2f32c A, int_fast32_t Diag) {
  exoblas::Call call("exo_stbmv_row_major_Upper_NonTrans_stride_1", exoblas::Call::kSync, ctxt);
  if (n < 0) return (void)call.fail_size("n", n);
  if (k < 0) return (void)call.fail_size("k", k);
  if (Diag < 0) return (void)call.fail_size("Diag", Diag);
  if ((true) && A.strides[1] != 1) return (void)call.fail_stride("A", A.strides[1]);
  if ((true) && n > 1 && A.strides[0] < k + 1) return (void)call.fail_ld("A", A.strides[0]);
  call.window(0, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(float), true, true);
  call.window(1, A.data, 0, exoblas::span2_hi(n, k + 1, A.strides[0]), sizeof(float), false, true);
  if (!call.stage()) return (void)call.status();
  if (!call.ok()) return (void)call.status();
  call.run([&] { exoblas::ops::tbmv_row_major_Upper_NonTrans<float>(call.ctx(), (int64_t)n, (int64_t)k, call.dev<float>(0), (int64_t)x.strides[0], call.dev<const float>(1), (int64_t)A.strides[0], (int64_t)Diag); });
  return (void)call.finish();
}
extern "C" int exoblas_stbmv_row_major_Upper_NonTrans_stride_1_async(hipStream_t stream, int_fast32_t n, int_fast32_t k, struct exo_win_1f32 x, struct exo_win_2f32c A, int_fast32_t Diag) {
  exoblas::Call call("exoblas_stbmv_row_major_Upper_NonTrans_stride_1_async", exoblas::Call::kAsync, stream);
  if (n < 0) return call.fail_size("n", n);
  if (k < 0) return call.fail_size("k", k);
  if (Diag < 0) return call.fail_size("Diag", Diag);
  if ((true) && A.strides[1] != 1) return call.fail_stride("A", A.strides[1]);
  if ((true) && n > 1 && A.strides[0] < k + 1) return call.fail_ld("A", A.strides[0]);
  call.window(0, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(float), true, true);
  call.window(1, A.data, 0, exoblas::span2_hi(n, k + 1, A.strides[0]), sizeof(float), false, true);
  if (!call.stage()) return call.status();
  if (!call.ok()) return call.status();
  call.run([&] { exoblas::ops::tbmv_row_major_Upper_NonTrans<float>(call.ctx(), (int64_t)n, (int64_t)k, call.dev<float>(0), (int64_t)x.strides[0], call.dev<const float>(1), (int64_t)A.strides[0], (int64_t)Diag); });
  return call.finish();
}

// src/level2/tbmv.py:12-43
extern "C" void exo_dtbmv_row_major_Upper_NonTrans_stride_any(void *ctxt, int_fast32_t n, int_fast32_t k, struct exo_win_1f64 x, struct exo_win_2f64c A, int_fast32_t Diag) {
  exoblas::Call call("exo_dtbmv_row_major_Upper_NonTrans_stride_any", exoblas::Call::kSync, ctxt);
  if (n < 0) return (void)call.fail_size("n", n);
  if (k < 0) return (void)call.fail_size("k", k);
  if (Diag < 0) return (void)call.fail_size("Diag", Diag);
  if ((true) && A.strides[1] != 1) return (void)call.fail_stride("A", A.strides[1]);
  if ((true) && n > 1 && A.strides[0] < k + 1) return (void)call.fail_ld("A", A.strides[0]);
  call.window(0, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(double), true, true);
  call.window(1, A.data, 0, exoblas::span2_hi(n, k + 1, A.strides[0]), sizeof(double), false, true);
  if (!call.stage()) return (void)call.status();
  if (!call.ok()) return (void)call.status();
  call.run([&] { exoblas::ops::tbmv_row_major_Upper_NonTrans<double>(call.ctx(), (int64_t)n, (int64_t)k, call.dev<double>(0), (int64_t)x.strides[0], call.dev<const double>(1), (int64_t)A.strides[0], (int64_t)Diag); });
  return (void)call.finish();
}
extern "C" int exoblas_dtbmv_row_major_Upper_NonTrans_stride_any_async(hipStream_t stream, int_fast32_t n, int_fast32_t k, struct exo_win_1f64 x, struct exo_win_2f64c A, int_fast32_t Diag) {
  exoblas::Call call("exoblas_dtbmv_row_major_Upper_NonTrans_stride_any_async", exoblas::Call::kAsync, stream);
  if (n < 0) return call.fail_size("n", n);
  if (k < 0) return call.fail_size("k", k);
  if (Diag < 0) return call.fail_size("Diag", Diag);
  if ((true) && A.strides[1] != 1) return call.fail_stride("A", A.strides[1]);
  if ((true) && n > 1 && A.strides[0] < k + 1) return call.fail_ld("A", A.strides[0]);
  call.window(0, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(double), true, true);
  call.window(1, A.data, 0, exoblas::span2_hi(n, k + 1, A.strides[0]), sizeof(double), false, true);
  if (!call.stage()) return call.status();
  if (!call.ok()) return call.status();
  call.run([&] { exoblas::ops::tbmv_row_major_Upper_NonTrans<double>(call.ctx(), (int64_t)n, (int64_t)k, call.dev<double>(0), (int64_t)x.strides[0], call.dev<const double>(1), (int64_t)A.strides[0], (int64_t)Diag); });
  return call.finish();
}

// src/level2/tbmv.py:12-43
extern "C" void exo_dtbmv_row_major_Upper_NonTrans_stride_1(void *ctxt, int_fast32_t n, int_fast32_t k, struct exo_win_1f64 x, struct exo_win_2f64c A, int_fast32_t Diag) {
  exoblas::Call call("exo_dtbmv_row_major_Upper_NonTrans_stride_1", exoblas::Call::kSync, ctxt);
  if (n < 0) return (void)call.fail_size("n", n);
  if (k < 0) return (void)call.fail_size("k", k);
  if (Diag < 0) return (void)call.fail_size("Diag", Diag);
  if ((true) && A.strides[1] != 1) return (void)call.fail_stride("A", A.strides[1]);
  if ((true) && n > 1 && A.strides[0] < k + 1) return (void)call.fail_ld("A", A.strides[0]);
  call.window(0, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(double), true, true);
  call.window(1, A.data, 0, exoblas::span2_hi(n, k + 1, A.strides[0]), sizeof(double), false, true);
  if (!call.stage()) return (void)call.status();
  if (!call.ok()) return (void)call.status();
  call.run([&] { exoblas::ops::tbmv_row_major_Upper_NonTrans<double>(call.ctx(), (int64_t)n, (int64_t)k, call.dev<double>(0), (int64_t)x.strides[0], call.dev<const double>(1), (int64_t)A.strides[0], (int64_t)Diag); });
  return (void)call.finish();
}
extern "C" int exoblas_dtbmv_row_major_Upper_NonTrans_stride_1_async(hipStream_t stream, int_fast32_t n, int_fast32_t k, struct exo_win_1f64 x, struct exo_win_2f64c A, int_fast32_t Diag) {
  exoblas::Call call("exoblas_dtbmv_row_major_Upper_NonTrans_stride_1_async", exoblas::Call::kAsync, stream);
  if (n < 0) return call.fail_size("n", n);
  if (k < 0) return call.fail_size("k", k);
  if (Diag < 0) return call.fail_size("Diag", Diag);
  if ((true) && A.strides[1] != 1) return call.fail_stride("A", A.strides[1]);
  if ((true) && n > 1 && A.strides[0] < k + 1) return call.fail_ld("A", A.strides[0]);
  call.window(0, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(double), true, true);
  call.window(1, A.data, 0, exoblas::span2_hi(n, k + 1, A.strides[0]), sizeof(double), false, true);
  if (!call.stage()) return call.status();
  if (!call.ok()) return call.status();
  call.run([&] { exoblas::ops::tbmv_row_major_Upper_NonTrans<double>(call.ctx(), (int64_t)n, (int64_t)k, call.dev<double>(0), (int64_t)x.strides[0], call.dev<const double>(1), (int64_t)A.strides[0], (int64_t)Diag); });
  return call.finish();
}

// src/level2/tbmv.py:46-77
extern "C" void exo_stbmv_row_major_Lower_NonTrans_stride_any(void *ctxt, int_fast32_t n, int_fast32_t k, struct exo_win_1f32 x, struct exo_win_2f32c A, int_fast32_t Diag) {
  exoblas::Call call("exo_stbmv_row_major_Lower_NonTrans_stride_any", exoblas::Call::kSync, ctxt);
  if (n < 0) return (void)call.fail_size("n", n);
  if (k < 0) return (void)call.fail_size("k", k);
  if (Diag < 0) return (void)call.fail_size("Diag", Diag);
  if ((true) && A.strides[1] != 1) return (void)call.fail_stride("A", A.strides[1]);
  if ((true) && n > 1 && A.strides[0] < k + 1) return (void)call.fail_ld("A", A.strides[0]);
  call.window(0, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(float), true, true);
  call.window(1, A.data, 0, exoblas::span2_hi(n, k + 1, A.strides[0]), sizeof(float), false, true);
  if (!call.stage()) return (void)call.status();
  if (!call.ok()) return (void)call.status();
  call.run([&] { exoblas::ops::tbmv_row_major_Lower_NonTrans<float>(call.ctx(), (int64_t)n, (int64_t)k, call.dev<float>(0), (int64_t)x.strides[0], call.dev<const float>(1), (int64_t)A.strides[0], (int64_t)Diag); });
  return (void)call.finish();
}
extern "C" int exoblas_stbmv_row_major_Lower_NonTrans_stride_any_async(hipStream_t stream, int_fast32_t n, int_fast32_t k, struct exo_win_1f32 x, struct exo_win_2f32c A, int_fast32_t Diag) {
  exoblas::Call call("exoblas_stbmv_row_major_Lower_NonTrans_stride_any_async", exoblas::Call::kAsync, stream);
  if (n < 0) return call.fail_size("n", n);
  if (k < 0) return call.fail_size("k", k);
  if (Diag < 0) return call.fail_size("Diag", Diag);
  if ((true) && A.strides[1] != 1) return call.fail_stride("A", A.strides[1]);
  if ((true) && n > 1 && A.strides[0] < k + 1) return call.fail_ld("A", A.strides[0]);
  call.window(0, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(float), true, true);
  call.window(1, A.data, 0, exoblas::span2_hi(n, k + 1, A.strides[0]), sizeof(float), false, true);
  if (!call.stage()) return call.status();
  if (!call.ok()) return call.status();
  call.run([&] { exoblas::ops::tbmv_row_major_Lower_NonTrans<float>(call.ctx(), (int64_t)n, (int64_t)k, call.dev<float>(0), (int64_t)x.strides[0], call.dev<const float>(1), (int64_t)A.strides[0], (int64_t)Diag); });
  return call.finish();
}

// src/level2/tbmv.py:46-77
extern "C" void exo_stbmv_row_major_Lower_NonTrans_stride_1(void *ctxt, int_fast32_t n, int_fast32_t k, struct exo_win_1f32 x, struct exo_win_2f32c A, int_fast32_t Diag) {
  exoblas::Call call("exo_stbmv_row_major_Lower_NonTrans_stride_1", exoblas::Call::kSync, ctxt);
  if (n < 0) return (void)call.fail_size("n", n);
  if (k < 0) return (void)call.fail_size("k", k);
  if (Diag < 0) return (void)call.fail_size("Diag", Diag);
  if ((true) && A.strides[1] != 1) return (void)call.fail_stride("A", A.strides[1]);
  if ((true) && n > 1 && A.strides[0] < k + 1) return (void)call.fail_ld("A", A.strides[0]);
  call.window(0, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(float), true, true);
  call.window(1, A.data, 0, exoblas::span2_hi(n, k + 1, A.strides[0]), sizeof(float), false, true);
  if (!call.stage()) return (void)call.status();
  if (!call.ok()) return (void)call.status();
  call.run([&] { exoblas::ops::tbmv_row_major_Lower_NonTrans<float>(call.ctx(), (int64_t)n, (int64_t)k, call.dev<float>(0), (int64_t)x.strides[0], call.dev<const float>(1), (int64_t)A.strides[0], (int64_t)Diag); });
  return (void)call.finish();
}
extern "C" int exoblas_stbmv_row_major_Lower_NonTrans_stride_1_async(hipStream_t stream, int_fast32_t n, int_fast32_t k, struct exo_win_1f32 x, struct exo_win_2f32c A, int_fast32_t Diag) {
  exoblas::Call call("exoblas_stbmv_row_major_Lower_NonTrans_stride_1_async", exoblas::Call::kAsync, stream);
  if (n < 0) return call.fail_size("n", n);
  if (k < 0) return call.fail_size("k", k);
  if (Diag < 0) return call.fail_size("Diag", Diag);
  if ((true) && A.strides[1] != 1) return call.fail_stride("A", A.strides[1]);
  if ((true) && n > 1 && A.strides[0] < k + 1) return call.fail_ld("A", A.strides[0]);
  call.window(0, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(float), true, true);
  call.window(1, A.data, 0, exoblas::span2_hi(n, k + 1, A.strides[0]), sizeof(float), false, true);
  if (!call.stage()) return call.status();
  if (!call.ok()) return call.status();
  call.run([&] { exoblas::ops::tbmv_row_major_Lower_NonTrans<float>(call.ctx(), (int64_t)n, (int64_t)k, call.dev<float>(0), (int64_t)x.strides[0], call.dev<const float>(1), (int64_t)A.strides[0], (int64_t)Diag); });
  return call.finish();
}

// src/level2/tbmv.py:46-77
extern "C" void exo_dtbmv_row_major_Lower_NonTrans_stride_any(void *ctxt, int_fast32_t n, int_fast32_t k, struct exo_win_1f64 x, struct exo_win_2f64c A, int_fast32_t Diag) {
  exoblas::Call call("exo_dtbmv_row_major_Lower_NonTrans_stride_any", exoblas::Call::kSync, ctxt);
  if (n < 0) return (void)call.fail_size("n", n);
  if (k < 0) return (void)call.fail_size("k", k);
  if (Diag < 0) return (void)call.fail_size("Diag", Diag);
  if ((true) && A.strides[1] != 1) return (void)call.fail_stride("A", A.strides[1]);
  if ((true) && n > 1 && A.strides[0] < k + 1) return (void)call.fail_ld("A", A.strides[0]);
  call.window(0, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(double), true, true);
  call.window(1, A.data, 0, exoblas::span2_hi(n, k + 1, A.strides[0]), sizeof(double), false, true);
  if (!call.stage()) return (void)call.status();
  if (!call.ok()) return (void)call.status();
  call.run([&] { exoblas::ops::tbmv_row_major_Lower_NonTrans<double>(call.ctx(), (int64_t)n, (int64_t)k, call.dev<double>(0), (int64_t)x.strides[0], call.dev<const double>(1), (int64_t)A.strides[0], (int64_t)Diag); });
  return (void)call.finish();
}
extern "C" int exoblas_dtbmv_row_major_Lower_NonTrans_stride_any_async(hipStream_t stream, int_fast32_t n, int_fast32_t k, struct exo_win_1f64 x, struct exo_win_2f64c A, int_fast32_t Diag) {
  exoblas::Call call("exoblas_dtbmv_row_major_Lower_NonTrans_stride_any_async", exoblas::Call::kAsync, stream);
  if (n < 0) return call.fail_size("n", n);
  if (k < 0) return call.fail_size("k", k);
  if (Diag < 0) return call.fail_size("Diag", Diag);
  if ((true) && A.strides[1] != 1) return call.fail_stride("A", A.strides[1]);
  if ((true) && n > 1 && A.strides[0] < k + 1) return call.fail_ld("A", A.strides[0]);
  call.window(0, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(double), true, true);
  call.window(1, A.data, 0, exoblas::span2_hi(n, k + 1, A.strides[0]), sizeof(double), false, true);
  if (!call.stage()) return call.status();
  if (!call.ok()) return call.status();
  call.run([&] { exoblas::ops::tbmv_row_major_Lower_NonTrans<double>(call.ctx(), (int64_t)n, (int64_t)k, call.dev<double>(0), (int64_t)x.strides[0], call.dev<const double>(1), (int64_t)A.strides[0], (int64_t)Diag); });
  return call.finish();
}

// src/level2/tbmv.py:46-77
extern "C" void exo_dtbmv_row_major_Lower_NonTrans_stride_1(void *ctxt, int_fast32_t n, int_fast32_t k, struct exo_win_1f64 x, struct exo_win_2f64c A, int_fast32_t Diag) {
  exoblas::Call call("exo_dtbmv_row_major_Lower_NonTrans_stride_1", exoblas::Call::kSync, ctxt);
  if (n < 0) return (void)call.fail_size("n", n);
  if (k < 0) return (void)call.fail_size("k", k);
  if (Diag < 0) return (void)call.fail_size("Diag", Diag);
  if ((true) && A.strides[1] != 1) return (void)call.fail_stride("A", A.strides[1]);
  if ((true) && n > 1 && A.strides[0] < k + 1) return (void)call.fail_ld("A", A.strides[0]);
  call.window(0, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(double), true, true);
  call.window(1, A.data, 0, exoblas::span2_hi(n, k + 1, A.strides[0]), sizeof(double), false, true);
  if (!call.stage()) return (void)call.status();
  if (!call.ok()) return (void)call.status();
  call.run([&] { exoblas::ops::tbmv_row_major_Lower_NonTrans<double>(call.ctx(), (int64_t)n, (int64_t)k, call.dev<double>(0), (int64_t)x.strides[0], call.dev<const double>(1), (int64_t)A.strides[0], (int64_t)Diag); });
  return (void)call.finish();
}
extern "C" int exoblas_dtbmv_row_major_Lower_NonTrans_stride_1_async(hipStream_t stream, int_fast32_t n, int_fast32_t k, struct exo_win_1f64 x, struct exo_win_2f64c A, int_fast32_t Diag) {
  exoblas::Call call("exoblas_dtbmv_row_major_Lower_NonTrans_stride_1_async", exoblas::Call::kAsync, stream);
  if (n < 0) return call.fail_size("n", n);
  if (k < 0) return call.fail_size("k", k);
  if (Diag < 0) return call.fail_size("Diag", Diag);
  if ((true) && A.strides[1] != 1) return call.fail_stride("A", A.strides[1]);
  if ((true) && n > 1 && A.strides[0] < k + 1) return call.fail_ld("A", A.strides[0]);
  call.window(0, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(double), true, true);
  call.window(1, A.data, 0, exoblas::span2_hi(n, k + 1, A.strides[0]), sizeof(double), false, true);
  if (!call.stage()) return call.status();
  if (!call.ok()) return call.status();
  call.run([&] { exoblas::ops::tbmv_row_major_Lower_NonTrans<double>(call.ctx(), (int64_t)n, (int64_t)k, call.dev<double>(0), (int64_t)x.strides[0], call.dev<const double>(1), (int64_t)A.strides[0], (int64_t)Diag); });
  return call.finish();
}

// src/level2/tbmv.py:80-108
extern "C" void exo_stbmv_row_major_Upper_Trans_stride_any(void *ctxt, int_fast32_t n, int_fast32_t k, struct exo_win_1f32 x, struct exo_win_2f32c A, int_fast32_t Diag) {
  exoblas::Call call("exo_stbmv_row_major_Upper_Trans_stride_any", exoblas::Call::kSync, ctxt);
  if (n < 0) return (void)call.fail_size("n", n);
  if (k < 0) return (void)call.fail_size("k", k);
  if (Diag < 0) return (void)call.fail_size("Diag", Diag);
  if ((true) && A.strides[1] != 1) return (void)call.fail_stride("A", A.strides[1]);
  if ((true) && n > 1 && A.strides[0] < k + 1) return (void)call.fail_ld("A", A.strides[0]);
  call.window(0, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(float), true, true);
  call.window(1, A.data, 0, exoblas::span2_hi(n, k + 1, A.strides[0]), sizeof(float), false, true);
  if (!call.stage()) return (void)call.status();
  if (!call.ok()) return (void)call.status();
  call.run([&] { exoblas::ops::tbmv_row_major_Upper_Trans<float>(call.ctx(), (int64_t)n, (int64_t)k, call.dev<float>(0), (int64_t)x.strides[0], call.dev<const float>(1), (int64_t)A.strides[0], (int64_t)Diag); });
  return (void)call.finish();
}
extern "C" int exoblas_stbmv_row_major_Upper_Trans_stride_any_async(hipStream_t stream, int_fast32_t n, int_fast32_t k, struct exo_win_1f32 x, struct exo_win_2f32c A, int_fast32_t Diag) {
  exoblas::Call call("exoblas_stbmv_row_major_Upper_Trans_stride_any_async", exoblas::Call::kAsync, stream);
  if (n < 0) return call.fail_size("n", n);
  if (k < 0) return call.fail_size("k", k);
  if (Diag < 0) return call.fail_size("Diag", Diag);
  if ((true) && A.strides[1] != 1) return call.fail_stride("A", A.strides[1]);
  if ((true) && n > 1 && A.strides[0] < k + 1) return call.fail_ld("A", A.strides[0]);
  call.window(0, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(float), true, true);
  call.window(1, A.data, 0, exoblas::span2_hi(n, k + 1, A.strides[0]), sizeof(float), false, true);
  if (!call.stage()) return call.status();
  if (!call.ok()) return call.status();
  call.run([&] { exoblas::ops::tbmv_row_major_Upper_Trans<float>(call.ctx(), (int64_t)n, (int64_t)k, call.dev<float>(0), (int64_t)x.strides[0], call.dev<const float>(1), (int64_t)A.strides[0], (int64_t)Diag); });
  return call.finish();
}

// src/level2/tbmv.py:80-108
extern "C" void exo_stbmv_row_major_Upper_Trans_stride_1(void *ctxt, int_fast32_t n, int_fast32_t k, struct exo_win_1f32 x, struct exo_win_2f32c A, int_fast32_t Diag) {
  exoblas::Call call("exo_stbmv_row_major_Upper_Trans_stride_1", exoblas::Call::kSync, ctxt);
  if (n < 0) return (void)call.fail_size("n", n);
  if (k < 0) return (void)call.fail_size("k", k);
  if (Diag < 0) return (void)call.fail_size("Diag", Diag);
  if ((true) && A.strides[1] != 1) return (void)call.fail_stride("A", A.strides[1]);
  if ((true) && n > 1 && A.strides[0] < k + 1) return (void)call.fail_ld("A", A.strides[0]);
  call.window(0, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(float), true, true);
  call.window(1, A.data, 0, exoblas::span2_hi(n, k + 1, A.strides[0]), sizeof(float), false, true);
  if (!call.stage()) return (void)call.status();
  if (!call.ok()) return (void)call.status();
  call.run([&] { exoblas::ops::tbmv_row_major_Upper_Trans<float>(call.ctx(), (int64_t)n, (int64_t)k, call.dev<float>(0), (int64_t)x.strides[0], call.dev<const float>(1), (int64_t)A.strides[0], (int64_t)Diag); });
  return (void)call.finish();
}
extern "C" int exoblas_stbmv_row_major_Upper_Trans_stride_1_async(hipStream_t stream, int_fast32_t n, int_fast32_t k, struct exo_win_1f32 x, struct exo_win_2f32c A, int_fast32_t Diag) {
  exoblas::Call call("exoblas_stbmv_row_major_Upper_Trans_stride_1_async", exoblas::Call::kAsync, stream);
  if (n < 0) return call.fail_size("n", n);
  if (k < 0) return call.fail_size("k", k);
  if (Diag < 0) return call.fail_size("Diag", Diag);
  if ((true) && A.strides[1] != 1) return call.fail_stride("A", A.strides[1]);
  if ((true) && n > 1 && A.strides[0] < k + 1) return call.fail_ld("A", A.strides[0]);
  call.window(0, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(float), true, true);
  call.window(1, A.data, 0, exoblas::span2_hi(n, k + 1, A.strides[0]), sizeof(float), false, true);
  if (!call.stage()) return call.status();
  if (!call.ok()) return call.status();
  call.run([&] { exoblas::ops::tbmv_row_major_Upper_Trans<float>(call.ctx(), (int64_t)n, (int64_t)k, call.dev<float>(0), (int64_t)x.strides[0], call.dev<const float>(1), (int64_t)A.strides[0], (int64_t)Diag); });
  return call.finish();
}

// src/level2/tbmv.py:80-108
extern "C" void exo_dtbmv_row_major_Upper_Trans_stride_any(void *ctxt, int_fast32_t n, int_fast32_t k, struct exo_win_1f64 x, struct exo_win_2f64c A, int_fast32_t Diag) {
  exoblas::Call call("exo_dtbmv_row_major_Upper_Trans_stride_any", exoblas::Call::kSync, ctxt);
  if (n < 0) return (void)call.fail_size("n", n);
  if (k < 0) return (void)call.fail_size("k", k);
  if (Diag < 0) return (void)call.fail_size("Diag", Diag);
  if ((true) && A.strides[1] != 1) return (void)call.fail_stride("A", A.strides[1]);
  if ((true) && n > 1 && A.strides[0] < k + 1) return (void)call.fail_ld("A", A.strides[0]);
  call.window(0, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(double), true, true);
  call.window(1, A.data, 0, exoblas::span2_hi(n, k + 1, A.strides[0]), sizeof(double), false, true);
  if (!call.stage()) return (void)call.status();
  if (!call.ok()) return (void)call.status();
  call.run([&] { exoblas::ops::tbmv_row_major_Upper_Trans<double>(call.ctx(), (int64_t)n, (int64_t)k, call.dev<double>(0), (int64_t)x.strides[0], call.dev<const double>(1), (int64_t)A.strides[0], (int64_t)Diag); });
  return (void)call.finish();
}
extern "C" int exoblas_dtbmv_row_major_Upper_Trans_stride_any_async(hipStream_t stream, int_fast32_t n, int_fast32_t k, struct exo_win_1f64 x, struct exo_win_2f64c A, int_fast32_t Diag) {
  exoblas::Call call("exoblas_dtbmv_row_major_Upper_Trans_stride_any_async", exoblas::Call::kAsync, stream);
  if (n < 0) return call.fail_size("n", n);
  if (k < 0) return call.fail_size("k", k);
  if (Diag < 0) return call.fail_size("Diag", Diag);
  if ((true) && A.strides[1] != 1) return call.fail_stride("A", A.strides[1]);
  if ((true) && n > 1 && A.strides[0] < k + 1) return call.fail_ld("A", A.strides[0]);
  call.window(0, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(double), true, true);
  call.window(1, A.data, 0, exoblas::span2_hi(n, k + 1, A.strides[0]), sizeof(double), false, true);
  if (!call.stage()) return call.status();
  if (!call.ok()) return call.status();
  call.run([&] { exoblas::ops::tbmv_row_major_Upper_Trans<double>(call.ctx(), (int64_t)n, (int64_t)k, call.dev<double>(0), (int64_t)x.strides[0], call.dev<const double>(1), (int64_t)A.strides[0], (int64_t)Diag); });
  return call.finish();
}

// src/level2/tbmv.py:80-108
extern "C" void exo_dtbmv_row_major_Upper_Trans_stride_1(void *ctxt, int_fast32_t n, int_fast32_t k, struct exo_win_1f64 x, struct exo_win_2f64c A, int_fast32_t Diag) {
  exoblas::Call call("exo_dtbmv_row_major_Upper_Trans_stride_1", exoblas::Call::kSync, ctxt);
  if (n < 0) return (void)call.fail_size("n", n);
  if (k < 0) return (void)call.fail_size("k", k);
  if (Diag < 0) return (void)call.fail_size("Diag", Diag);
  if ((true) && A.strides[1] != 1) return (void)call.fail_stride("A", A.strides[1]);
  if ((true) && n > 1 && A.strides[0] < k + 1) return (void)call.fail_ld("A", A.strides[0]);
  call.window(0, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(double), true, true);
  call.window(1, A.data, 0, exoblas::span2_hi(n, k + 1, A.strides[0]), sizeof(double), false, true);
  if (!call.stage()) return (void)call.status();
  if (!call.ok()) return (void)call.status();
  call.run([&] { exoblas::ops::tbmv_row_major_Upper_Trans<double>(call.ctx(), (int64_t)n, (int64_t)k, call.dev<double>(0), (int64_t)x.strides[0], call.dev<const double>(1), (int64_t)A.strides[0], (int64_t)Diag); });
  return (void)call.finish();
}
extern "C" int exoblas_dtbmv_row_major_Upper_Trans_stride_1_async(hipStream_t stream, int_fast32_t n, int_fast32_t k, struct exo_win_1f64 x, struct exo_win_2f64c A, int_fast32_t Diag) {
  exoblas::Call call("exoblas_dtbmv_row_major_Upper_Trans_stride_1_async", exoblas::Call::kAsync, stream);
  if (n < 0) return call.fail_size("n", n);
  if (k < 0) return call.fail_size("k", k);
  if (Diag < 0) return call.fail_size("Diag", Diag);
  if ((true) && A.strides[1] != 1) return call.fail_stride("A", A.strides[1]);
  if ((true) && n > 1 && A.strides[0] < k + 1) return call.fail_ld("A", A.strides[0]);
  call.window(0, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(double), true, true);
  call.window(1, A.data, 0, exoblas::span2_hi(n, k + 1, A.strides[0]), sizeof(double), false, true);
  if (!call.stage()) return call.status();
  if (!call.ok()) return call.status();
  call.run([&] { exoblas::ops::tbmv_row_major_Upper_Trans<double>(call.ctx(), (int64_t)n, (int64_t)k, call.dev<double>(0), (int64_t)x.strides[0], call.dev<const double>(1), (int64_t)A.strides[0], (int64_t)Diag); });
  return call.finish();
}

// src/level2/tbmv.py:111-141
extern "C" void exo_stbmv_row_major_Lower_Trans_stride_any(void *ctxt, int_fast32_t n, int_fast32_t k, struct exo_win_1f32 x, struct exo_win_2f32c A, int_fast32_t Diag) {
  exoblas::Call call("exo_stbmv_row_major_Lower_Trans_stride_any", exoblas::Call::kSync, ctxt);
  if (n < 0) return (void)call.fail_size("n", n);
  if (k < 0) return (void)call.fail_size("k", k);
  if (Diag < 0) return (void)call.fail_size("Diag", Diag);
  if ((true) && A.strides[1] != 1) return (void)call.fail_stride("A", A.strides[1]);
  if ((true) && n > 1 && A.strides[0] < k + 1) return (void)call.fail_ld("A", A.strides[0]);
  call.window(0, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(float), true, true);
  call.window(1, A.data, 0, exoblas::span2_hi(n, k + 1, A.strides[0]), sizeof(float), false, true);
  if (!call.stage()) return (void)call.status();
  if (!call.ok()) return (void)call.status();
  call.run([&] { exoblas::ops::tbmv_row_major_Lower_Trans<float>(call.ctx(), (int64_t)n, (int64_t)k, call.dev<float>(0), (int64_t)x.strides[0], call.dev<const float>(1), (int64_t)A.strides[0], (int64_t)Diag); });
  return (void)call.finish();
}
extern "C" int exoblas_stbmv_row_major_Lower_Trans_stride_any_async(hipStream_t stream, int_fast32_t n, int_fast32_t k, struct exo_win_1f32 x, struct exo_win_2f32c A, int_fast32_t Diag) {
  exoblas::Call call("exoblas_stbmv_row_major_Lower_Trans_stride_any_async", exoblas::Call::kAsync, stream);
  if (n < 0) return call.fail_size("n", n);
  if (k < 0) return call.fail_size("k", k);
  if (Diag < 0) return call.fail_size("Diag", Diag);
  if ((true) && A.strides[1] != 1) return call.fail_stride("A", A.strides[1]);
  if ((true) && n > 1 && A.strides[0] < k + 1) return call.fail_ld("A", A.strides[0]);
  call.window(0, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(float), true, true);
  call.window(1, A.data, 0, exoblas::span2_hi(n, k + 1, A.strides[0]), sizeof(float), false, true);
  if (!call.stage()) return call.status();
  if (!call.ok()) return call.status();
  call.run([&] { exoblas::ops::tbmv_row_major_Lower_Trans<float>(call.ctx(), (int64_t)n, (int64_t)k, call.dev<float>(0), (int64_t)x.strides[0], call.dev<const float>(1), (int64_t)A.strides[0], (int64_t)Diag); });
  return call.finish();
}

// src/level2/tbmv.py:111-141
extern "C" void exo_stbmv_row_major_Lower_Trans_stride_1(void *ctxt, int_fast32_t n, int_fast32_t k, struct exo_win_1f32 x, struct exo_win_2f32c A, int_fast32_t Diag) {
  exoblas::Call call("exo_stbmv_row_major_Lower_Trans_stride_1", exoblas::Call::kSync, ctxt);
  if (n < 0) return (void)call.fail_size("n", n);
  if (k < 0) return (void)call.fail_size("k", k);
  if (Diag < 0) return (void)call.fail_size("Diag", Diag);
  if ((true) && A.strides[1] != 1) return (void)call.fail_stride("A", A.strides[1]);
  if ((true) && n > 1 && A.strides[0] < k + 1) return (void)call.fail_ld("A", A.strides[0]);
  call.window(0, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(float), true, true);
  call.window(1, A.data, 0, exoblas::span2_hi(n, k + 1, A.strides[0]), sizeof(float), false, true);
  if (!call.stage()) return (void)call.status();
  if (!call.ok()) return (void)call.status();
  call.run([&] { exoblas::ops::tbmv_row_major_Lower_Trans<float>(call.ctx(), (int64_t)n, (int64_t)k, call.dev<float>(0), (int64_t)x.strides[0], call.dev<const float>(1), (int64_t)A.strides[0], (int64_t)Diag); });
  return (void)call.finish();
}
extern "C" int exoblas_stbmv_row_major_Lower_Trans_stride_1_async(hipStream_t stream, int_fast32_t n, int_fast32_t k, struct exo_win_1f32 x, struct exo_win_2f32c A, int_fast32_t Diag) {
  exoblas::Call call("exoblas_stbmv_row_major_Lower_Trans_stride_1_async", exoblas::Call::kAsync, stream);
  if (n < 0) return call.fail_size("n", n);
  if (k < 0) return call.fail_size("k", k);
  if (Diag < 0) return call.fail_size("Diag", Diag);
  if ((true) && A.strides[1] != 1) return call.fail_stride("A", A.strides[1]);
  if ((true) && n > 1 && A.strides[0] < k + 1) return call.fail_ld("A", A.strides[0]);
  call.window(0, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(float), true, true);
  call.window(1, A.data, 0, exoblas::span2_hi(n, k + 1, A.strides[0]), sizeof(float), false, true);
  if (!call.stage()) return call.status();
  if (!call.ok()) return call.status();
  call.run([&] { exoblas::ops::tbmv_row_major_Lower_Trans<float>(call.ctx(), (int64_t)n, (int64_t)k, call.dev<float>(0), (int64_t)x.strides[0], call.dev<const float>(1), (int64_t)A.strides[0], (int64_t)Diag); });
  return call.finish();
}

// src/level2/tbmv.py:111-141
extern "C" void exo_dtbmv_row_major_Lower_Trans_stride_any(void *ctxt, int_fast32_t n, int_fast32_t k, struct exo_win_1f64 x, struct exo_win_2f64c A, int_fast32_t Diag) {
  exoblas::Call call("exo_dtbmv_row_major_Lower_Trans_stride_any", exoblas::Call::kSync, ctxt);
  if (n < 0) return (void)call.fail_size("n", n);
  if (k < 0) return (void)call.fail_size("k", k);
  if (Diag < 0) return (void)call.fail_size("Diag", Diag);
  if ((true) && A.strides[1] != 1) return (void)call.fail_stride("A", A.strides[1]);
  if ((true) && n > 1 && A.strides[0] < k + 1) return (void)call.fail_ld("A", A.strides[0]);
  call.window(0, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(double), true, true);
  call.window(1, A.data, 0, exoblas::span2_hi(n, k + 1, A.strides[0]), sizeof(double), false, true);
  if (!call.stage()) return (void)call.status();
  if (!call.ok()) return (void)call.status();
  call.run([&] { exoblas::ops::tbmv_row_major_Lower_Trans<double>(call.ctx(), (int64_t)n, (int64_t)k, call.dev<double>(0), (int64_t)x.strides[0], call.dev<const double>(1), (int64_t)A.strides[0], (int64_t)Diag); });
  return (void)call.finish();
}
extern "C" int exoblas_dtbmv_row_major_Lower_Trans_stride_any_async(hipStream_t stream, int_fast32_t n, int_fast32_t k, struct exo_win_1f64 x, struct exo_win_2f64c A, int_fast32_t Diag) {
  exoblas::Call call("exoblas_dtbmv_row_major_Lower_Trans_stride_any_async", exoblas::Call::kAsync, stream);
  if (n < 0) return call.fail_size("n", n);
  if (k < 0) return call.fail_size("k", k);
  if (Diag < 0) return call.fail_size("Diag", Diag);
  if ((true) && A.strides[1] != 1) return call.fail_stride("A", A.strides[1]);
  if ((true) && n > 1 && A.strides[0] < k + 1) return call.fail_ld("A", A.strides[0]);
  call.window(0, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(double), true, true);
  call.window(1, A.data, 0, exoblas::span2_hi(n, k + 1, A.strides[0]), sizeof(double), false, true);
  if (!call.stage()) return call.status();
  if (!call.ok()) return call.status();
  call.run([&] { exoblas::ops::tbmv_row_major_Lower_Trans<double>(call.ctx(), (int64_t)n, (int64_t)k, call.dev<double>(0), (int64_t)x.strides[0], call.dev<const double>(1), (int64_t)A.strides[0], (int64_t)Diag); });
  return call.finish();
}

// src/level2/tbmv.py:111-141
extern "C" void exo_dtbmv_row_major_Lower_Trans_stride_1(void *ctxt, int_fast32_t n, int_fast32_t k, struct exo_win_1f64 x, struct exo_win_2f64c A, int_fast32_t Diag) {
  exoblas::Call call("exo_dtbmv_row_major_Lower_Trans_stride_1", exoblas::Call::kSync, ctxt);
  if (n < 0) return (void)call.fail_size("n", n);
  if (k < 0) return (void)call.fail_size("k", k);
  if (Diag < 0) return (void)call.fail_size("Diag", Diag);
  if ((true) && A.strides[1] != 1) return (void)call.fail_stride("A", A.strides[1]);
  if ((true) && n > 1 && A.strides[0] < k + 1) return (void)call.fail_ld("A", A.strides[0]);
  call.window(0, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(double), true, true);
  call.window(1, A.data, 0, exoblas::span2_hi(n, k + 1, A.strides[0]), sizeof(double), false, true);
  if (!call.stage()) return (void)call.status();
  if (!call.ok()) return (void)call.status();
  call.run([&] { exoblas::ops::tbmv_row_major_Lower_Trans<double>(call.ctx(), (int64_t)n, (int64_t)k, call.dev<double>(0), (int64_t)x.strides[0], call.dev<const double>(1), (int64_t)A.strides[0], (int64_t)Diag); });
  return (void)call.finish();
}
extern "C" int exoblas_dtbmv_row_major_Lower_Trans_stride_1_async(hipStream_t stream, int_fast32_t n, int_fast32_t k, struct exo_win_1f64 x, struct exo_win_2f64c A, int_fast32_t Diag) {
  exoblas::Call call("exoblas_dtbmv_row_major_Lower_Trans_stride_1_async", exoblas::Call::kAsync, stream);
  if (n < 0) return call.fail_size("n", n);
  if (k < 0) return call.fail_size("k", k);
  if (Diag < 0) return call.fail_size("Diag", Diag);
  if ((true) && A.strides[1] != 1) return call.fail_stride("A", A.strides[1]);
  if ((true) && n > 1 && A.strides[0] < k + 1) return call.fail_ld("A", A.strides[0]);
  call.window(0, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(double), true, true);
  call.window(1, A.data, 0, exoblas::span2_hi(n, k + 1, A.strides[0]), sizeof(double), false, true);
  if (!call.stage()) return call.status();
  if (!call.ok()) return call.status();
  call.run([&] { exoblas::ops::tbmv_row_major_Lower_Trans<double>(call.ctx(), (int64_t)n, (int64_t)k, call.dev<double>(0), (int64_t)x.strides[0], call.dev<const double>(1), (int64_t)A.strides[0], (int64_t)Diag); });
  return call.finish();
}

// src/level2/tbsv.py:12-45
extern "C" void exo_stbsv_row_major_Upper_NonTrans_stride_any(void *ctxt, int_fast32_t n, int_fast32_t k, struct exo_win_1f32 x, struct exo_win_2f32c A, int_fast32_t Diag) {
  exoblas::Call call("exo_stbsv_row_major_Upper_NonTrans_stride_any", exoblas::Call::kSync, ctxt);
  if (n < 0) return (void)call.fail_size("n", n);
  if (k < 0) return (void)call.fail_size("k", k);
  if (Diag < 0) return (void)call.fail_size("Diag", Diag);
  if ((true) && A.strides[1] != 1) return (void)call.fail_stride("A", A.strides[1]);
  if ((true) && n > 1 && A.strides[0] < k + 1) return (void)call.fail_ld("A", A.strides[0]);
  call.window(0, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(float), true, true);
  call.window(1, A.data, 0, exoblas::span2_hi(n, k + 1, A.strides[0]), sizeof(float), false, true);
  if (!call.stage()) return (void)call.status();
  if (!call.ok()) return (void)call.status();
  call.run([&] { exoblas::ops::tbsv_row_major_Upper_NonTrans<float>(call.ctx(), (int64_t)n, (int64_t)k, call.dev<float>(0), (int64_t)x.strides[0], call.dev<const float>(1), (int64_t)A.strides[0], (int64_t)Diag); });
  return (void)call.finish();
}
extern "C" int exoblas_stbsv_row_major_Upper_NonTrans_stride_any_async(hipStream_t stream, int_fast32_t n, int_fast32_t k, struct exo_win_1f32 x, struct exo_win_2f32c A, int_fast32_t Diag) {
  exoblas::Call call("exoblas_stbsv_row_major_Upper_NonTrans_stride_any_async", exoblas::Call::kAsync, stream);
  if (n < 0) return call.fail_size("n", n);
  if (k < 0) return call.fail_size("k", k);
  if (Diag < 0) return call.fail_size("Diag", Diag);
  if ((true) && A.strides[1] != 1) return call.fail_stride("A", A.strides[1]);
  if ((true) && n > 1 && A.strides[0] < k + 1) return call.fail_ld("A", A.strides[0]);
  call.window(0, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(float), true, true);
  call.window(1, A.data, 0, exoblas::span2_hi(n, k + 1, A.strides[0]), sizeof(float), false, true);
  if (!call.stage()) return call.status();
  if (!call.ok()) return call.status();
  call.run([&] { exoblas::ops::tbsv_row_major_Upper_NonTrans<float>(call.ctx(), (int64_t)n, (int64_t)k, call.dev<float>(0), (int64_t)x.strides[0], call.dev<const float>(1), (int64_t)A.strides[0], (int64_t)Diag); });
  return call.finish();
}

// src/level2/tbsv.py:12-45
extern "C" void exo_stbsv_row_major_Upper_NonTrans_stride_1(void *ctxt, int_fast32_t n, int_fast32_t k, struct exo_win_1f32 x, struct exo_win_2f32c A, int_fast32_t Diag) {
  exoblas::Call call("exo_stbsv_row_major_Upper_NonTrans_stride_1", exoblas::Call::kSync, ctxt);
  if (n < 0) return (void)call.fail_size("n", n);
  if (k < 0) return (void)call.fail_size("k", k);
  if (Diag < 0) return (void)call.fail_size("Diag", Diag);
  if ((true) && A.strides[1] != 1) return (void)call.fail_stride("A", A.strides[1]);
  if ((true) && n > 1 && A.strides[0] < k + 1) return (void)call.fail_ld("A", A.strides[0]);
  call.window(0, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(float), true, true);
  call.window(1, A.data, 0, exoblas::span2_hi(n, k + 1, A.strides[0]), sizeof(float), false, true);
  if (!call.stage()) return (void)call.status();
  if (!call.ok()) return (void)call.status();
  call.run([&] { exoblas::ops::tbsv_row_major_Upper_NonTrans<float>(call.ctx(), (int64_t)n, (int64_t)k, call.dev<float>(0), (int64_t)x.strides[0], call.dev<const float>(1), (int64_t)A.strides[0], (int64_t)Diag); });
  return (void)call.finish();
}
extern "C" int exoblas_stbsv_row_major_Upper_NonTrans_stride_1_async(hipStream_t stream, int_fast32_t n, int_fast32_t k, struct exo_win_1f32 x, struct exo_win_2f32c A, int_fast32_t Diag) {
  exoblas::Call call("exoblas_stbsv_row_major_Upper_NonTrans_stride_1_async", exoblas::Call::kAsync, stream);
  if (n < 0) return call.fail_size("n", n);
  if (k < 0) return call.fail_size("k", k);
  if (Diag < 0) return call.fail_size("Diag", Diag);
  if ((true) && A.strides[1] != 1) return call.fail_stride("A", A.strides[1]);
  if ((true) && n > 1 && A.strides[0] < k + 1) return call.fail_ld("A", A.strides[0]);
  call.window(0, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(float), true, true);
  call.window(1, A.data, 0, exoblas::span2_hi(n, k + 1, A.strides[0]), sizeof(float), false, true);
  if (!call.stage()) return call.status();
  if (!call.ok()) return call.status();
  call.run([&] { exoblas::ops::tbsv_row_major_Upper_NonTrans<float>(call.ctx(), (int64_t)n, (int64_t)k, call.dev<float>(0), (int64_t)x.strides[0], call.dev<const float>(1), (int64_t)A.strides[0], (int64_t)Diag); });
  return call.finish();
}

// src/level2/tbsv.py:12-45
extern "C" void exo_dtbsv_row_major_Upper_NonTrans_stride_any(void *ctxt, int_fast32_t n, int_fast32_t k, struct exo_win_1f64 x, struct exo_win_2f64c A, int_fast32_t Diag) {
  exoblas::Call call("exo_dtbsv_row_major_Upper_NonTrans_stride_any", exoblas::Call::kSync, ctxt);
  if (n < 0) return (void)call.fail_size("n", n);
  if (k < 0) return (void)call.fail_size("k", k);
  if (Diag < 0) return (void)call.fail_size("Diag", Diag);
  if ((true) && A.strides[1] != 1) return (void)call.fail_stride("A", A.strides[1]);
  if ((true) && n > 1 && A.strides[0] < k + 1) return (void)call.fail_ld("A", A.strides[0]);
  call.window(0, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(double), true, true);
  call.window(1, A.data, 0, exoblas::span2_hi(n, k + 1, A.strides[0]), sizeof(double), false, true);
  if (!call.stage()) return (void)call.status();
  if (!call.ok()) return (void)call.status();
  call.run([&] { exoblas::ops::tbsv_row_major_Upper_NonTrans<double>(call.ctx(), (int64_t)n, (int64_t)k, call.dev<double>(0), (int64_t)x.strides[0], call.dev<const double>(1), (int64_t)A.strides[0], (int64_t)Diag); });
  return (void)call.finish();
}
extern "C" int exoblas_dtbsv_row_major_Upper_NonTrans_stride_any_async(hipStream_t stream, int_fast32_t n, int_fast32_t k, struct exo_win_1f64 x, struct exo_win_2f64c A, int_fast32_t Diag) {
  exoblas::Call call("exoblas_dtbsv_row_major_Upper_NonTrans_stride_any_async", exoblas::Call::kAsync, stream);
  if (n < 0) return call.fail_size("n", n);
  if (k < 0) return call.fail_size("k", k);
  if (Diag < 0) return call.fail_size("Diag", Diag);
  if ((true) && A.strides[1] != 1) return call.fail_stride("A", A.strides[1]);
  if ((true) && n > 1 && A.strides[0] < k + 1) return call.fail_ld("A", A.strides[0]);
  call.window(0, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(double), true, true);
  call.window(1, A.data, 0, exoblas::span2_hi(n, k + 1, A.strides[0]), sizeof(double), false, true);
  if (!call.stage()) return call.status();
  if (!call.ok()) return call.status();
  call.run([&] { exoblas::ops::tbsv_row_major_Upper_NonTrans<double>(call.ctx(), (int64_t)n, (int64_t)k, call.dev<double>(0), (int64_t)x.strides[0], call.dev<const double>(1), (int64_t)A.strides[0], (int64_t)Diag); });
  return call.finish();
}

// src/level2/tbsv.py:12-45
extern "C" void exo_dtbsv_row_major_Upper_NonTrans_stride_1(void *ctxt, int_fast32_t n, int_fast32_t k, struct exo_win_1f64 x, struct exo_win_2f64c A, int_fast32_t Diag) {
  exoblas::Call call("exo_dtbsv_row_major_Upper_NonTrans_stride_1", exoblas::Call::kSync, ctxt);
  if (n < 0) return (void)call.fail_size("n", n);
  if (k < 0) return (void)call.fail_size("k", k);
  if (Diag < 0) return (void)call.fail_size("Diag", Diag);
  if ((true) && A.strides[1] != 1) return (void)call.fail_stride("A", A.strides[1]);
  if ((true) && n > 1 && A.strides[0] < k + 1) return (void)call.fail_ld("A", A.strides[0]);
  call.window(0, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(double), true, true);
  call.window(1, A.data, 0, exoblas::span2_hi(n, k + 1, A.strides[0]), sizeof(double), false, true);
  if (!call.stage()) return (void)call.status();
  if (!call.ok()) return (void)call.status();
  call.run([&] { exoblas::ops::tbsv_row_major_Upper_NonTrans<double>(call.ctx(), (int64_t)n, (int64_t)k, call.dev<double>(0), (int64_t)x.strides[0], call.dev<const double>(1), (int64_t)A.strides[0], (int64_t)Diag); });
  return (void)call.finish();
}
extern "C" int exoblas_dtbsv_row_major_Upper_NonTrans_stride_1_async(hipStream_t stream, int_fast32_t n, int_fast32_t k, struct exo_win_1f64 x, struct exo_win_2f64c A, int_fast32_t Diag) {
  exoblas::Call call("exoblas_dtbsv_row_major_Upper_NonTrans_stride_1_async", exoblas::Call::kAsync, stream);
  if (n < 0) return call.fail_size("n", n);
  if (k < 0) return call.fail_size("k", k);
  if (Diag < 0) return call.fail_size("Diag", Diag);
  if ((true) && A.strides[1] != 1) return call.fail_stride("A", A.strides[1]);
  if ((true) && n > 1 && A.strides[0] < k + 1) return call.fail_ld("A", A.strides[0]);
  call.window(0, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(double), true, true);
  call.window(1, A.data, 0, exoblas::span2_hi(n, k + 1, A.strides[0]), sizeof(double), false, true);
  if (!call.stage()) return call.status();
  if (!call.ok()) return call.status();
  call.run([&] { exoblas::ops::tbsv_row_major_Upper_NonTrans<double>(call.ctx(), (int64_t)n, (int64_t)k, call.dev<double>(0), (int64_t)x.strides[0], call.dev<const double>(1), (int64_t)A.strides[0], (int64_t)Diag); });
  return call.finish();
}

// src/level2/tbsv.py:48-81
extern "C" void exo_stbsv_row_major_Lower_NonTrans_stride_any(void *ctxt, int_fast32_t n, int_fast32_t k, struct exo_win_1f32 x, struct exo_win_2f32c A, int_fast32_t Diag) {
  exoblas::Call call("exo_stbsv_row_major_Lower_NonTrans_stride_any", exoblas::Call::kSync, ctxt);
  if (n < 0) return (void)call.fail_size("n", n);
  if (k < 0) return (void)call.fail_size("k", k);
  if (Diag < 0) return (void)call.fail_size("Diag", Diag);
  if ((true) && A.strides[1] != 1) return (void)call.fail_stride("A", A.strides[1]);
  if ((true) && n > 1 && A.strides[0] < k + 1) return (void)call.fail_ld("A", A.strides[0]);
  call.window(0, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(float), true, true);
  call.window(1, A.data, 0, exoblas::span2_hi(n, k + 1, A.strides[0]), sizeof(float), false, true);
  if (!call.stage()) return (void)call.status();
  if (!call.ok()) return (void)call.status();
  call.run([&] { exoblas::ops::tbsv_row_major_Lower_NonTrans<float>(call.ctx(), (int64_t)n, (int64_t)k, call.dev<float>(0), (int64_t)x.strides[0], call.dev<const float>(1), (int64_t)A.strides[0], (int64_t)Diag); });
  return (void)call.finish();
}
extern "C" int exoblas_stbsv_row_major_Lower_NonTrans_stride_any_async(hipStream_t stream, int_fast32_t n, int_fast32_t k, struct exo_win_1f32 x, struct exo_win_2f32c A, int_fast32_t Diag) {
  exoblas::Call call("exoblas_stbsv_row_major_Lower_NonTrans_stride_any_async", exoblas::Call::kAsync, stream);
  if (n < 0) return call.fail_size("n", n);
  if (k < 0) return call.fail_size("k", k);
  if (Diag < 0) return call.fail_size("Diag", Diag);
  if ((true) && A.strides[1] != 1) return call.fail_stride("A", A.strides[1]);
  if ((true) && n > 1 && A.strides[0] < k + 1) return call.fail_ld("A", A.strides[0]);
  call.window(0, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(float), true, true);
  call.window(1, A.data, 0, exoblas::span2_hi(n, k + 1, A.strides[0]), sizeof(float), false, true);
  if (!call.stage()) return call.status();
  if (!call.ok()) return call.status();
  call.run([&] { exoblas::ops::tbsv_row_major_Lower_NonTrans<float>(call.ctx(), (int64_t)n, (int64_t)k, call.dev<float>(0), (int64_t)x.strides[0], call.dev<const float>(1), (int64_t)A.strides[0], (int64_t)Diag); });
  return call.finish();
}

// src/level2/tbsv.py:48-81
extern "C" void exo_stbsv_row_major_Lower_NonTrans_stride_1(void *ctxt, int_fast32_t n, int_fast32_t k, struct exo_win_1f32 x, struct exo_win_2f32c A, int_fast32_t Diag) {
  exoblas::Call call("exo_stbsv_row_major_Lower_NonTrans_stride_1", exoblas::Call::kSync, ctxt);
  if (n < 0) return (void)call.fail_size("n", n);
  if (k < 0) return (void)call.fail_size("k", k);
  if (Diag < 0) return (void)call.fail_size("Diag", Diag);
  if ((true) && A.strides[1] != 1) return (void)call.fail_stride("A", A.strides[1]);
  if ((true) && n > 1 && A.strides[0] < k + 1) return (void)call.fail_ld("A", A.strides[0]);
  call.window(0, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(float), true, true);
  call.window(1, A.data, 0, exoblas::span2_hi(n, k + 1, A.strides[0]), sizeof(float), false, true);
  if (!call.stage()) return (void)call.status();
  if (!call.ok()) return (void)call.status();
  call.run([&] { exoblas::ops::tbsv_row_major_Lower_NonTrans<float>(call.ctx(), (int64_t)n, (int64_t)k, call.dev<float>(0), (int64_t)x.strides[0], call.dev<const float>(1), (int64_t)A.strides[0], (int64_t)Diag); });
  return (void)call.finish();
}
extern "C" int exoblas_stbsv_row_major_Lower_NonTrans_stride_1_async(hipStream_t stream, int_fast32_t n, int_fast32_t k, struct exo_win_1f32 x, struct exo_win_2f32c A, int_fast32_t Diag) {
  exoblas::Call call("exoblas_stbsv_row_major_Lower_NonTrans_stride_1_async", exoblas::Call::kAsync, stream);
  if (n < 0) return call.fail_size("n", n);
  if (k < 0) return call.fail_size("k", k);
  if (Diag < 0) return call.fail_size("Diag", Diag);
  if ((true) && A.strides[1] != 1) return call.fail_stride("A", A.strides[1]);
  if ((true) && n > 1 && A.strides[0] < k + 1) return call.fail_ld("A", A.strides[0]);
  call.window(0, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(float), true, true);
  call.window(1, A.data, 0, exoblas::span2_hi(n, k + 1, A.strides[0]), sizeof(float), false, true);
  if (!call.stage()) return call.status();
  if (!call.ok()) return call.status();
  call.run([&] { exoblas::ops::tbsv_row_major_Lower_NonTrans<float>(call.ctx(), (int64_t)n, (int64_t)k, call.dev<float>(0), (int64_t)x.strides[0], call.dev<const float>(1), (int64_t)A.strides[0], (int64_t)Diag); });
  return call.finish();
}

// src/level2/tbsv.py:48-81
extern "C" void exo_dtbsv_row_major_Lower_NonTrans_stride_any(void *ctxt, int_fast32_t n, int_fast32_t k, struct exo_win_1f64 x, struct exo_win_2f64c A, int_fast32_t Diag) {
  exoblas::Call call("exo_dtbsv_row_major_Lower_NonTrans_stride_any", exoblas::Call::kSync, ctxt);
  if (n < 0) return (void)call.fail_size("n", n);
  if (k < 0) return (void)call.fail_size("k", k);
  if (Diag < 0) return (void)call.fail_size("Diag", Diag);
  if ((true) && A.strides[1] != 1) return (void)call.fail_stride("A", A.strides[1]);
  if ((true) && n > 1 && A.strides[0] < k + 1) return (void)call.fail_ld("A", A.strides[0]);
  call.window(0, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(double), true, true);
  call.window(1, A.data, 0, exoblas::span2_hi(n, k + 1, A.strides[0]), sizeof(double), false, true);
  if (!call.stage()) return (void)call.status();
  if (!call.ok()) return (void)call.status();
  call.run([&] { exoblas::ops::tbsv_row_major_Lower_NonTrans<double>(call.ctx(), (int64_t)n, (int64_t)k, call.dev<double>(0), (int64_t)x.strides[0], call.dev<const double>(1), (int64_t)A.strides[0], (int64_t)Diag); });
  return (void)call.finish();
}
extern "C" int exoblas_dtbsv_row_major_Lower_NonTrans_stride_any_async(hipStream_t stream, int_fast32_t n, int_fast32_t k, struct exo_win_1f64 x, struct exo_win_2f64c A, int_fast32_t Diag) {
  exoblas::Call call("exoblas_dtbsv_row_major_Lower_NonTrans_stride_any_async", exoblas::Call::kAsync, stream);
  if (n < 0) return call.fail_size("n", n);
  if (k < 0) return call.fail_size("k", k);
  if (Diag < 0) return call.fail_size("Diag", Diag);
  if ((true) && A.strides[1] != 1) return call.fail_stride("A", A.strides[1]);
  if ((true) && n > 1 && A.strides[0] < k + 1) return call.fail_ld("A", A.strides[0]);
  call.window(0, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(double), true, true);
  call.window(1, A.data, 0, exoblas::span2_hi(n, k + 1, A.strides[0]), sizeof(double), false, true);
  if (!call.stage()) return call.status();
  if (!call.ok()) return call.status();
  call.run([&] { exoblas::ops::tbsv_row_major_Lower_NonTrans<double>(call.ctx(), (int64_t)n, (int64_t)k, call.dev<double>(0), (int64_t)x.strides[0], call.dev<const double>(1), (int64_t)A.strides[0], (int64_t)Diag); });
  return call.finish();
}

// src/level2/tbsv.py:48-81
extern "C" void exo_dtbsv_row_major_Lower_NonTrans_stride_1(void *ctxt, int_fast32_t n, int_fast32_t k, struct exo_win_1f64 x, struct exo_win_2f64c A, int_fast32_t Diag) {
  exoblas::Call call("exo_dtbsv_row_major_Lower_NonTrans_stride_1", exoblas::Call::kSync, ctxt);
  if (n < 0) return (void)call.fail_size("n", n);
  if (k < 0) return (void)call.fail_size("k", k);
  if (Diag < 0) return (void)call.fail_size("Diag", Diag);
  if ((true) && A.strides[1] != 1) return (void)call.fail_stride("A", A.strides[1]);
  if ((true) && n > 1 && A.strides[0] < k + 1) return (void)call.fail_ld("A", A.strides[0]);
  call.window(0, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(double), true, true);
  call.window(1, A.data, 0, exoblas::span2_hi(n, k + 1, A.strides[0]), sizeof(double), false, true);
  if (!call.stage()) return (void)call.status();
  if (!call.ok()) return (void)call.status();
  call.run([&] { exoblas::ops::tbsv_row_major_Lower_NonTrans<double>(call.ctx(), (int64_t)n, (int64_t)k, call.dev<double>(0), (int64_t)x.strides[0], call.dev<const double>(1), (int64_t)A.strides[0], (int64_t)Diag); });
  return (void)call.finish();
}
extern "C" int exoblas_dtbsv_row_major_Lower_NonTrans_stride_1_async(hipStream_t stream, int_fast32_t n, int_fast32_t k, struct exo_win_1f64 x, struct exo_win_2f64c A, int_fast32_t Diag) {
  exoblas::Call call("exoblas_dtbsv_row_major_Lower_NonTrans_stride_1_async", exoblas::Call::kAsync, stream);
  if (n < 0) return call.fail_size("n", n);
  if (k < 0) return call.fail_size("k", k);
  if (Diag < 0) return call.fail_size("Diag", Diag);
  if ((true) && A.strides[1] != 1) return call.fail_stride("A", A.strides[1]);
  if ((true) && n > 1 && A.strides[0] < k + 1) return call.fail_ld("A", A.strides[0]);
  call.window(0, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(double), true, true);
  call.window(1, A.data, 0, exoblas::span2_hi(n, k + 1, A.strides[0]), sizeof(double), false, true);
  if (!call.stage()) return call.status();
  if (!call.ok()) return call.status();
  call.run([&] { exoblas::ops::tbsv_row_major_Lower_NonTrans<double>(call.ctx(), (int64_t)n, (int64_t)k, call.dev<double>(0), (int64_t)x.strides[0], call.dev<const double>(1), (int64_t)A.strides[0], (int64_t)Diag); });
  return call.finish();
}

// src/level2/tbsv.py:84-105
extern "C" void exo_stbsv_row_major_Upper_Trans_stride_any(void *ctxt, int_fast32_t n, int_fast32_t k, struct exo_win_1f32 x, struct exo_win_2f32c A, int_fast32_t Diag) {
  exoblas::Call call("exo_stbsv_row_major_Upper_Trans_stride_any", exoblas::Call::kSync, ctxt);
  if (n < 0) return (void)call.fail_size("n", n);
  if (k < 0) return (void)call.fail_size("k", k);
  if (Diag < 0) return (void)call.fail_size("Diag", Diag);
  if ((true) && A.strides[1] != 1) return (void)call.fail_stride("A", A.strides[1]);
  if ((true) && n > 1 && A.strides[0] < k + 1) return (void)call.fail_ld("A", A.strides[0]);
  call.window(0, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(float), true, true);
  call.window(1, A.data, 0, exoblas::span2_hi(n, k + 1, A.strides[0]), sizeof(float), false, true);
  if (!call.stage()) return (void)call.status();
  if (!call.ok()) return (void)call.status();
  call.run([&] { exoblas::ops::tbsv_row_major_Upper_Trans<float>(call.ctx(), (int64_t)n, (int64_t)k, call.dev<float>(0), (int64_t)x.strides[0], call.dev<const float>(1), (int64_t)A.strides[0], (int64_t)Diag); });
  return (void)call.finish();
}
extern "C" int exoblas_stbsv_row_major_Upper_Trans_stride_any_async(hipStream_t stream, int_fast32_t n, int_fast32_t k, struct exo_win_1f32 x, struct exo_win_2f32c A, int_fast32_t Diag) {
  exoblas::Call call("exoblas_stbsv_row_major_Upper_Trans_stride_any_async", exoblas::Call::kAsync, stream);
  if (n < 0) return call.fail_size("n", n);
  if (k < 0) return call.fail_size("k", k);
  if (Diag < 0) return call.fail_size("Diag", Diag);
  if ((true) && A.strides[1] != 1) return call.fail_stride("A", A.strides[1]);
  if ((true) && n > 1 && A.strides[0] < k + 1) return call.fail_ld("A", A.strides[0]);
  call.window(0, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(float), true, true);
  call.window(1, A.data, 0, exoblas::span2_hi(n, k + 1, A.strides[0]), sizeof(float), false, true);
  if (!call.stage()) return call.status();
  if (!call.ok()) return call.status();
  call.run([&] { exoblas::ops::tbsv_row_major_Upper_Trans<float>(call.ctx(), (int64_t)n, (int64_t)k, call.dev<float>(0), (int64_t)x.strides[0], call.dev<const float>(1), (int64_t)A.strides[0], (int64_t)Diag); });
  return call.finish();
}

// src/level2/tbsv.py:84-105
extern "C" void exo_dtbsv_row_major_Upper_Trans_stride_any(void *ctxt, int_fast32_t n, int_fast32_t k, struct exo_win_1f64 x, struct exo_win_2f64c A, int_fast32_t Diag) {
  exoblas::Call call("exo_dtbsv_row_major_Upper_Trans_stride_any", exoblas::Call::kSync, ctxt);
  if (n < 0) return (void)call.fail_size("n", n);
  if (k < 0) return (void)call.fail_size("k", k);
  if (Diag < 0) return (void)call.fail_size("Diag", Diag);
  if ((true) && A.strides[1] != 1) return (void)call.fail_stride("A", A.strides[1]);
  if ((true) && n > 1 && A.strides[0] < k + 1) return (void)call.fail_ld("A", A.strides[0]);
  call.window(0, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(double), true, true);
  call.window(1, A.data, 0, exoblas::span2_hi(n, k + 1, A.strides[0]), sizeof(double), false, true);
  if (!call.stage()) return (void)call.status();
  if (!call.ok()) return (void)call.status();
  call.run([&] { exoblas::ops::tbsv_row_major_Upper_Trans<double>(call.ctx(), (int64_t)n, (int64_t)k, call.dev<double>(0), (int64_t)x.strides[0], call.dev<const double>(1), (int64_t)A.strides[0], (int64_t)Diag); });
  return (void)call.finish();
}
extern "C" int exoblas_dtbsv_row_major_Upper_Trans_stride_any_async(hipStream_t stream, int_fast32_t n, int_fast32_t k, struct exo_win_1f64 x, struct exo_win_2f64c A, int_fast32_t Diag) {
  exoblas::Call call("exoblas_dtbsv_row_major_Upper_Trans_stride_any_async", exoblas::Call::kAsync, stream);
  if (n < 0) return call.fail_size("n", n);
  if (k < 0) return call.fail_size("k", k);
  if (Diag < 0) return call.fail_size("Diag", Diag);
  if ((true) && A.strides[1] != 1) return call.fail_stride("A", A.strides[1]);
  if ((true) && n > 1 && A.strides[0] < k + 1) return call.fail_ld("A", A.strides[0]);
  call.window(0, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(double), true, true);
  call.window(1, A.data, 0, exoblas::span2_hi(n, k + 1, A.strides[0]), sizeof(double), false, true);
  if (!call.stage()) return call.status();
  if (!call.ok()) return call.status();
  call.run([&] { exoblas::ops::tbsv_row_major_Upper_Trans<double>(call.ctx(), (int64_t)n, (int64_t)k, call.dev<double>(0), (int64_t)x.strides[0], call.dev<const double>(1), (int64_t)A.strides[0], (int64_t)Diag); });
  return call.finish();
}

// src/level2/tbsv.py:108-129
extern "C" void exo_stbsv_row_major_Lower_Trans_stride_any(void *ctxt, int_fast32_t n, int_fast32_t k, struct exo_win_1f32 x, struct exo_win_2f32c A, int_fast32_t Diag) {
  exoblas::Call call("exo_stbsv_row_major_Lower_Trans_stride_any", exoblas::Call::kSync, ctxt);
  if (n < 0) return (void)call.fail_size("n", n);
  if (k < 0) return (void)call.fail_size("k", k);
  if (Diag < 0) return (void)call.fail_size("Diag", Diag);
  if ((true) && A.strides[1] != 1) return (void)call.fail_stride("A", A.strides[1]);
  if ((true) && n > 1 && A.strides[0] < k + 1) return (void)call.fail_ld("A", A.strides[0]);
  call.window(0, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(float), true, true);
  call.window(1, A.data, 0, exoblas::span2_hi(n, k + 1, A.strides[0]), sizeof(float), false, true);
  if (!call.stage()) return (void)call.status();
  if (!call.ok()) return (void)call.status();
  call.run([&] { exoblas::ops::tbsv_row_major_Lower_Trans<float>(call.ctx(), (int64_t)n, (int64_t)k, call.dev<float>(0), (int64_t)x.strides[0], call.dev<const float>(1), (int64_t)A.strides[0], (int64_t)Diag); });
  return (void)call.finish();
}
extern "C" int exoblas_stbsv_row_major_Lower_Trans_stride_any_async(hipStream_t stream, int_fast32_t n, int_fast32_t k, struct exo_win_1f32 x, struct exo_win_2f32c A, int_fast32_t Diag) {
  exoblas::Call call("exoblas_stbsv_row_major_Lower_Trans_stride_any_async", exoblas::Call::kAsync, stream);
  if (n < 0) return call.fail_size("n", n);
  if (k < 0) return call.fail_size("k", k);
  if (Diag < 0) return call.fail_size("Diag", Diag);
  if ((true) && A.strides[1] != 1) return call.fail_stride("A", A.strides[1]);
  if ((true) && n > 1 && A.strides[0] < k + 1) return call.fail_ld("A", A.strides[0]);
  call.window(0, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(float), true, true);
  call.window(1, A.data, 0, exoblas::span2_hi(n, k + 1, A.strides[0]), sizeof(float), false, true);
  if (!call.stage()) return call.status();
  if (!call.ok()) return call.status();
  call.run([&] { exoblas::ops::tbsv_row_major_Lower_Trans<float>(call.ctx(), (int64_t)n, (int64_t)k, call.dev<float>(0), (int64_t)x.strides[0], call.dev<const float>(1), (int64_t)A.strides[0], (int64_t)Diag); });
  return call.finish();
}

// src/level2/tbsv.py:108-129
extern "C" void exo_dtbsv_row_major_Lower_Trans_stride_any(void *ctxt, int_fast32_t n, int_fast32_t k, struct exo_win_1f64 x, struct exo_win_2f64c A, int_fast32_t Diag) {
  exoblas::Call call("exo_dtbsv_row_major_Lower_Trans_stride_any", exoblas::Call::kSync, ctxt);
  if (n < 0) return (void)call.fail_size("n", n);
  if (k < 0) return (void)call.fail_size("k", k);
  if (Diag < 0) return (void)call.fail_size("Diag", Diag);
  if ((true) && A.strides[1] != 1) return (void)call.fail_stride("A", A.strides[1]);
  if ((true) && n > 1 && A.strides[0] < k + 1) return (void)call.fail_ld("A", A.strides[0]);
  call.window(0, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(double), true, true);
  call.window(1, A.data, 0, exoblas::span2_hi(n, k + 1, A.strides[0]), sizeof(double), false, true);
  if (!call.stage()) return (void)call.status();
  if (!call.ok()) return (void)call.status();
  call.run([&] { exoblas::ops::tbsv_row_major_Lower_Trans<double>(call.ctx(), (int64_t)n, (int64_t)k, call.dev<double>(0), (int64_t)x.strides[0], call.dev<const double>(1), (int64_t)A.strides[0], (int64_t)Diag); });
  return (void)call.finish();
}
extern "C" int exoblas_dtbsv_row_major_Lower_Trans_stride_any_async(hipStream_t stream, int_fast32_t n, int_fast32_t k, struct exo_win_1f64 x, struct exo_win_2f64c A, int_fast32_t Diag) {
  exoblas::Call call("exoblas_dtbsv_row_major_Lower_Trans_stride_any_async", exoblas::Call::kAsync, stream);
  if (n < 0) return call.fail_size("n", n);
  if (k < 0) return call.fail_size("k", k);
  if (Diag < 0) return call.fail_size("Diag", Diag);
  if ((true) && A.strides[1] != 1) return call.fail_stride("A", A.strides[1]);
  if ((true) && n > 1 && A.strides[0] < k + 1) return call.fail_ld("A", A.strides[0]);
  call.window(0, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(double), true, true);
  call.window(1, A.data, 0, exoblas::span2_hi(n, k + 1, A.strides[0]), sizeof(double), false, true);
  if (!call.stage()) return call.status();
  if (!call.ok()) return call.status();
  call.run([&] { exoblas::ops::tbsv_row_major_Lower_Trans<double>(call.ctx(), (int64_t)n, (int64_t)k, call.dev<double>(0), (int64_t)x.strides[0], call.dev<const double>(1), (int64_t)A.strides[0], (int64_t)Diag); });
  return call.finish();
}

// src/level3/gemm.py:17-47
extern "C" void exo_sgemm_stride_any(void *ctxt, int_fast32_t TransA, int_fast32_t TransB, int_fast32_t M, int_fast32_t N, int_fast32_t K, const float *alpha, struct exo_win_2f32c A, struct exo_win_2f32c B, struct exo_win_2f32c AT, struct exo_win_2f32c BT, struct exo_win_2f32 C) {
  exoblas::Call call("exo_sgemm_stride_any", exoblas::Call::kSync, ctxt);
  if (!(TransA == 111 || TransA == 112 || TransA == 113)) return (void)call.fail_enum("TransA", TransA);
  if (!(TransB == 111 || TransB == 112 || TransB == 113)) return (void)call.fail_enum("TransB", TransB);
  if (M < 0) return (void)call.fail_size("M", M);
  if (N < 0) return (void)call.fail_size("N", N);
  if (K < 0) return (void)call.fail_size("K", K);
  if ((TransA == 111) && A.strides[1] != 1) return (void)call.fail_stride("A", A.strides[1]);
  if ((TransA == 111) && M > 1 && A.strides[0] < K) return (void)call.fail_ld("A", A.strides[0]);
  if ((TransB == 111) && B.strides[1] != 1) return (void)call.fail_stride("B", B.strides[1]);
  if ((TransB == 111) && K > 1 && B.strides[0] < N) return (void)call.fail_ld("B", B.strides[0]);
  if ((TransA != 111) && AT.strides[1] != 1) return (void)call.fail_stride("AT", AT.strides[1]);
  if ((TransA != 111) && K > 1 && AT.strides[0] < M) return (void)call.fail_ld("AT", AT.strides[0]);
  if ((TransB != 111) && BT.strides[1] != 1) return (void)call.fail_stride("BT", BT.strides[1]);
  if ((TransB != 111) && N > 1 && BT.strides[0] < K) return (void)call.fail_ld("BT", BT.strides[0]);
  if ((true) && C.strides[1] != 1) return (void)call.fail_stride("C", C.strides[1]);
  if ((true) && M > 1 && C.strides[0] < N) return (void)call.fail_ld("C", C.strides[0]);
  call.window(0, A.data, 0, exoblas::span2_hi(M, K, A.strides[0]), sizeof(float), false, TransA == 111);
  call.window(1, B.data, 0, exoblas::span2_hi(K, N, B.strides[0]), sizeof(float), false, TransB == 111);
  call.window(2, AT.data, 0, exoblas::span2_hi(K, M, AT.strides[0]), sizeof(float), false, TransA != 111);
  call.window(3, BT.data, 0, exoblas::span2_hi(N, K, BT.strides[0]), sizeof(float), false, TransB != 111);
  call.window(4, C.data, 0, exoblas::span2_hi(M, N, C.strides[0]), sizeof(float), true, true);
  if (!call.stage()) return (void)call.status();
  const float alpha_v = call.scalar(alpha);
  if (!call.ok()) return (void)call.status();
  call.run([&] { exoblas::ops::gemm<float>(call.ctx(), (int64_t)TransA, (int64_t)TransB, (int64_t)M, (int64_t)N, (int64_t)K, alpha_v, call.dev<const float>(0), (int64_t)A.strides[0], call.dev<const float>(1), (int64_t)B.strides[0], call.dev<const float>(2), (int64_t)AT.strides[0], call.dev<const float>(3), (int64_t)BT.strides[0], call.dev<float>(4), (int64_t)C.strides[0]); });
  return (void)call.finish();
}
extern "C" int exoblas_sgemm_stride_any_async(hipStream_t stream, int_fast32_t TransA, int_fast32_t TransB, int_fast32_t M, int_fast32_t N, int_fast32_t K, const float *alpha, struct exo_win_2f32c A, struct exo_win_2f32c B, struct exo_win_2f32c AT, struct exo_win_2f32c BT, struct exo_win_2f32 C) {
  exoblas::Call call("exoblas_sgemm_stride_any_async", exoblas::Call::kAsync, stream);
  if (!(TransA == 111 || TransA == 112 || TransA == 113)) return call.fail_enum("TransA", TransA);
  if (!(TransB == 111 || TransB == 112 || TransB == 113)) return call.fail_enum("TransB", TransB);
  if (M < 0) return call.fail_size("M", M);
  if (N < 0) return call.fail_size("N", N);
  if (K < 0) return call.fail_size("K", K);
  if ((TransA == 111) && A.strides[1] != 1) return call.fail_stride("A", A.strides[1]);
  if ((TransA == 111) && M > 1 && A.strides[0] < K) return call.fail_ld("A", A.strides[0]);
  if ((TransB == 111) && B.strides[1] != 1) return call.fail_stride("B", B.strides[1]);
  if ((TransB == 111) && K > 1 && B.strides[0] < N) return call.fail_ld("B", B.strides[0]);
  if ((TransA != 111) && AT.strides[1] != 1) return call.fail_stride("AT", AT.strides[1]);
  if ((TransA != 111) && K > 1 && AT.strides[0] < M) return call.fail_ld("AT", AT.strides[0]);
  if ((TransB != 111) && BT.strides[1] != 1) return call.fail_stride("BT", BT.strides[1]);
  if ((TransB != 111) && N > 1 && BT.strides[0] < K) return call.fail_ld("BT", BT.strides[0]);
  if ((true) && C.strides[1] != 1) return call.fail_stride("C", C.strides[1]);
  if ((true) && M > 1 && C.strides[0] < N) return call.fail_ld("C", C.strides[0]);
  call.window(0, A.data, 0, exoblas::span2_hi(M, K, A.strides[0]), sizeof(float), false, TransA == 111);
  call.window(1, B.data, 0, exoblas::span2_hi(K, N, B.strides[0]), sizeof(float), false, TransB == 111);
  call.window(2, AT.data, 0, exoblas::span2_hi(K, M, AT.strides[0]), sizeof(float), false, TransA != 111);
  call.window(3, BT.data, 0, exoblas::span2_hi(N, K, BT.strides[0]), sizeof(float), false, TransB != 111);
  call.window(4, C.data, 0, exoblas::span2_hi(M, N, C.strides[0]), sizeof(float), true, true);
  if (!call.stage()) return call.status();
  const float alpha_v = call.scalar(alpha);
  if (!call.ok()) return call.status();
  call.run([&] { exoblas::ops::gemm<float>(call.ctx(), (int64_t)TransA, (int64_t)TransB, (int64_t)M, (int64_t)N, (int64_t)K, alpha_v, call.dev<const float>(0), (int64_t)A.strides[0], call.dev<const float>(1), (int64_t)B.strides[0], call.dev<const float>(2), (int64_t)AT.strides[0], call.dev<const float>(3), (int64_t)BT.strides[0], call.dev<float>(4), (int64_t)C.strides[0]); });
  return call.finish();
}

// src/level3/gemm.py:17-47
extern "C" void exo_sgemm_stride_1(void *ctxt, int_fast32_t TransA, int_fast32_t TransB, int_fast32_t M, int_fast32_t N, int_fast32_t K, const float *alpha, struct exo_win_2f32c A, struct exo_win_2f32c B, struct exo_win_2f32c AT, struct exo_win_2f32c BT, struct exo_win_2f32 C) {
  exoblas::Call call("exo_sgemm_stride_1", exoblas::Call::kSync, ctxt);
  if (!(TransA == 111 || TransA == 112 || TransA == 113)) return (void)call.fail_enum("TransA", TransA);
  if (!(TransB == 111 || TransB == 112 || TransB == 113)) return (void)call.fail_enum("TransB", TransB);
  if (M < 0) return (void)call.fail_size("M", M);
  if (N < 0) return (void)call.fail_size("N", N);
  if (K < 0) return (void)call.fail_size("K", K);
  if ((TransA == 111) && A.strides[1] != 1) return (void)call.fail_stride("A", A.strides[1]);
  if ((TransA == 111) && M > 1 && A.strides[0] < K) return (void)call.fail_ld("A", A.strides[0]);
  if ((TransB == 111) && B.strides[1] != 1) return (void)call.fail_stride("B", B.strides[1]);
  if ((TransB == 111) && K > 1 && B.strides[0] < N) return (void)call.fail_ld("B", B.strides[0]);
  if ((TransA != 111) && AT.strides[1] != 1) return (void)call.fail_stride("AT", AT.strides[1]);
  if ((TransA != 111) && K > 1 && AT.strides[0] < M) return (void)call.fail_ld("AT", AT.strides[0]);
  if ((TransB != 111) && BT.strides[1] != 1) return (void)call.fail_stride("BT", BT.strides[1]);
  if ((TransB != 111) && N > 1 && BT.strides[0] < K) return (void)call.fail_ld("BT", BT.strides[0]);
  if ((true) && C.strides[1] != 1) return (void)call.fail_stride("C", C.strides[1]);
  if ((true) && M > 1 && C.strides[0] < N) return (void)call.fail_ld("C", C.strides[0]);
  call.window(0, A.data, 0, exoblas::span2_hi(M, K, A.strides[0]), sizeof(float), false, TransA == 111);
  call.window(1, B.data, 0, exoblas::span2_hi(K, N, B.strides[0]), sizeof(float), false, TransB == 111);
  call.window(2, AT.data, 0, exoblas::span2_hi(K, M, AT.strides[0]), sizeof(float), false, TransA != 111);
  call.window(3, BT.data, 0, exoblas::span2_hi(N, K, BT.strides[0]), sizeof(float), false, TransB != 111);
  call.window(4, C.data, 0, exoblas::span2_hi(M, N, C.strides[0]), sizeof(float), true, true);
  if (!call.stage()) return (void)call.status();
  const float alpha_v = call.scalar(alpha);
  if (!call.ok()) return (void)call.status();
  call.run([&] { exoblas::ops::gemm<float>(call.ctx(), (int64_t)TransA, (int64_t)TransB, (int64_t)M, (int64_t)N, (int64_t)K, alpha_v, call.dev<const float>(0), (int64_t)A.strides[0], call.dev<const float>(1), (int64_t)B.strides[0], call.dev<const float>(2), (int64_t)AT.strides[0], call.dev<const float>(3), (int64_t)BT.strides[0], call.dev<float>(4), (int64_t)C.strides[0]); });
  return (void)call.finish();
}
extern "C" int exoblas_sgemm_stride_1_async(hipStream_t stream, int_fast32_t TransA, int_fast32_t TransB, int_fast32_t M, int_fast32_t N, int_fast32_t K, const float *alpha, struct exo_win_2f32c A, struct exo_win_2f32c B, struct exo_win_2f32c AT, struct exo_win_2f32c BT, struct exo_win_2f32 C) {
  exoblas::Call call("exoblas_sgemm_stride_1_async", exoblas::Call::kAsync, stream);
  if (!(TransA == 111 || TransA == 112 || TransA == 113)) return call.fail_enum("TransA", TransA);
  if (!(TransB == 111 || TransB == 112 || TransB == 113)) return call.fail_enum("TransB", TransB);
  if (M < 0) return call.fail_size("M", M);
  if (N < 0) return call.fail_size("N", N);
  if (K < 0) return call.fail_size("K", K);
  if ((TransA == 111) && A.strides[1] != 1) return call.fail_stride("A", A.strides[1]);
  if ((TransA == 111) && M > 1 && A.strides[0] < K) return call.fail_ld("A", A.strides[0]);
  if ((TransB == 111) && B.strides[1] != 1) return call.fail_stride("B", B.strides[1]);
  if ((TransB == 111) && K > 1 && B.strides[0] < N) return call.fail_ld("B", B.strides[0]);
  if ((TransA != 111) && AT.strides[1] != 1) return call.fail_stride("AT", AT.strides[1]);
  if ((TransA != 111) && K > 1 && AT.strides[0] < M) return call.fail_ld("AT", AT.strides[0]);
  if ((TransB != 111) && BT.strides[1] != 1) return call.fail_stride("BT", BT.strides[1]);
  if ((TransB != 111) && N > 1 && BT.strides[0] < K) return call.fail_ld("BT", BT.strides[0]);
  if ((true) && C.strides[1] != 1) return call.fail_stride("C", C.strides[1]);
  if ((true) && M > 1 && C.strides[0] < N) return call.fail_ld("C", C.strides[0]);
  call.window(0, A.data, 0, exoblas::span2_hi(M, K, A.strides[0]), sizeof(float), false, TransA == 111);
  call.window(1, B.data, 0, exoblas::span2_hi(K, N, B.strides[0]), sizeof(float), false, TransB == 111);
  call.window(2, AT.data, 0, exoblas::span2_hi(K, M, AT.strides[0]), sizeof(float), false, TransA != 111);
  call.window(3, BT.data, 0, exoblas::span2_hi(N, K, BT.strides[0]), sizeof(float), false, TransB != 111);
  call.window(4, C.data, 0, exoblas::span2_hi(M, N, C.strides[0]), sizeof(float), true, true);
  if (!call.stage()) return call.status();
  const float alpha_v = call.scalar(alpha);
  if (!call.ok()) return call.status();
  call.run([&] { exoblas::ops::gemm<float>(call.ctx(), (int64_t)TransA, (int64_t)TransB, (int64_t)M, (int64_t)N, (int64_t)K, alpha_v, call.dev<const float>(0), (int64_t)A.strides[0], call.dev<const float>(1), (int64_t)B.strides[0], call.dev<const float>(2), (int64_t)AT.strides[0], call.dev<const float>(3), (int64_t)BT.strides[0], call.dev<float>(4), (int64_t)C.strides[0]); });
  return call.finish();
}

// src/level3/gemm.py:17-47
extern "C" void exo_dgemm_stride_any(void *ctxt, int_fast32_t TransA, int_fast32_t TransB, int_fast32_t M, int_fast32_t N, int_fast32_t K, const double *alpha, struct exo_win_2f64c A, struct exo_win_2f64c B, struct exo_win_2f64c AT, struct exo_win_2f64c BT, struct exo_win_2f64 C) {
  exoblas::Call call("exo_dgemm_stride_any", exoblas::Call::kSync, ctxt);
  if (!(TransA == 111 || TransA == 112 || TransA == 113)) return (void)call.fail_enum("TransA", TransA);
  if (!(TransB == 111 || TransB == 112 || TransB == 113)) return (void)call.fail_enum("TransB", TransB);
  if (M < 0) return (void)call.fail_size("M", M);
  if (N < 0) return (void)call.fail_size("N", N);
  if (K < 0) return (void)call.fail_size("K", K);
  if ((TransA == 111) && A.strides[1] != 1) return (void)call.fail_stride("A", A.strides[1]);
  if ((TransA == 111) && M > 1 && A.strides[0] < K) return (void)call.fail_ld("A", A.strides[0]);
  if ((TransB == 111) && B.strides[1] != 1) return (void)call.fail_stride("B", B.strides[1]);
  if ((TransB == 111) && K > 1 && B.strides[0] < N) return (void)call.fail_ld("B", B.strides[0]);
  if ((TransA != 111) && AT.strides[1] != 1) return (void)call.fail_stride("AT", AT.strides[1]);
  if ((TransA != 111) && K > 1 && AT.strides[0] < M) return (void)call.fail_ld("AT", AT.strides[0]);
  if ((TransB != 111) && BT.strides[1] != 1) return (void)call.fail_stride("BT", BT.strides[1]);
  if ((TransB != 111) && N > 1 && BT.strides[0] < K) return (void)call.fail_ld("BT", BT.strides[0]);
  if ((true) && C.strides[1] != 1) return (void)call.fail_stride("C", C.strides[1]);
  if ((true) && M > 1 && C.strides[0] < N) return (void)call.fail_ld("C", C.strides[0]);
  call.window(0, A.data, 0, exoblas::span2_hi(M, K, A.strides[0]), sizeof(double), false, TransA == 111);
  call.window(1, B.data, 0, exoblas::span2_hi(K, N, B.strides[0]), sizeof(double), false, TransB == 111);
  call.window(2, AT.data, 0, exoblas::span2_hi(K, M, AT.strides[0]), sizeof(double), false, TransA != 111);
  call.window(3, BT.data, 0, exoblas::span2_hi(N, K, BT.strides[0]), sizeof(double), false, TransB != 111);
  call.window(4, C.data, 0, exoblas::span2_hi(M, N, C.strides[0]), sizeof(double), true, true);
  if (!call.stage()) return (void)call.status();
  const double alpha_v = call.scalar(alpha);
  if (!call.ok()) return (void)call.status();
  call.run([&] { exoblas::ops::gemm<double>(call.ctx(), (int64_t)TransA, (int64_t)TransB, (int64_t)M, (int64_t)N, (int64_t)K, alpha_v, call.dev<const double>(0), (int64_t)A.strides[0], call.dev<const double>(1), (int64_t)B.strides[0], call.dev<const double>(2), (int64_t)AT.strides[0], call.dev<const double>(3), (int64_t)BT.strides[0], call.dev<double>(4), (int64_t)C.strides[0]); });
  return (void)call.finish();
}
extern "C" int exoblas_dgemm_stride_any_async(hipStream_t stream, int_fast32_t TransA, int_fast32_t TransB, int_fast32_t M, int_fast32_t N, int_fast32_t K, const double *alpha, struct exo_win_2f64c A, struct exo_win_2f64c B, struct exo_win_2f64c AT, struct exo_win_2f64c BT, struct exo_win_2f64 C) {
  exoblas::Call call("exoblas_dgemm_stride_any_async", exoblas::Call::kAsync, stream);
  if (!(TransA == 111 || TransA == 112 || TransA == 113)) return call.fail_enum("TransA", TransA);
  if (!(TransB == 111 || TransB == 112 || TransB == 113)) return call.fail_enum("TransB", TransB);
  if (M < 0) return call.fail_size("M", M);
  if (N < 0) return call.fail_size("N", N);
  if (K < 0) return call.fail_size("K", K);
  if ((TransA == 111) && A.strides[1] != 1) return call.fail_stride("A", A.strides[1]);
  if ((TransA == 111) && M > 1 && A.strides[0] < K) return call.fail_ld("A", A.strides[0]);
  if ((TransB == 111) && B.strides[1] != 1) return call.fail_stride("B", B.strides[1]);
  if ((TransB == 111) && K > 1 && B.strides[0] < N) return call.fail_ld("B", B.strides[0]);
  if ((TransA != 111) && AT.strides[1] != 1) return call.fail_stride("AT", AT.strides[1]);
  if ((TransA != 111) && K > 1 && AT.strides[0] < M) return call.fail_ld("AT", AT.strides[0]);
  if ((TransB != 111) && BT.strides[1] != 1) return call.fail_stride("BT", BT.strides[1]);
  if ((TransB != 111) && N > 1 && BT.strides[0] < K) return call.fail_ld("BT", BT.strides[0]);
  if ((true) && C.strides[1] != 1) return call.fail_stride("C", C.strides[1]);
  if ((true) && M > 1 && C.strides[0] < N) return call.fail_ld("C", C.strides[0]);
  call.window(0, A.data, 0, exoblas::span2_hi(M, K, A.strides[0]), sizeof(double), false, TransA == 111);
  call.window(1, B.data, 0, exoblas::span2_hi(K, N, B.strides[0]), sizeof(double), false, TransB == 111);
  call.window(2, AT.data, 0, exoblas::span2_hi(K, M, AT.strides[0]), sizeof(double), false, TransA != 111);
  call.window(3, BT.data, 0, exoblas::span2_hi(N, K, BT.strides[0]), sizeof(double), false, TransB != 111);
  call.window(4, C.data, 0, exoblas::span2_hi(M, N, C.strides[0]), sizeof(double), true, true);
  if (!call.stage()) return call.status();
  const double alpha_v = call.scalar(alpha);
  if (!call.ok()) return call.status();
  call.run([&] { exoblas::ops::gemm<double>(call.ctx(), (int64_t)TransA, (int64_t)TransB, (int64_t)M, (int64_t)N, (int64_t)K, alpha_v, call.dev<const double>(0), (int64_t)A.strides[0], call.dev<const double>(1), (int64_t)B.strides[0], call.dev<const double>(2), (int64_t)AT.strides[0], call.dev<const double>(3), (int64_t)BT.strides[0], call.dev<double>(4), (int64_t)C.strides[0]); });
  return call.finish();
}

// src/level3/gemm.py:17-47
extern "C" void exo_dgemm_stride_1(void *ctxt, int_fast32_t TransA, int_fast32_t TransB, int_fast32_t M, int_fast32_t N, int_fast32_t K, const double *alpha, struct exo_win_2f64c A, struct exo_win_2f64c B, struct exo_win_2f64c AT, struct exo_win_2f64c BT, struct exo_win_2f64 C) {
  exoblas::Call call("exo_dgemm_stride_1", exoblas::Call::kSync, ctxt);
  if (!(TransA == 111 || TransA == 112 || TransA == 113)) return (void)call.fail_enum("TransA", TransA);
  if (!(TransB == 111 || TransB == 112 || TransB == 113)) return (void)call.fail_enum("TransB", TransB);
  if (M < 0) return (void)call.fail_size("M", M);
  if (N < 0) return (void)call.fail_size("N", N);
  if (K < 0) return (void)call.fail_size("K", K);
  if ((TransA == 111) && A.strides[1] != 1) return (void)call.fail_stride("A", A.strides[1]);
  if ((TransA == 111) && M > 1 && A.strides[0] < K) return (void)call.fail_ld("A", A.strides[0]);
  if ((TransB == 111) && B.strides[1] != 1) return (void)call.fail_stride("B", B.strides[1]);
  if ((TransB == 111) && K > 1 && B.strides[0] < N) return (void)call.fail_ld("B", B.strides[0]);
  if ((TransA != 111) && AT.strides[1] != 1) return (void)call.fail_stride("AT", AT.strides[1]);
  if ((TransA != 111) && K > 1 && AT.strides[0] < M) return (void)call.fail_ld("AT", AT.strides[0]);
  if ((TransB != 111) && BT.strides[1] != 1) return (void)call.fail_stride("BT", BT.strides[1]);
  if ((TransB != 111) && N > 1 && BT.strides[0] < K) return (void)call.fail_ld("BT", BT.strides[0]);
  if ((true) && C.strides[1] != 1) return (void)call.fail_stride("C", C.strides[1]);
  if ((true) && M > 1 && C.strides[0] < N) return (void)call.fail_ld("C", C.strides[0]);
  call.window(0, A.data, 0, exoblas::span2_hi(M, K, A.strides[0]), sizeof(double), false, TransA == 111);
  call.window(1, B.data, 0, exoblas::span2_hi(K, N, B.strides[0]), sizeof(double), false, TransB == 111);
  call.window(2, AT.data, 0, exoblas::span2_hi(K, M, AT.strides[0]), sizeof(double), false, TransA != 111);
  call.window(3, BT.data, 0, exoblas::span2_hi(N, K, BT.strides[0]), sizeof(double), false, TransB != 111);
  call.window(4, C.data, 0, exoblas::span2_hi(M, N, C.strides[0]), sizeof(double), true, true);
  if (!call.stage()) return (void)call.status();
  const double alpha_v = call.scalar(alpha);
  if (!call.ok()) return (void)call.status();
  call.run([&] { exoblas::ops::gemm<double>(call.ctx(), (int64_t)TransA, (int64_t)TransB, (int64_t)M, (int64_t)N, (int64_t)K, alpha_v, call.dev<const double>(0), (int64_t)A.strides[0], call.dev<const double>(1), (int64_t)B.strides[0], call.dev<const double>(2), (int64_t)AT.strides[0], call.dev<const double>(3), (int64_t)BT.strides[0], call.dev<double>(4), (int64_t)C.strides[0]); });
  return (void)call.finish();
}
extern "C" int exoblas_dgemm_stride_1_async(hipStream_t stream, int_fast32_t TransA, int_fast32_t TransB, int_fast32_t M, int_fast32_t N, int_fast32_t K, const double *alpha, struct exo_win_2f64c A, struct exo_win_2f64c B, struct exo_win_2f64c AT, struct exo_win_2f64c BT, struct exo_win_2f64 C) {
  exoblas::Call call("exoblas_dgemm_stride_1_async", exoblas::Call::kAsync, stream);
  if (!(TransA == 111 || TransA == 112 || TransA == 113)) return call.fail_enum("TransA", TransA);
  if (!(TransB == 111 || TransB == 112 || TransB == 113)) return call.fail_enum("TransB", TransB);
  if (M < 0) return call.fail_size("M", M);
  if (N < 0) return call.fail_size("N", N);
  if (K < 0) return call.fail_size("K", K);
  if ((TransA == 111) && A.strides[1] != 1) return call.fail_stride("A", A.strides[1]);
  if ((TransA == 111) && M > 1 && A.strides[0] < K) return call.fail_ld("A", A.strides[0]);
  if ((TransB == 111) && B.strides[1] != 1) return call.fail_stride("B", B.strides[1]);
  if ((TransB == 111) && K > 1 && B.strides[0] < N) return call.fail_ld("B", B.strides[0]);
  if ((TransA != 111) && AT.strides[1] != 1) return call.fail_stride("AT", AT.strides[1]);
  if ((TransA != 111) && K > 1 && AT.strides[0] < M) return call.fail_ld("AT", AT.strides[0]);
  if ((TransB != 111) && BT.strides[1] != 1) return call.fail_stride("BT", BT.strides[1]);
  if ((TransB != 111) && N > 1 && BT.strides[0] < K) return call.fail_ld("BT", BT.strides[0]);
  if ((true) && C.strides[1] != 1) return call.fail_stride("C", C.strides[1]);
  if ((true) && M > 1 && C.strides[0] < N) return call.fail_ld("C", C.strides[0]);
  call.window(0, A.data, 0, exoblas::span2_hi(M, K, A.strides[0]), sizeof(double), false, TransA == 111);
  call.window(1, B.data, 0, exoblas::span2_hi(K, N, B.strides[0]), sizeof(double), false, TransB == 111);
  call.window(2, AT.data, 0, exoblas::span2_hi(K, M, AT.strides[0]), sizeof(double), false, TransA != 111);
  call.window(3, BT.data, 0, exoblas::span2_hi(N, K, BT.strides[0]), sizeof(double), false, TransB != 111);
  call.window(4, C.data, 0, exoblas::span2_hi(M, N, C.strides[0]), sizeof(double), true, true);
  if (!call.stage()) return call.status();
  const double alpha_v = call.scalar(alpha);
  if (!call.ok()) return call.status();
  call.run([&] { exoblas::ops::gemm<double>(call.ctx(), (int64_t)TransA, (int64_t)TransB, (int64_t)M, (int64_t)N, (int64_t)K, alpha_v, call.dev<const double>(0), (int64_t)A.strides[0], call.dev<const double>(1), (int64_t)B.strides[0], call.dev<const double>(2), (int64_t)AT.strides[0], call.dev<const double>(3), (int64_t)BT.strides[0], call.dev<double>(4), (int64_t)C.strides[0]); });
  return call.finish();
}

// src/level3/mscal.py:12-17
extern "C" void exo_smscal_rm_stride_any(void *ctxt, int_fast32_t M, int_fast32_t N, const float *alpha, struct exo_win_2f32 A) {
  exoblas::Call call("exo_smscal_rm_stride_any", exoblas::Call::kSync, ctxt);
  if (M < 0) return (void)call.fail_size("M", M);
  if (N < 0) return (void)call.fail_size("N", N);
  if ((true) && A.strides[1] != 1) return (void)call.fail_stride("A", A.strides[1]);
  if ((true) && M > 1 && A.strides[0] < N) return (void)call.fail_ld("A", A.strides[0]);
  call.window(0, A.data, 0, exoblas::span2_hi(M, N, A.strides[0]), sizeof(float), true, true);
  if (!call.stage()) return (void)call.status();
  const float alpha_v = call.scalar(alpha);
  if (!call.ok()) return (void)call.status();
  call.run([&] { exoblas::ops::mscal_rm<float>(call.ctx(), (int64_t)M, (int64_t)N, alpha_v, call.dev<float>(0), (int64_t)A.strides[0]); });
  return (void)call.finish();
}
extern "C" int exoblas_smscal_rm_stride_any_async(hipStream_t stream, int_fast32_t M, int_fast32_t N, const float *alpha, struct exo_win_2f32 A) {
  exoblas::Call call("exoblas_smscal_rm_stride_any_async", exoblas::Call::kAsync, stream);
  if (M < 0) return call.fail_size("M", M);
  if (N < 0) return call.fail_size("N", N);
  if ((true) && A.strides[1] != 1) return call.fail_stride("A", A.strides[1]);
  if ((true) && M > 1 && A.strides[0] < N) return call.fail_ld("A", A.strides[0]);
  call.window(0, A.data, 0, exoblas::span2_hi(M, N, A.strides[0]), sizeof(float), true, true);
  if (!call.stage()) return call.status();
  const float alpha_v = call.scalar(alpha);
  if (!call.ok()) return call.status();
  call.run([&] { exoblas::ops::mscal_rm<float>(call.ctx(), (int64_t)M, (int64_t)N, alpha_v, call.dev<float>(0), (int64_t)A.strides[0]); });
  return call.finish();
}

// src/level3/mscal.py:12-17
extern "C" void exo_smscal_rm_stride_1(void *ctxt, int_fast32_t M, int_fast32_t N, const float *alpha, struct exo_win_2f32 A) {
  exoblas::Call call("exo_smscal_rm_stride_1", exoblas::Call::kSync, ctxt);
  if (M < 0) return (void)call.fail_size("M", M);
  if (N < 0) return (void)call.fail_size("N", N);
  if ((true) && A.strides[1] != 1) return (void)call.fail_stride("A", A.strides[1]);
  if ((true) && M > 1 && A.strides[0] < N) return (void)call.fail_ld("A", A.strides[0]);
  call.window(0, A.data, 0, exoblas::span2_hi(M, N, A.strides[0]), sizeof(float), true, true);
  if (!call.stage()) return (void)call.status();
  const float alpha_v = call.scalar(alpha);
  if (!call.ok()) return (void)call.status();
  call.run([&] { exoblas::ops::mscal_rm<float>(call.ctx(), (int64_t)M, (int64_t)N, alpha_v, call.dev<float>(0), (int64_t)A.strides[0]); });
  return (void)call.finish();
}
extern "C" int exoblas_smscal_rm_stride_1_async(hipStream_t stream, int_fast32_t M, int_fast32_t N, const float *alpha, struct exo_win_2f32 A) {
  exoblas::Call call("exoblas_smscal_rm_stride_1_async", exoblas::Call::kAsync, stream);
  if (M < 0) return call.fail_size("M", M);
  if (N < 0) return call.fail_size("N", N);
  if ((true) && A.strides[1] != 1) return call.fail_stride("A", A.strides[1]);
  if ((true) && M > 1 && A.strides[0] < N) return call.fail_ld("A", A.strides[0]);
  call.window(0, A.data, 0, exoblas::span2_hi(M, N, A.strides[0]), sizeof(float), true, true);
  if (!call.stage()) return call.status();
  const float alpha_v = call.scalar(alpha);
  if (!call.ok()) return call.status();
  call.run([&] { exoblas::ops::mscal_rm<float>(call.ctx(), (int64_t)M, (int64_t)N, alpha_v, call.dev<float>(0), (int64_t)A.strides[0]); });
  return call.finish();
}

// src/level3/mscal.py:12-17
extern "C" void exo_dmscal_rm_stride_any(void *ctxt, int_fast32_t M, int_fast32_t N, const double *alpha, struct exo_win_2f64 A) {
  exoblas::Call call("exo_dmscal_rm_stride_any", exoblas::Call::kSync, ctxt);
  if (M < 0) return (void)call.fail_size("M", M);
  if (N < 0) return (void)call.fail_size("N", N);
  if ((true) && A.strides[1] != 1) return (void)call.fail_stride("A", A.strides[1]);
  if ((true) && M > 1 && A.strides[0] < N) return (void)call.fail_ld("A", A.strides[0]);
  call.window(0, A.data, 0, exoblas::span2_hi(M, N, A.strides[0]), sizeof(double), true, true);
  if (!call.stage()) return (void)call.status();
  const double alpha_v = call.scalar(alpha);
  if (!call.ok()) return (void)call.status();
  call.run([&] { exoblas::ops::mscal_rm<double>(call.ctx(), (int64_t)M, (int64_t)N, alpha_v, call.dev<double>(0), (int64_t)A.strides[0]); });
  return (void)call.finish();
}
extern "C" int exoblas_dmscal_rm_stride_any_async(hipStream_t stream, int_fast32_t M, int_fast32_t N, const double *alpha, struct exo_win_2f64 A) {
  exoblas::Call call("exoblas_dmscal_rm_stride_any_async", exoblas::Call::kAsync, stream);
  if (M < 0) return call.fail_size("M", M);
  if (N < 0) return call.fail_size("N", N);
  if ((true) && A.strides[1] != 1) return call.fail_stride("A", A.strides[1]);
  if ((true) && M > 1 && A.strides[0] < N) return call.fail_ld("A", A.strides[0]);
  call.window(0, A.data, 0, exoblas::span2_hi(M, N, A.strides[0]), sizeof(double), true, true);
  if (!call.stage()) return call.status();
  const double alpha_v = call.scalar(alpha);
  if (!call.ok()) return call.status();
  call.run([&] { exoblas::ops::mscal_rm<double>(call.ctx(), (int64_t)M, (int64_t)N, alpha_v, call.dev<double>(0), (int64_t)A.strides[0]); });
  return call.finish();
}

// src/level3/mscal.py:12-17
extern "C" void exo_dmscal_rm_stride_1(void *ctxt, int_fast32_t M, int_fast32_t N, const double *alpha, struct exo_win_2f64 A) {
  exoblas::Call call("exo_dmscal_rm_stride_1", exoblas::Call::kSync, ctxt);
  if (M < 0) return (void)call.fail_size("M", M);
  if (N < 0) return (void)call.fail_size("N", N);
  if ((true) && A.strides[1] != 1) return (void)call.fail_stride("A", A.strides[1]);
  if ((true) && M > 1 && A.strides[0] < N) return (void)call.fail_ld("A", A.strides[0]);
  call.window(0, A.data, 0, exoblas::span2_hi(M, N, A.strides[0]), sizeof(double), true, true);
  if (!call.stage()) return (void)call.status();
  const double alpha_v = call.scalar(alpha);
  if (!call.ok()) return (void)call.status();
  call.run([&] { exoblas::ops::mscal_rm<double>(call.ctx(), (int64_t)M, (int64_t)N, alpha_v, call.dev<double>(0), (int64_t)A.strides[0]); });
  return (void)call.finish();
}
extern "C" int exoblas_dmscal_rm_stride_1_async(hipStream_t stream, int_fast32_t M, int_fast32_t N, const double *alpha, struct exo_win_2f64 A) {
  exoblas::Call call("exoblas_dmscal_rm_stride_1_async", exoblas::Call::kAsync, stream);
  if (M < 0) return call.fail_size("M", M);
  if (N < 0) return call.fail_size("N", N);
  if ((true) && A.strides[1] != 1) return call.fail_stride("A", A.strides[1]);
  if ((true) && M > 1 && A.strides[0] < N) return call.fail_ld("A", A.strides[0]);
  call.window(0, A.data, 0, exoblas::span2_hi(M, N, A.strides[0]), sizeof(double), true, true);
  if (!call.stage()) return call.status();
  const double alpha_v = call.scalar(alpha);
  if (!call.ok()) return call.status();
  call.run([&] { exoblas::ops::mscal_rm<double>(call.ctx(), (int64_t)M, (int64_t)N, alpha_v, call.dev<double>(0), (int64_t)A.strides[0]); });
  return call.finish();
}

// src/level3/mscal.py:20-27
extern "C" void exo_strmscal_rm_stride_any(void *ctxt, int_fast32_t Uplo, int_fast32_t N, const float *alpha, struct exo_win_2f32 A) {
  exoblas::Call call("exo_strmscal_rm_stride_any", exoblas::Call::kSync, ctxt);
  if (!(Uplo == 121 || Uplo == 122)) return (void)call.fail_enum("Uplo", Uplo);
  if (N < 0) return (void)call.fail_size("N", N);
  if ((true) && A.strides[1] != 1) return (void)call.fail_stride("A", A.strides[1]);
  if ((true) && N > 1 && A.strides[0] < N) return (void)call.fail_ld("A", A.strides[0]);
  call.window(0, A.data, 0, exoblas::span2_hi(N, N, A.strides[0]), sizeof(float), true, true);
  if (!call.stage()) return (void)call.status();
  const float alpha_v = call.scalar(alpha);
  if (!call.ok()) return (void)call.status();
  call.run([&] { exoblas::ops::trmscal_rm<float>(call.ctx(), (int64_t)Uplo, (int64_t)N, alpha_v, call.dev<float>(0), (int64_t)A.strides[0]); });
  return (void)call.finish();
}
extern "C" int exoblas_strmscal_rm_stride_any_async(hipStream_t stream, int_fast32_t Uplo, int_fast32_t N, const float *alpha, struct exo_win_2f32 A) {
  exoblas::Call call("exoblas_strmscal_rm_stride_any_async", exoblas::Call::kAsync, stream);
  if (!(Uplo == 121 || Uplo == 122)) return call.fail_enum("Uplo", Uplo);
  if (N < 0) return call.fail_size("N", N);
  if ((true) && A.strides[1] != 1) return call.fail_stride("A", A.strides[1]);
  if ((true) && N > 1 && A.strides[0] < N) return call.fail_ld("A", A.strides[0]);
  call.window(0, A.data, 0, exoblas::span2_hi(N, N, A.strides[0]), sizeof(float), true, true);
  if (!call.stage()) return call.status();
  const float alpha_v = call.scalar(alpha);
  if (!call.ok()) return call.status();
  call.run([&] { exoblas::ops::trmscal_rm<float>(call.ctx(), (int64_t)Uplo, (int64_t)N, alpha_v, call.dev<float>(0), (int64_t)A.strides[0]); });
  return call.finish();
}

// src/level3/mscal.py:20-27
extern "C" void exo_strmscal_rm_stride_1(void *ctxt, int_fast32_t Uplo, int_fast32_t N, const float *alpha, struct exo_win_2f32 A) {
  exoblas::Call call("exo_strmscal_rm_stride_1", exoblas::Call::kSync, ctxt);
  if (!(Uplo == 121 || Uplo == 122)) return (void)call.fail_enum("Uplo", Uplo);
  if (N < 0) return (void)call.fail_size("N", N);
  if ((true) && A.strides[1] != 1) return (void)call.fail_stride("A", A.strides[1]);
  if ((true) && N > 1 && A.strides[0] < N) return (void)call.fail_ld("A", A.strides[0]);
  call.window(0, A.data, 0, exoblas::span2_hi(N, N, A.strides[0]), sizeof(float), true, true);
  if (!call.stage()) return (void)call.status();
  const float alpha_v = call.scalar(alpha);
  if (!call.ok()) return (void)call.status();
  call.run([&] { exoblas::ops::trmscal_rm<float>(call.ctx(), (int64_t)Uplo, (int64_t)N, alpha_v, call.dev<float>(0), (int64_t)A.strides[0]); });
  return (void)call.finish();
}
extern "C" int exoblas_strmscal_rm_stride_1_async(hipStream_t stream, int_fast32_t Uplo, int_fast32_t N, const float *alpha, struct exo_win_2f32 A) {
  exoblas::Call call("exoblas_strmscal_rm_stride_1_async", exoblas::Call::kAsync, stream);
  if (!(Uplo == 121 || Uplo == 122)) return call.fail_enum("Uplo", Uplo);
  if (N < 0) return call.fail_size("N", N);
  if ((true) && A.strides[1] != 1) return call.fail_stride("A", A.strides[1]);
  if ((true) && N > 1 && A.strides[0] < N) return call.fail_ld("A", A.strides[0]);
  call.window(0, A.data, 0, exoblas::span2_hi(N, N, A.strides[0]), sizeof(float), true, true);
  if (!call.stage()) return call.status();
  const float alpha_v = call.scalar(alpha);
  if (!call.ok()) return call.status();
  call.run([&] { exoblas::ops::trmscal_rm<float>(call.ctx(), (int64_t)Uplo, (int64_t)N, alpha_v, call.dev<float>(0), (int64_t)A.strides[0]); });
  return call.finish();
}

// src/level3/mscal.py:20-27
extern "C" void exo_dtrmscal_rm_stride_any(void *ctxt, int_fast32_t Uplo, int_fast32_t N, const double *alpha, struct exo_win_2f64 A) {
  exoblas::Call call("exo_dtrmscal_rm_stride_any", exoblas::Call::kSync, ctxt);
  if (!(Uplo == 121 || Uplo == 122)) return (void)call.fail_enum("Uplo", Uplo);
  if (N < 0) return (void)call.fail_size("N", N);
  if ((true) && A.strides[1] != 1) return (void)call.fail_stride("A", A.strides[1]);
  if ((true) && N > 1 && A.strides[0] < N) return (void)call.fail_ld("A", A.strides[0]);
  call.window(0, A.data, 0, exoblas::span2_hi(N, N, A.strides[0]), sizeof(double), true, true);
  if (!call.stage()) return (void)call.status();
  const double alpha_v = call.scalar(alpha);
  if (!call.ok()) return (void)call.status();
  call.run([&] { exoblas::ops::trmscal_rm<double>(call.ctx(), (int64_t)Uplo, (int64_t)N, alpha_v, call.dev<double>(0), (int64_t)A.strides[0]); });
  return (void)call.finish();
}
extern "C" int exoblas_dtrmscal_rm_stride_any_async(hipStream_t stream, int_fast32_t Uplo, int_fast32_t N, const double *alpha, struct exo_win_2f64 A) {
  exoblas::Call call("exoblas_dtrmscal_rm_stride_any_async", exoblas::Call::kAsync, stream);
  if (!(Uplo == 121 || Uplo == 122)) return call.fail_enum("Uplo", Uplo);
  if (N < 0) return call.fail_size("N", N);
  if ((true) && A.strides[1] != 1) return call.fail_stride("A", A.strides[1]);
  if ((true) && N > 1 && A.strides[0] < N) return call.fail_ld("A", A.strides[0]);
  call.window(0, A.data, 0, exoblas::span2_hi(N, N, A.strides[0]), sizeof(double), true, true);
  if (!call.stage()) return call.status();
  const double alpha_v = call.scalar(alpha);
  if (!call.ok()) return call.status();
  call.run([&] { exoblas::ops::trmscal_rm<double>(call.ctx(), (int64_t)Uplo, (int64_t)N, alpha_v, call.dev<double>(0), (int64_t)A.strides[0]); });
  return call.finish();
}

// src/level3/mscal.py:20-27
extern "C" void exo_dtrmscal_rm_stride_1(void *ctxt, int_fast32_t Uplo, int_fast32_t N, const double *alpha, struct exo_win_2f64 A) {
  exoblas::Call call("exo_dtrmscal_rm_stride_1", exoblas::Call::kSync, ctxt);
  if (!(Uplo == 121 || Uplo == 122)) return (void)call.fail_enum("Uplo", Uplo);
  if (N < 0) return (void)call.fail_size("N", N);
  if ((true) && A.strides[1] != 1) return (void)call.fail_stride("A", A.strides[1]);
  if ((true) && N > 1 && A.strides[0] < N) return (void)call.fail_ld("A", A.strides[0]);
  call.window(0, A.data, 0, exoblas::span2_hi(N, N, A.strides[0]), sizeof(double), true, true);
  if (!call.stage()) return (void)call.status();
  const double alpha_v = call.scalar(alpha);
  if (!call.ok()) return (void)call.status();
  call.run([&] { exoblas::ops::trmscal_rm<double>(call.ctx(), (int64_t)Uplo, (int64_t)N, alpha_v, call.dev<double>(0), (int64_t)A.strides[0]); });
  return (void)call.finish();
}
extern "C" int exoblas_dtrmscal_rm_stride_1_async(hipStream_t stream, int_fast32_t Uplo, int_fast32_t N, const double *alpha, struct exo_win_2f64 A) {
  exoblas::Call call("exoblas_dtrmscal_rm_stride_1_async", exoblas::Call::kAsync, stream);
  if (!(Uplo == 121 || Uplo == 122)) return call.fail_enum("Uplo", Uplo);
  if (N < 0) return call.fail_size("N", N);
  if ((true) && A.strides[1] != 1) return call.fail_stride("A", A.strides[1]);
  if ((true) && N > 1 && A.strides[0] < N) return call.fail_ld("A", A.strides[0]);
  call.window(0, A.data, 0, exoblas::span2_hi(N, N, A.strides[0]), sizeof(double), true, true);
  if (!call.stage()) return call.status();
  const double alpha_v = call.scalar(alpha);
  if (!call.ok()) return call.status();
  call.run([&] { exoblas::ops::trmscal_rm<double>(call.ctx(), (int64_t)Uplo, (int64_t)N, alpha_v, call.dev<double>(0), (int64_t)A.strides[0]); });
  return call.finish();
}

// src/level3/symm.py:16-47
extern "C" void exo_ssymm_rm_stride_any(void *ctxt, int_fast32_t Side, int_fast32_t Uplo, int_fast32_t M, int_fast32_t N, const float *alpha, struct exo_win_2f32c Aleft, struct exo_win_2f32c Aright, struct exo_win_2f32c B, struct exo_win_2f32 C) {
  exoblas::Call call("exo_ssymm_rm_stride_any", exoblas::Call::kSync, ctxt);
  if (!(Side == 141 || Side == 142)) return (void)call.fail_enum("Side", Side);
  if (!(Uplo == 121 || Uplo == 122)) return (void)call.fail_enum("Uplo", Uplo);
  if (M < 0) return (void)call.fail_size("M", M);
  if (N < 0) return (void)call.fail_size("N", N);
  if ((Side == 141) && Aleft.strides[1] != 1) return (void)call.fail_stride("Aleft", Aleft.strides[1]);
  if ((Side == 141) && M > 1 && Aleft.strides[0] < M) return (void)call.fail_ld("Aleft", Aleft.strides[0]);
  if ((Side != 141) && Aright.strides[1] != 1) return (void)call.fail_stride("Aright", Aright.strides[1]);
  if ((Side != 141) && N > 1 && Aright.strides[0] < N) return (void)call.fail_ld("Aright", Aright.strides[0]);
  if ((true) && B.strides[1] != 1) return (void)call.fail_stride("B", B.strides[1]);
  if ((true) && M > 1 && B.strides[0] < N) return (void)call.fail_ld("B", B.strides[0]);
  if ((true) && C.strides[1] != 1) return (void)call.fail_stride("C", C.strides[1]);
  if ((true) && M > 1 && C.strides[0] < N) return (void)call.fail_ld("C", C.strides[0]);
  call.window(0, Aleft.data, 0, exoblas::span2_hi(M, M, Aleft.strides[0]), sizeof(float), false, Side == 141);
  call.window(1, Aright.data, 0, exoblas::span2_hi(N, N, Aright.strides[0]), sizeof(float), false, Side != 141);
  call.window(2, B.data, 0, exoblas::span2_hi(M, N, B.strides[0]), sizeof(float), false, true);
  call.window(3, C.data, 0, exoblas::span2_hi(M, N, C.strides[0]), sizeof(float), true, true);
  if (!call.stage()) return (void)call.status();
  const float alpha_v = call.scalar(alpha);
  if (!call.ok()) return (void)call.status();
  call.run([&] { exoblas::ops::symm_rm<float>(call.ctx(), (int64_t)Side, (int64_t)Uplo, (int64_t)M, (int64_t)N, alpha_v, call.dev<const float>(0), (int64_t)Aleft.strides[0], call.dev<const float>(1), (int64_t)Aright.strides[0], call.dev<const float>(2), (int64_t)B.strides[0], call.dev<float>(3), (int64_t)C.strides[0]); });
  return (void)call.finish();
}
extern "C" int exoblas_ssymm_rm_stride_any_async(hipStream_t stream, int_fast32_t Side, int_fast32_t Uplo, int_fast32_t M, int_fast32_t N, const float *alpha, struct exo_win_2f32c Aleft, struct exo_win_2f32c Aright, struct exo_win_2f32c B, struct exo_win_2f32 C) {
  exoblas::Call call("exoblas_ssymm_rm_stride_any_async", exoblas::Call::kAsync, stream);
  if (!(Side == 141 || Side == 142)) return call.fail_enum("Side", Side);
  if (!(Uplo == 121 || Uplo == 122)) return call.fail_enum("Uplo", Uplo);
  if (M < 0) return call.fail_size("M", M);
  if (N < 0) return call.fail_size("N", N);
  if ((Side == 141) && Aleft.strides[1] != 1) return call.fail_stride("Aleft", Aleft.strides[1]);
  if ((Side == 141) && M > 1 && Aleft.strides[0] < M) return call.fail_ld("Aleft", Aleft.strides[0]);
  if ((Side != 141) && Aright.strides[1] != 1) return call.fail_stride("Aright", Aright.strides[1]);
  if ((Side != 141) && N > 1 && Aright.strides[0] < N) return call.fail_ld("Aright", Aright.strides[0]);
  if ((true) && B.strides[1] != 1) return call.fail_stride("B", B.strides[1]);
  if ((true) && M > 1 && B.strides[0] < N) return call.fail_ld("B", B.strides[0]);
  if ((true) && C.strides[1] != 1) return call.fail_stride("C", C.strides[1]);
  if ((true) && M > 1 && C.strides[0] < N) return call.fail_ld("C", C.strides[0]);
  call.window(0, Aleft.data, 0, exoblas::span2_hi(M, M, Aleft.strides[0]), sizeof(float), false, Side == 141);
  call.window(1, Aright.data, 0, exoblas::span2_hi(N, N, Aright.strides[0]), sizeof(float), false, Side != 141);
  call.window(2, B.data, 0, exoblas::span2_hi(M, N, B.strides[0]), sizeof(float), false, true);
  call.window(3, C.data, 0, exoblas::span2_hi(M, N, C.strides[0]), sizeof(float), true, true);
  if (!call.stage()) return call.status();
  const float alpha_v = call.scalar(alpha);
  if (!call.ok()) return call.status();
  call.run([&] { exoblas::ops::symm_rm<float>(call.ctx(), (int64_t)Side, (int64_t)Uplo, (int64_t)M, (int64_t)N, alpha_v, call.dev<const float>(0), (int64_t)Aleft.strides[0], call.dev<const float>(1), (int64_t)Aright.strides[0], call.dev<const float>(2), (int64_t)B.strides[0], call.dev<float>(3), (int64_t)C.strides[0]); });
  return call.finish();
}

// src/level3/symm.py:16-47
extern "C" void exo_ssymm_rm_stride_1(void *ctxt, int_fast32_t Side, int_fast32_t Uplo, int_fast32_t M, int_fast32_t N, const float *alpha, struct exo_win_2f32c Aleft, struct exo_win_2f32c Aright, struct exo_win_2f32c B, struct exo_win_2f32 C) {
  exoblas::Call call("exo_ssymm_rm_stride_1", exoblas::Call::kSync, ctxt);
  if (!(Side == 141 || Side == 142)) return (void)call.fail_enum("Side", Side);
  if (!(Uplo == 121 || Uplo == 122)) return (void)call.fail_enum("Uplo", Uplo);
  if (M < 0) return (void)call.fail_size("M", M);
  if (N < 0) return (void)call.fail_size("N", N);
  if ((Side == 141) && Aleft.strides[1] != 1) return (void)call.fail_stride("Aleft", Aleft.strides[1]);
  if ((Side == 141) && M > 1 && Aleft.strides[0] < M) return (void)call.fail_ld("Aleft", Aleft.strides[0]);
  if ((Side != 141) && Aright.strides[1] != 1) return (void)call.fail_stride("Aright", Aright.strides[1]);
  if ((Side != 141) && N > 1 && Aright.strides[0] < N) return (void)call.fail_ld("Aright", Aright.strides[0]);
  if ((true) && B.strides[1] != 1) return (void)call.fail_stride("B", B.strides[1]);
  if ((true) && M > 1 && B.strides[0] < N) return (void)call.fail_ld("B", B.strides[0]);
  if ((true) && C.strides[1] != 1) return (void)call.fail_stride("C", C.strides[1]);
  if ((true) && M > 1 && C.strides[0] < N) return (void)call.fail_ld("C", C.strides[0]);
  call.window(0, Aleft.data, 0, exoblas::span2_hi(M, M, Aleft.strides[0]), sizeof(float), false, Side == 141);
  call.window(1, Aright.data, 0, exoblas::span2_hi(N, N, Aright.strides[0]), sizeof(float), false, Side != 141);
  call.window(2, B.data, 0, exoblas::span2_hi(M, N, B.strides[0]), sizeof(float), false, true);
  call.window(3, C.data, 0, exoblas::span2_hi(M, N, C.strides[0]), sizeof(float), true, true);
  if (!call.stage()) return (void)call.status();
  const float alpha_v = call.scalar(alpha);
  if (!call.ok()) return (void)call.status();
  call.run([&] { exoblas::ops::symm_rm<float>(call.ctx(), (int64_t)Side, (int64_t)Uplo, (int64_t)M, (int64_t)N, alpha_v, call.dev<const float>(0), (int64_t)Aleft.strides[0], call.dev<const float>(1), (int64_t)Aright.strides[0], call.dev<const float>(2), (int64_t)B.strides[0], call.dev<float>(3), (int64_t)C.strides[0]); });
  return (void)call.finish();
}
extern "C" int exoblas_ssymm_rm_stride_1_async(hipStream_t stream, int_fast32_t Side, int_fast32_t Uplo, int_fast32_t M, int_fast32_t N, const float *alpha, struct exo_win_2f32c Aleft, struct exo_win_2f32c Aright, struct exo_win_2f32c B, struct exo_win_2f32 C) {
  exoblas::Call call("exoblas_ssymm_rm_stride_1_async", exoblas::Call::kAsync, stream);
  if (!(Side == 141 || Side == 142)) return call.fail_enum("Side", Side);
  if (!(Uplo == 121 || Uplo == 122)) return call.fail_enum("Uplo", Uplo);
  if (M < 0) return call.fail_size("M", M);
  if (N < 0) return call.fail_size("N", N);
  if ((Side == 141) && Aleft.strides[1] != 1) return call.fail_stride("Aleft", Aleft.strides[1]);
  if ((Side == 141) && M > 1 && Aleft.strides[0] < M) return call.fail_ld("Aleft", Aleft.strides[0]);
  if ((Side != 141) && Aright.strides[1] != 1) return call.fail_stride("Aright", Aright.strides[1]);
  if ((Side != 141) && N > 1 && Aright.strides[0] < N) return call.fail_ld("Aright", Aright.strides[0]);
  if ((true) && B.strides[1] != 1) return call.fail_stride("B", B.strides[1]);
  if ((true) && M > 1 && B.strides[0] < N) return call.fail_ld("B", B.strides[0]);
  if ((true) && C.strides[1] != 1) return call.fail_stride("C", C.strides[1]);
  if ((true) && M > 1 && C.strides[0] < N) return call.fail_ld("C", C.strides[0]);
  call.window(0, Aleft.data, 0, exoblas::span2_hi(M, M, Aleft.strides[0]), sizeof(float), false, Side == 141);
  call.window(1, Aright.data, 0, exoblas::span2_hi(N, N, Aright.strides[0]), sizeof(float), false, Side != 141);
  call.window(2, B.data, 0, exoblas::span2_hi(M, N, B.strides[0]), sizeof(float), false, true);
  call.window(3, C.data, 0, exoblas::span2_hi(M, N, C.strides[0]), sizeof(float), true, true);
  if (!call.stage()) return call.status();
  const float alpha_v = call.scalar(alpha);
  if (!call.ok()) return call.status();
  call.run([&] { exoblas::ops::symm_rm<float>(call.ctx(), (int64_t)Side, (int64_t)Uplo, (int64_t)M, (int64_t)N, alpha_v, call.dev<const float>(0), (int64_t)Aleft.strides[0], call.dev<const float>(1), (int64_t)Aright.strides[0], call.dev<const float>(2), (int64_t)B.strides[0], call.dev<float>(3), (int64_t)C.strides[0]); });
  return call.finish();
}

// src/level3/symm.py:16-47
extern "C" void exo_dsymm_rm_stride_any(void *ctxt, int_fast32_t Side, int_fast32_t Uplo, int_fast32_t M, int_fast32_t N, const double *alpha, struct exo_win_2f64c Aleft, struct exo_win_2f64c Aright, struct exo_win_2f64c B, struct exo_win_2f64 C) {
  exoblas::Call call("exo_dsymm_rm_stride_any", exoblas::Call::kSync, ctxt);
  if (!(Side == 141 || Side == 142)) return (void)call.fail_enum("Side", Side);
  if (!(Uplo == 121 || Uplo == 122)) return (void)call.fail_enum("Uplo", Uplo);
  if (M < 0) return (void)call.fail_size("M", M);
  if (N < 0) return (void)call.fail_size("N", N);
  if ((Side == 141) && Aleft.strides[1] != 1) return (void)call.fail_stride("Aleft", Aleft.strides[1]);
  if ((Side == 141) && M > 1 && Aleft.strides[0] < M) return (void)call.fail_ld("Aleft", Aleft.strides[0]);
  if ((Side != 141) && Aright.strides[1] != 1) return (void)call.fail_stride("Aright", Aright.strides[1]);
  if ((Side != 141) && N > 1 && Aright.strides[0] < N) return (void)call.fail_ld("Aright", Aright.strides[0]);
  if ((true) && B.strides[1] != 1) return (void)call.fail_stride("B", B.strides[1]);
  if ((true) && M > 1 && B.strides[0] < N) return (void)call.fail_ld("B", B.strides[0]);
  if ((true) && C.strides[1] != 1) return (void)call.fail_stride("C", C.strides[1]);
  if ((true) && M > 1 && C.strides[0] < N) return (void)call.fail_ld("C", C.strides[0]);
  call.window(0, Aleft.data, 0, exoblas::span2_hi(M, M, Aleft.strides[0]), sizeof(double), false, Side == 141);
  call.window(1, Aright.data, 0, exoblas::span2_hi(N, N, Aright.strides[0]), sizeof(double), false, Side != 141);
  call.window(2, B.data, 0, exoblas::span2_hi(M, N, B.strides[0]), sizeof(double), false, true);
  call.window(3, C.data, 0, exoblas::span2_hi(M, N, C.strides[0]), sizeof(double), true, true);
  if (!call.stage()) return (void)call.status();
  const double alpha_v = call.scalar(alpha);
  if (!call.ok()) return (void)call.status();
  call.run([&] { exoblas::ops::symm_rm<double>(call.ctx(), (int64_t)Side, (int64_t)Uplo, (int64_t)M, (int64_t)N, alpha_v, call.dev<const double>(0), (int64_t)Aleft.strides[0], call.dev<const double>(1), (int64_t)Aright.strides[0], call.dev<const double>(2), (int64_t)B.strides[0], call.dev<double>(3), (int64_t)C.strides[0]); });
  return (void)call.finish();
}
extern "C" int exoblas_dsymm_rm_stride_any_async(hipStream_t stream, int_fast32_t Side, int_fast32_t Uplo, int_fast32_t M, int_fast32_t N, const double *alpha, struct exo_win_2f64c Aleft, struct exo_win_2f64c Aright, struct exo_win_2f64c B, struct exo_win_2f64 C) {
  exoblas::Call call("exoblas_dsymm_rm_stride_any_async", exoblas::Call::kAsync, stream);
  if (!(Side == 141 || Side == 142)) return call.fail_enum("Side", Side);
  if (!(Uplo == 121 || Uplo == 122)) return call.fail_enum("Uplo", Uplo);
  if (M < 0) return call.fail_size("M", M);
  if (N < 0) return call.fail_size("N", N);
  if ((Side == 141) && Aleft.strides[1] != 1) return call.fail_stride("Aleft", Aleft.strides[1]);
  if ((Side == 141) && M > 1 && Aleft.strides[0] < M) return call.fail_ld("Aleft", Aleft.strides[0]);
  if ((Side != 141) && Aright.strides[1] != 1) return call.fail_stride("Aright", Aright.strides[1]);
  if ((Side != 141) && N > 1 && Aright.strides[0] < N) return call.fail_ld("Aright", Aright.strides[0]);
  if ((true) && B.strides[1] != 1) return call.fail_stride("B", B.strides[1]);
  if ((true) && M > 1 && B.strides[0] < N) return call.fail_ld("B", B.strides[0]);
  if ((true) && C.strides[1] != 1) return call.fail_stride("C", C.strides[1]);
  if ((true) && M > 1 && C.strides[0] < N) return call.fail_ld("C", C.strides[0]);
  call.window(0, Aleft.data, 0, exoblas::span2_hi(M, M, Aleft.strides[0]), sizeof(double), false, Side == 141);
  call.window(1, Aright.data, 0, exoblas::span2_hi(N, N, Aright.strides[0]), sizeof(double), false, Side != 141);
  call.window(2, B.data, 0, exoblas::span2_hi(M, N, B.strides[0]), sizeof(double), false, true);
  call.window(3, C.data, 0, exoblas::span2_hi(M, N, C.strides[0]), sizeof(double), true, true);
  if (!call.stage()) return call.status();
  const double alpha_v = call.scalar(alpha);
  if (!call.ok()) return call.status();
  call.run([&] { exoblas::ops::symm_rm<double>(call.ctx(), (int64_t)Side, (int64_t)Uplo, (int64_t)M, (int64_t)N, alpha_v, call.dev<const double>(0), (int64_t)Aleft.strides[0], call.dev<const double>(1), (int64_t)Aright.strides[0], call.dev<const double>(2), (int64_t)B.strides[0], call.dev<double>(3), (int64_t)C.strides[0]); });
  return call.finish();
}

// src/level3/symm.py:16-47
extern "C" void exo_dsymm_rm_stride_1(void *ctxt, int_fast32_t Side, int_fast32_t Uplo, int_fast32_t M, int_fast32_t N, const double *alpha, struct exo_win_2f64c Aleft, struct exo_win_2f64c Aright, struct exo_win_2f64c B, struct exo_win_2f64 C) {
  exoblas::Call call("exo_dsymm_rm_stride_1", exoblas::Call::kSync, ctxt);
  if (!(Side == 141 || Side == 142)) return (void)call.fail_enum("Side", Side);
  if (!(Uplo == 121 || Uplo == 122)) return (void)call.fail_enum("Uplo", Uplo);
  if (M < 0) return (void)call.fail_size("M", M);
  if (N < 0) return (void)call.fail_size("N", N);
  if ((Side == 141) && Aleft.strides[1] != 1) return (void)call.fail_stride("Aleft", Aleft.strides[1]);
  if ((Side == 141) && M > 1 && Aleft.strides[0] < M) return (void)call.fail_ld("Aleft", Aleft.strides[0]);
  if ((Side != 141) && Aright.strides[1] != 1) return (void)call.fail_stride("Aright", Aright.strides[1]);
  if ((Side != 141) && N > 1 && Aright.strides[0] < N) return (void)call.fail_ld("Aright", Aright.strides[0]);
  if ((true) && B.strides[1] != 1) return (void)call.fail_stride("B", B.strides[1]);
  if ((true) && M > 1 && B.strides[0] < N) return (void)call.fail_ld("B", B.strides[0]);
  if ((true) && C.strides[1] != 1) return (void)call.fail_stride("C", C.strides[1]);
  if ((true) && M > 1 && C.strides[0] < N) return (void)call.fail_ld("C", C.strides[0]);
  call.window(0, Aleft.data, 0, exoblas::span2_hi(M, M, Aleft.strides[0]), sizeof(double), false, Side == 141);
  call.window(1, Aright.data, 0, exoblas::span2_hi(N, N, Aright.strides[0]), sizeof(double), false, Side != 141);
  call.window(2, B.data, 0, exoblas::span2_hi(M, N, B.strides[0]), sizeof(double), false, true);
  call.window(3, C.data, 0, exoblas::span2_hi(M, N, C.strides[0]), sizeof(double), true, true);
  if (!call.stage()) return (void)call.status();
  const double alpha_v = call.scalar(alpha);
  if (!call.ok()) return (void)call.status();
  call.run([&] { exoblas::ops::symm_rm<double>(call.ctx(), (int64_t)Side, (int64_t)Uplo, (int64_t)M, (int64_t)N, alpha_v, call.dev<const double>(0), (int64_t)Aleft.strides[0], call.dev<const double>(1), (int64_t)Aright.strides[0], call.dev<const double>(2), (int64_t)B.strides[0], call.dev<double>(3), (int64_t)C.strides[0]); });
  return (void)call.finish();
}
extern "C" int exoblas_dsymm_rm_stride_1_async(hipStream_t stream, int_fast32_t Side, int_fast32_t Uplo, int_fast32_t M, int_fast32_t N, const double *alpha, struct exo_win_2f64c Aleft, struct exo_win_2f64c Aright, struct exo_win_2f64c B, struct exo_win_2f64 C) {
  exoblas::Call call("exoblas_dsymm_rm_stride_1_async", exoblas::Call::kAsync, stream);
  if (!(Side == 141 || Side == 142)) return call.fail_enum("Side", Side);
  if (!(Uplo == 121 || Uplo == 122)) return call.fail_enum("Uplo", Uplo);
  if (M < 0) return call.fail_size("M", M);
  if (N < 0) return call.fail_size("N", N);
  if ((Side == 141) && Aleft.strides[1] != 1) return call.fail_stride("Aleft", Aleft.strides[1]);
  if ((Side == 141) && M > 1 && Aleft.strides[0] < M) return call.fail_ld("Aleft", Aleft.strides[0]);
  if ((Side != 141) && Aright.strides[1] != 1) return call.fail_stride("Aright", Aright.strides[1]);
  if ((Side != 141) && N > 1 && Aright.strides[0] < N) return call.fail_ld("Aright", Aright.strides[0]);
  if ((true) && B.strides[1] != 1) return call.fail_stride("B", B.strides[1]);
  if ((true) && M > 1 && B.strides[0] < N) return call.fail_ld("B", B.strides[0]);
  if ((true) && C.strides[1] != 1) return call.fail_stride("C", C.strides[1]);
  if ((true) && M > 1 && C.strides[0] < N) return call.fail_ld("C", C.strides[0]);
  call.window(0, Aleft.data, 0, exoblas::span2_hi(M, M, Aleft.strides[0]), sizeof(double), false, Side == 141);
  call.window(1, Aright.data, 0, exoblas::span2_hi(N, N, Aright.strides[0]), sizeof(double), false, Side != 141);
  call.window(2, B.data, 0, exoblas::span2_hi(M, N, B.strides[0]), sizeof(double), false, true);
  call.window(3, C.data, 0, exoblas::span2_hi(M, N, C.strides[0]), sizeof(double), true, true);
  if (!call.stage()) return call.status();
  const double alpha_v = call.scalar(alpha);
  if (!call.ok()) return call.status();
  call.run([&] { exoblas::ops::symm_rm<double>(call.ctx(), (int64_t)Side, (int64_t)Uplo, (int64_t)M, (int64_t)N, alpha_v, call.dev<const double>(0), (int64_t)Aleft.strides[0], call.dev<const double>(1), (int64_t)Aright.strides[0], call.dev<const double>(2), (int64_t)B.strides[0], call.dev<double>(3), (int64_t)C.strides[0]); });
  return call.finish();
}

// src/level3/syrk.py:17-32
extern "C" void exo_ssyrk_rm_stride_any(void *ctxt, int_fast32_t Uplo, int_fast32_t Trans, int_fast32_t N, int_fast32_t K, const float *alpha, struct exo_win_2f32c A, struct exo_win_2f32c A_, struct exo_win_2f32c AT, struct exo_win_2f32c AT_, struct exo_win_2f32 C) {
  exoblas::Call call("exo_ssyrk_rm_stride_any", exoblas::Call::kSync, ctxt);
  if (!(Uplo == 121 || Uplo == 122)) return (void)call.fail_enum("Uplo", Uplo);
  if (!(Trans == 111 || Trans == 112 || Trans == 113)) return (void)call.fail_enum("Trans", Trans);
  if (N < 0) return (void)call.fail_size("N", N);
  if (K < 0) return (void)call.fail_size("K", K);
  if ((Trans == 111) && A.strides[1] != 1) return (void)call.fail_stride("A", A.strides[1]);
  if ((Trans == 111) && N > 1 && A.strides[0] < K) return (void)call.fail_ld("A", A.strides[0]);
  if ((Trans == 111) && A_.strides[1] != 1) return (void)call.fail_stride("A_", A_.strides[1]);
  if ((Trans == 111) && N > 1 && A_.strides[0] < K) return (void)call.fail_ld("A_", A_.strides[0]);
  if ((Trans != 111) && AT.strides[1] != 1) return (void)call.fail_stride("AT", AT.strides[1]);
  if ((Trans != 111) && K > 1 && AT.strides[0] < N) return (void)call.fail_ld("AT", AT.strides[0]);
  if ((Trans != 111) && AT_.strides[1] != 1) return (void)call.fail_stride("AT_", AT_.strides[1]);
  if ((Trans != 111) && K > 1 && AT_.strides[0] < N) return (void)call.fail_ld("AT_", AT_.strides[0]);
  if ((true) && C.strides[1] != 1) return (void)call.fail_stride("C", C.strides[1]);
  if ((true) && N > 1 && C.strides[0] < N) return (void)call.fail_ld("C", C.strides[0]);
  call.window(0, A.data, 0, exoblas::span2_hi(N, K, A.strides[0]), sizeof(float), false, Trans == 111);
  call.window(1, A_.data, 0, exoblas::span2_hi(N, K, A_.strides[0]), sizeof(float), false, Trans == 111);
  call.window(2, AT.data, 0, exoblas::span2_hi(K, N, AT.strides[0]), sizeof(float), false, Trans != 111);
  call.window(3, AT_.data, 0, exoblas::span2_hi(K, N, AT_.strides[0]), sizeof(float), false, Trans != 111);
  call.window(4, C.data, 0, exoblas::span2_hi(N, N, C.strides[0]), sizeof(float), true, true);
  if (!call.stage()) return (void)call.status();
  const float alpha_v = call.scalar(alpha);
  if (!call.ok()) return (void)call.status();
  call.run([&] { exoblas::ops::syrk_rm<float>(call.ctx(), (int64_t)Uplo, (int64_t)Trans, (int64_t)N, (int64_t)K, alpha_v, call.dev<const float>(0), (int64_t)A.strides[0], call.dev<const float>(1), (int64_t)A_.strides[0], call.dev<const float>(2), (int64_t)AT.strides[0], call.dev<const float>(3), (int64_t)AT_.strides[0], call.dev<float>(4), (int64_t)C.strides[0]); });
  return (void)call.finish();
}
extern "C" int exoblas_ssyrk_rm_stride_any_async(hipStream_t stream, int_fast32_t Uplo, int_fast32_t Trans, int_fast32_t N, int_fast32_t K, const float *alpha, struct exo_win_2f32c A, struct exo_win_2f32c A_, struct exo_win_2f32c AT, struct exo_win_2f32c AT_, struct exo_win_2f32 C) {
  exoblas::Call call("exoblas_ssyrk_rm_stride_any_async", exoblas::Call::kAsync, stream);
  if (!(Uplo == 121 || Uplo == 122)) return call.fail_enum("Uplo", Uplo);
  if (!(Trans == 111 || Trans == 112 || Trans == 113)) return call.fail_enum("Trans", Trans);
  if (N < 0) return call.fail_size("N", N);
  if (K < 0) return call.fail_size("K", K);
  if ((Trans == 111) && A.strides[1] != 1) return call.fail_stride("A", A.strides[1]);
  if ((Trans == 111) && N > 1 && A.strides[0] < K) return call.fail_ld("A", A.strides[0]);
  if ((Trans == 111) && A_.strides[1] != 1) return call.fail_stride("A_", A_.strides[1]);
  if ((Trans == 111) && N > 1 && A_.strides[0] < K) return call.fail_ld("A_", A_.strides[0]);
  if ((Trans != 111) && AT.strides[1] != 1) return call.fail_stride("AT", AT.strides[1]);
  if ((Trans != 111) && K > 1 && AT.strides[0] < N) return call.fail_ld("AT", AT.strides[0]);
  if ((Trans != 111) && AT_.strides[1] != 1) return call.fail_stride("AT_", AT_.strides[1]);
  if ((Trans != 111) && K > 1 && AT_.strides[0] < N) return call.fail_ld("AT_", AT_.strides[0]);
  if ((true) && C.strides[1] != 1) return call.fail_stride("C", C.strides[1]);
  if ((true) && N > 1 && C.strides[0] < N) return call.fail_ld("C", C.strides[0]);
  call.window(0, A.data, 0, exoblas::span2_hi(N, K, A.strides[0]), sizeof(float), false, Trans == 111);
  call.window(1, A_.data, 0, exoblas::span2_hi(N, K, A_.strides[0]), sizeof(float), false, Trans == 111);
  call.window(2, AT.data, 0, exoblas::span2_hi(K, N, AT.strides[0]), sizeof(float), false, Trans != 111);
  call.window(3, AT_.data, 0, exoblas::span2_hi(K, N, AT_.strides[0]), sizeof(float), false, Trans != 111);
  call.window(4, C.data, 0, exoblas::span2_hi(N, N, C.strides[0]), sizeof(float), true, true);
  if (!call.stage()) return call.status();
  const float alpha_v = call.scalar(alpha);
  if (!call.ok()) return call.status();
  call.run([&] { exoblas::ops::syrk_rm<float>(call.ctx(), (int64_t)Uplo, (int64_t)Trans, (int64_t)N, (int64_t)K, alpha_v, call.dev<const float>(0), (int64_t)A.strides[0], call.dev<const float>(1), (int64_t)A_.strides[0], call.dev<const float>(2), (int64_t)AT.strides[0], call.dev<const float>(3), (int64_t)AT_.strides[0], call.dev<float>(4), (int64_t)C.strides[0]); });
  return call.finish();
}

// src/level3/syrk.py:17-32
extern "C" void exo_ssyrk_rm_stride_1(void *ctxt, int_fast32_t Uplo, int_fast32_t Trans, int_fast32_t N, int_fast32_t K, const float *alpha, struct exo_win_2f32c A, struct exo_win_2f32c A_, struct exo_win_2f32c AT, struct exo_win_2f32c AT_, struct exo_win_2f32 C) {
  exoblas::Call call("exo_ssyrk_rm_stride_1", exoblas::Call::kSync, ctxt);
  if (!(Uplo == 121 || Uplo == 122)) return (void)call.fail_enum("Uplo", Uplo);
  if (!(Trans == 111 || Trans == 112 || Trans == 113)) return (void)call.fail_enum("Trans", Trans);
  if (N < 0) return (void)call.fail_size("N", N);
  if (K < 0) return (void)call.fail_size("K", K);
  if ((Trans == 111) && A.strides[1] != 1) return (void)call.fail_stride("A", A.strides[1]);
  if ((Trans == 111) && N > 1 && A.strides[0] < K) return (void)call.fail_ld("A", A.strides[0]);
  if ((Trans == 111) && A_.strides[1] != 1) return (void)call.fail_stride("A_", A_.strides[1]);
  if ((Trans == 111) && N > 1 && A_.strides[0] < K) return (void)call.fail_ld("A_", A_.strides[0]);
  if ((Trans != 111) && AT.strides[1] != 1) return (void)call.fail_stride("AT", AT.strides[1]);
  if ((Trans != 111) && K > 1 && AT.strides[0] < N) return (void)call.fail_ld("AT", AT.strides[0]);
  if ((Trans != 111) && AT_.strides[1] != 1) return (void)call.fail_stride("AT_", AT_.strides[1]);
  if ((Trans != 111) && K > 1 && AT_.strides[0] < N) return (void)call.fail_ld("AT_", AT_.strides[0]);
  if ((true) && C.strides[1] != 1) return (void)call.fail_stride("C", C.strides[1]);
  if ((true) && N > 1 && C.strides[0] < N) return (void)call.fail_ld("C", C.strides[0]);
  call.window(0, A.data, 0, exoblas::span2_hi(N, K, A.strides[0]), sizeof(float), false, Trans == 111);
  call.window(1, A_.data, 0, exoblas::span2_hi(N, K, A_.strides[0]), sizeof(float), false, Trans == 111);
  call.window(2, AT.data, 0, exoblas::span2_hi(K, N, AT.strides[0]), sizeof(float), false, Trans != 111);
  call.window(3, AT_.data, 0, exoblas::span2_hi(K, N, AT_.strides[0]), sizeof(float), false, Trans != 111);
  call.window(4, C.data, 0, exoblas::span2_hi(N, N, C.strides[0]), sizeof(float), true, true);
  if (!call.stage()) return (void)call.status();
  const float alpha_v = call.scalar(alpha);
  if (!call.ok()) return (void)call.status();
  call.run([&] { exoblas::ops::syrk_rm<float>(call.ctx(), (int64_t)Uplo, (int64_t)Trans, (int64_t)N, (int64_t)K, alpha_v, call.dev<const float>(0), (int64_t)A.strides[0], call.dev<const float>(1), (int64_t)A_.strides[0], call.dev<const float>(2), (int64_t)AT.strides[0], call.dev<const float>(3), (int64_t)AT_.strides[0], call.dev<float>(4), (int64_t)C.strides[0]); });
  return (void)call.finish();
}
extern "C" int exoblas_ssyrk_rm_stride_1_async(hipStream_t stream, int_fast32_t Uplo, int_fast32_t Trans, int_fast32_t N, int_fast32_t K, const float *alpha, struct exo_win_2f32c A, struct exo_win_2f32c A_, struct exo_win_2f32c AT, struct exo_win_2f32c AT_, struct exo_win_2f32 C) {
  exoblas::Call call("exoblas_ssyrk_rm_stride_1_async", exoblas::Call::kAsync, stream);
  if (!(Uplo == 121 || Uplo == 122)) return call.fail_enum("Uplo", Uplo);
  if (!(Trans == 111 || Trans == 112 || Trans == 113)) return call.fail_enum("Trans", Trans);
  if (N < 0) return call.fail_size("N", N);
  if (K < 0) return call.fail_size("K", K);
  if ((Trans == 111) && A.strides[1] != 1) return call.fail_stride("A", A.strides[1]);
  if ((Trans == 111) && N > 1 && A.strides[0] < K) return call.fail_ld("A", A.strides[0]);
  if ((Trans == 111) && A_.strides[1] != 1) return call.fail_stride("A_", A_.strides[1]);
  if ((Trans == 111) && N > 1 && A_.strides[0] < K) return call.fail_ld("A_", A_.strides[0]);
  if ((Trans != 111) && AT.strides[1] != 1) return call.fail_stride("AT", AT.strides[1]);
  if ((Trans != 111) && K > 1 && AT.strides[0] < N) return call.fail_ld("AT", AT.strides[0]);
  if ((Trans != 111) && AT_.strides[1] != 1) return call.fail_stride("AT_", AT_.strides[1]);
  if ((Trans != 111) && K > 1 && AT_.strides[0] < N) return call.fail_ld("AT_", AT_.strides[0]);
  if ((true) && C.strides[1] != 1) return call.fail_stride("C", C.strides[1]);
  if ((true) && N > 1 && C.strides[0] < N) return call.fail_ld("C", C.strides[0]);
  call.window(0, A.data, 0, exoblas::span2_hi(N, K, A.strides[0]), sizeof(float), false, Trans == 111);
  call.window(1, A_.data, 0, exoblas::span2_hi(N, K, A_.strides[0]), sizeof(float), false, Trans == 111);
  call.window(2, AT.data, 0, exoblas::span2_hi(K, N, AT.strides[0]), sizeof(float), false, Trans != 111);
  call.window(3, AT_.data, 0, exoblas::span2_hi(K, N, AT_.strides[0]), sizeof(float), false, Trans != 111);
  call.window(4, C.data, 0, exoblas::span2_hi(N, N, C.strides[0]), sizeof(float), true, true);
  if (!call.stage()) return call.status();
  const float alpha_v = call.scalar(alpha);
  if (!call.ok()) return call.status();
  call.run([&] { exoblas::ops::syrk_rm<float>(call.ctx(), (int64_t)Uplo, (int64_t)Trans, (int64_t)N, (int64_t)K, alpha_v, call.dev<const float>(0), (int64_t)A.strides[0], call.dev<const float>(1), (int64_t)A_.strides[0], call.dev<const float>(2), (int64_t)AT.strides[0], call.dev<const float>(3), (int64_t)AT_.strides[0], call.dev<float>(4), (int64_t)C.strides[0]); });
  return call.finish();
}

// src/level3/syrk.py:17-32
extern "C" void exo_dsyrk_rm_stride_any(void *ctxt, int_fast32_t Uplo, int_fast32_t Trans, int_fast32_t N, int_fast32_t K, const double *alpha, struct exo_win_2f64c A, struct exo_win_2f64c A_, struct exo_win_2f64c AT, struct exo_win_2f64c AT_, struct exo_win_2f64 C) {
  exoblas::Call call("exo_dsyrk_rm_stride_any", exoblas::Call::kSync, ctxt);
  if (!(Uplo == 121 || Uplo == 122)) return (void)call.fail_enum("Uplo", Uplo);
  if (!(Trans == 111 || Trans == 112 || Trans == 113)) return (void)call.fail_enum("Trans", Trans);
  if (N < 0) return (void)call.fail_size("N", N);
  if (K < 0) return (void)call.fail_size("K", K);
  if ((Trans == 111) && A.strides[1] != 1) return (void)call.fail_stride("A", A.strides[1]);
  if ((Trans == 111) && N > 1 && A.strides[0] < K) return (void)call.fail_ld("A", A.strides[0]);
  if ((Trans == 111) && A_.strides[1] != 1) return (void)call.fail_stride("A_", A_.strides[1]);
  if ((Trans == 111) && N > 1 && A_.strides[0] < K) return (void)call.fail_ld("A_", A_.strides[0]);
  if ((Trans != 111) && AT.strides[1] != 1) return (void)call.fail_stride("AT", AT.strides[1]);
  if ((Trans != 111) && K > 1 && AT.strides[0] < N) return (void)call.fail_ld("AT", AT.strides[0]);
  if ((Trans != 111) && AT_.strides[1] != 1) return (void)call.fail_stride("AT_", AT_.strides[1]);
  if ((Trans != 111) && K > 1 && AT_.strides[0] < N) return (void)call.fail_ld("AT_", AT_.strides[0]);
  if ((true) && C.strides[1] != 1) return (void)call.fail_stride("C", C.strides[1]);
  if ((true) && N > 1 && C.strides[0] < N) return (void)call.fail_ld("C", C.strides[0]);
  call.window(0, A.data, 0, exoblas::span2_hi(N, K, A.strides[0]), sizeof(double), false, Trans == 111);
  call.window(1, A_.data, 0, exoblas::span2_hi(N, K, A_.strides[0]), sizeof(double), false, Trans == 111);
  call.window(2, AT.data, 0, exoblas::span2_hi(K, N, AT.strides[0]), sizeof(double), false, Trans != 111);
  call.window(3, AT_.data, 0, exoblas::span2_hi(K, N, AT_.strides[0]), sizeof(double), false, Trans != 111);
  call.window(4, C.data, 0, exoblas::span2_hi(N, N, C.strides[0]), sizeof(double), true, true);
  if (!call.stage()) return (void)call.status();
  const double alpha_v = call.scalar(alpha);
  if (!call.ok()) return (void)call.status();
  call.run([&] { exoblas::ops::syrk_rm<double>(call.ctx(), (int64_t)Uplo, (int64_t)Trans, (int64_t)N, (int64_t)K, alpha_v, call.dev<const double>(0), (int64_t)A.strides[0], call.dev<const double>(1), (int64_t)A_.strides[0], call.dev<const double>(2), (int64_t)AT.strides[0], call.dev<const double>(3), (int64_t)AT_.strides[0], call.dev<double>(4), (int64_t)C.strides[0]); });
  return (void)call.finish();
}
extern "C" int exoblas_dsyrk_rm_stride_any_async(hipStream_t stream, int_fast32_t Uplo, int_fast32_t Trans, int_fast32_t N, int_fast32_t K, const double *alpha, struct exo_win_2f64c A, struct exo_win_2f64c A_, struct exo_win_2f64c AT, struct exo_win_2f64c AT_, struct exo_win_2f64 C) {
  exoblas::Call call("exoblas_dsyrk_rm_stride_any_async", exoblas::Call::kAsync, stream);
  if (!(Uplo == 121 || Uplo == 122)) return call.fail_enum("Uplo", Uplo);
  if (!(Trans == 111 || Trans == 112 || Trans == 113)) return call.fail_enum("Trans", Trans);
  if (N < 0) return call.fail_size("N", N);
  if (K < 0) return call.fail_size("K", K);
  if ((Trans == 111) && A.strides[1] != 1) return call.fail_stride("A", A.strides[1]);
  if ((Trans == 111) && N > 1 && A.strides[0] < K) return call.fail_ld("A", A.strides[0]);
  if ((Trans == 111) && A_.strides[1] != 1) return call.fail_stride("A_", A_.strides[1]);
  if ((Trans == 111) && N > 1 && A_.strides[0] < K) return call.fail_ld("A_", A_.strides[0]);
  if ((Trans != 111) && AT.strides[1] != 1) return call.fail_stride("AT", AT.strides[1]);
  if ((Trans != 111) && K > 1 && AT.strides[0] < N) return call.fail_ld("AT", AT.strides[0]);
  if ((Trans != 111) && AT_.strides[1] != 1) return call.fail_stride("AT_", AT_.strides[1]);
  if ((Trans != 111) && K > 1 && AT_.strides[0] < N) return call.fail_ld("AT_", AT_.strides[0]);
  if ((true) && C.strides[1] != 1) return call.fail_stride("C", C.strides[1]);
  if ((true) && N > 1 && C.strides[0] < N) return call.fail_ld("C", C.strides[0]);
  call.window(0, A.data, 0, exoblas::span2_hi(N, K, A.strides[0]), sizeof(double), false, Trans == 111);
  call.window(1, A_.data, 0, exoblas::span2_hi(N, K, A_.strides[0]), sizeof(double), false, Trans == 111);
  call.window(2, AT.data, 0, exoblas::span2_hi(K, N, AT.strides[0]), sizeof(double), false, Trans != 111);
  call.window(3, AT_.data, 0, exoblas::span2_hi(K, N, AT_.strides[0]), sizeof(double), false, Trans != 111);
  call.window(4, C.data, 0, exoblas::span2_hi(N, N, C.strides[0]), sizeof(double), true, true);
  if (!call.stage()) return call.status();
  const double alpha_v = call.scalar(alpha);
  if (!call.ok()) return call.status();
  call.run([&] { exoblas::ops::syrk_rm<double>(call.ctx(), (int64_t)Uplo, (int64_t)Trans, (int64_t)N, (int64_t)K, alpha_v, call.dev<const double>(0), (int64_t)A.strides[0], call.dev<const double>(1), (int64_t)A_.strides[0], call.dev<const double>(2), (int64_t)AT.strides[0], call.dev<const double>(3), (int64_t)AT_.strides[0], call.dev<double>(4), (int64_t)C.strides[0]); });
  return call.finish();
}

// src/level3/syrk.py:17-32
extern "C" void exo_dsyrk_rm_stride_1(void *ctxt, int_fast32_t Uplo, int_fast32_t Trans, int_fast32_t N, int_fast32_t K, const double *alpha, struct exo_win_2f64c A, struct exo_win_2f64c A_, struct exo_win_2f64c AT, struct exo_win_2f64c AT_, struct exo_win_2f64 C) {
  exoblas::Call call("exo_dsyrk_rm_stride_1", exoblas::Call::kSync, ctxt);
  if (!(Uplo == 121 || Uplo == 122)) return (void)call.fail_enum("Uplo", Uplo);
  if (!(Trans == 111 || Trans == 112 || Trans == 113)) return (void)call.fail_enum("Trans", Trans);
  if (N < 0) return (void)call.fail_size("N", N);
  if (K < 0) return (void)call.fail_size("K", K);
  if ((Trans == 111) && A.strides[1] != 1) return (void)call.fail_stride("A", A.strides[1]);
  if ((Trans == 111) && N > 1 && A.strides[0] < K) return (void)call.fail_ld("A", A.strides[0]);
  if ((Trans == 111) && A_.strides[1] != 1) return (void)call.fail_stride("A_", A_.strides[1]);
  if ((Trans == 111) && N > 1 && A_.strides[0] < K) return (void)call.fail_ld("A_", A_.strides[0]);
  if ((Trans != 111) && AT.strides[1] != 1) return (void)call.fail_stride("AT", AT.strides[1]);
  if ((Trans != 111) && K > 1 && AT.strides[0] < N) return (void)call.fail_ld("AT", AT.strides[0]);
  if ((Trans != 111) && AT_.strides[1] != 1) return (void)call.fail_stride("AT_", AT_.strides[1]);
  if ((Trans != 111) && K > 1 && AT_.strides[0] < N) return (void)call.fail_ld("AT_", AT_.strides[0]);
  if ((true) && C.strides[1] != 1) return (void)call.fail_stride("C", C.strides[1]);
  if ((true) && N > 1 && C.strides[0] < N) return (void)call.fail_ld("C", C.strides[0]);
  call.window(0, A.data, 0, exoblas::span2_hi(N, K, A.strides[0]), sizeof(double), false, Trans == 111);
  call.window(1, A_.data, 0, exoblas::span2_hi(N, K, A_.strides[0]), sizeof(double), false, Trans == 111);
  call.window(2, AT.data, 0, exoblas::span2_hi(K, N, AT.strides[0]), sizeof(double), false, Trans != 111);
  call.window(3, AT_.data, 0, exoblas::span2_hi(K, N, AT_.strides[0]), sizeof(double), false, Trans != 111);
  call.window(4, C.data, 0, exoblas::span2_hi(N, N, C.strides[0]), sizeof(double), true, true);
  if (!call.stage()) return (void)call.status();
  const double alpha_v = call.scalar(alpha);
  if (!call.ok()) return (void)call.status();
  call.run([&] { exoblas::ops::syrk_rm<double>(call.ctx(), (int64_t)Uplo, (int64_t)Trans, (int64_t)N, (int64_t)K, alpha_v, call.dev<const double>(0), (int64_t)A.strides[0], call.dev<const double>(1), (int64_t)A_.strides[0], call.dev<const double>(2), (int64_t)AT.strides[0], call.dev<const double>(3), (int64_t)AT_.strides[0], call.dev<double>(4), (int64_t)C.strides[0]); });
  return (void)call.finish();
}
extern "C" int exoblas_dsyrk_rm_stride_1_async(hipStream_t stream, int_fast32_t Uplo, int_fast32_t Trans, int_fast32_t N, int_fast32_t K, const double *alpha, struct exo_win_2f64c A, struct exo_win_2f64c A_, struct exo_win_2f64c AT, struct exo_win_2f64c AT_, struct exo_win_2f64 C) {
  exoblas::Call call("exoblas_dsyrk_rm_stride_1_async", exoblas::Call::kAsync, stream);
  if (!(Uplo == 121 || Uplo == 122)) return call.fail_enum("Uplo", Uplo);
  if (!(Trans == 111 || Trans == 112 || Trans == 113)) return call.fail_enum("Trans", Trans);
  if (N < 0) return call.fail_size("N", N);
  if (K < 0) return call.fail_size("K", K);
  if ((Trans == 111) && A.strides[1] != 1) return call.fail_stride("A", A.strides[1]);
  if ((Trans == 111) && N > 1 && A.strides[0] < K) return call.fail_ld("A", A.strides[0]);
  if ((Trans == 111) && A_.strides[1] != 1) return call.fail_stride("A_", A_.strides[1]);
  if ((Trans == 111) && N > 1 && A_.strides[0] < K) return call.fail_ld("A_", A_.strides[0]);
  if ((Trans != 111) && AT.strides[1] != 1) return call.fail_stride("AT", AT.strides[1]);
  if ((Trans != 111) && K > 1 && AT.strides[0] < N) return call.fail_ld("AT", AT.strides[0]);
  if ((Trans != 111) && AT_.strides[1] != 1) return call.fail_stride("AT_", AT_.strides[1]);
  if ((Trans != 111) && K > 1 && AT_.strides[0] < N) return call.fail_ld("AT_", AT_.strides[0]);
  if ((true) && C.strides[1] != 1) return call.fail_stride("C", C.strides[1]);
  if ((true) && N > 1 && C.strides[0] < N) return call.fail_ld("C", C.strides[0]);
  call.window(0, A.data, 0, exoblas::span2_hi(N, K, A.strides[0]), sizeof(double), false, Trans == 111);
  call.window(1, A_.data, 0, exoblas::span2_hi(N, K, A_.strides[0]), sizeof(double), false, Trans == 111);
  call.window(2, AT.data, 0, exoblas::span2_hi(K, N, AT.strides[0]), sizeof(double), false, Trans != 111);
  call.window(3, AT_.data, 0, exoblas::span2_hi(K, N, AT_.strides[0]), sizeof(double), false, Trans != 111);
  call.window(4, C.data, 0, exoblas::span2_hi(N, N, C.strides[0]), sizeof(double), true, true);
  if (!call.stage()) return call.status();
  const double alpha_v = call.scalar(alpha);
  if (!call.ok()) return call.status();
  call.run([&] { exoblas::ops::syrk_rm<double>(call.ctx(), (int64_t)Uplo, (int64_t)Trans, (int64_t)N, (int64_t)K, alpha_v, call.dev<const double>(0), (int64_t)A.strides[0], call.dev<const double>(1), (int64_t)A_.strides[0], call.dev<const double>(2), (int64_t)AT.strides[0], call.dev<const double>(3), (int64_t)AT_.strides[0], call.dev<double>(4), (int64_t)C.strides[0]); });
  return call.finish();
}

// src/level3/syr2k.py:14-39
extern "C" void exo_ssyr2k_rm_stride_any(void *ctxt, int_fast32_t Uplo, int_fast32_t Trans, int_fast32_t N, int_fast32_t K, const float *alpha, struct exo_win_2f32c A, struct exo_win_2f32c A_, struct exo_win_2f32c B, struct exo_win_2f32c B_, struct exo_win_2f32c AT, struct exo_win_2f32c AT_, struct exo_win_2f32c BT, struct exo_win_2f32c BT_, struct exo_win_2f32 C) {
  exoblas::Call call("exo_ssyr2k_rm_stride_any", exoblas::Call::kSync, ctxt);
  if (!(Uplo == 121 || Uplo == 122)) return (void)call.fail_enum("Uplo", Uplo);
  if (!(Trans == 111 || Trans == 112 || Trans == 113)) return (void)call.fail_enum("Trans", Trans);
  if (N < 0) return (void)call.fail_size("N", N);
  if (K < 0) return (void)call.fail_size("K", K);
  if ((Trans == 111) && A.strides[1] != 1) return (void)call.fail_stride("A", A.strides[1]);
  if ((Trans == 111) && N > 1 && A.strides[0] < K) return (void)call.fail_ld("A", A.strides[0]);
  if ((Trans == 111) && A_.strides[1] != 1) return (void)call.fail_stride("A_", A_.strides[1]);
  if ((Trans == 111) && N > 1 && A_.strides[0] < K) return (void)call.fail_ld("A_", A_.strides[0]);
  if ((Trans == 111) && B.strides[1] != 1) return (void)call.fail_stride("B", B.strides[1]);
  if ((Trans == 111) && N > 1 && B.strides[0] < K) return (void)call.fail_ld("B", B.strides[0]);
  if ((Trans == 111) && B_.strides[1] != 1) return (void)call.fail_stride("B_", B_.strides[1]);
  if ((Trans == 111) && N > 1 && B_.strides[0] < K) return (void)call.fail_ld("B_", B_.strides[0]);
  if ((Trans != 111) && AT.strides[1] != 1) return (void)call.fail_stride("AT", AT.strides[1]);
  if ((Trans != 111) && K > 1 && AT.strides[0] < N) return (void)call.fail_ld("AT", AT.strides[0]);
  if ((Trans != 111) && AT_.strides[1] != 1) return (void)call.fail_stride("AT_", AT_.strides[1]);
  if ((Trans != 111) && K > 1 && AT_.strides[0] < N) return (void)call.fail_ld("AT_", AT_.strides[0]);
  if ((Trans != 111) && BT.strides[1] != 1) return (void)call.fail_stride("BT", BT.strides[1]);
  if ((Trans != 111) && K > 1 && BT.strides[0] < N) return (void)call.fail_ld("BT", BT.strides[0]);
  if ((Trans != 111) && BT_.strides[1] != 1) return (void)call.fail_stride("BT_", BT_.strides[1]);
  if ((Trans != 111) && K > 1 && BT_.strides[0] < N) return (void)call.fail_ld("BT_", BT_.strides[0]);
  if ((true) && C.strides[1] != 1) return (void)call.fail_stride("C", C.strides[1]);
  if ((true) && N > 1 && C.strides[0] < N) return (void)call.fail_ld("C", C.strides[0]);
  call.window(0, A.data, 0, exoblas::span2_hi(N, K, A.strides[0]), sizeof(float), false, Trans == 111);
  call.window(1, A_.data, 0, exoblas::span2_hi(N, K, A_.strides[0]), sizeof(float), false, Trans == 111);
  call.window(2, B.data, 0, exoblas::span2_hi(N, K, B.strides[0]), sizeof(float), false, Trans == 111);
  call.window(3, B_.data, 0, exoblas::span2_hi(N, K, B_.strides[0]), sizeof(float), false, Trans == 111);
  call.window(4, AT.data, 0, exoblas::span2_hi(K, N, AT.strides[0]), sizeof(float), false, Trans != 111);
  call.window(5, AT_.data, 0, exoblas::span2_hi(K, N, AT_.strides[0]), sizeof(float), false, Trans != 111);
  call.window(6, BT.data, 0, exoblas::span2_hi(K, N, BT.strides[0]), sizeof(float), false, Trans != 111);
  call.window(7, BT_.data, 0, exoblas::span2_hi(K, N, BT_.strides[0]), sizeof(float), false, Trans != 111);
  call.window(8, C.data, 0, exoblas::span2_hi(N, N, C.strides[0]), sizeof(float), true, true);
  if (!call.stage()) return (void)call.status();
  const float alpha_v = call.scalar(alpha);
  if (!call.ok()) return (void)call.status();
  call.run([&] { exoblas::ops::syr2k_rm<float>(call.ctx(), (int64_t)Uplo, (int64_t)Trans, (int64_t)N, (int64_t)K, alpha_v, call.dev<const float>(0), (int64_t)A.strides[0], call.dev<const float>(1), (int64_t)A_.strides[0], call.dev<const float>(2), (int64_t)B.strides[0], call.dev<const float>(3), (int64_t)B_.strides[0], call.dev<const float>(4), (int64_t)AT.strides[0], call.dev<const float>(5), (int64_t)AT_.strides[0], call.dev<const float>(6), (int64_t)BT.strides[0], call.dev<const float>(7), (int64_t)BT_.strides[0], call.dev<float>(8), (int64_t)C.strides[0]); });
  return (void)call.finish();
}
extern "C" int exoblas_ssyr2k_rm_stride_any_async(hipStream_t stream, int_fast32_t Uplo, int_fast32_t Trans, int_fast32_t N, int_fast32_t K, const float *alpha, struct exo_win_2f32c A, struct exo_win_2f32c A_, struct exo_win_2f32c B, struct exo_win_2f32c B_, struct exo_win_2f32c AT, struct exo_win_2f32c AT_, struct exo_win_2f32c BT, struct exo_win_2f32c BT_, struct exo_win_2f32 C) {
  exoblas::Call call("exoblas_ssyr2k_rm_stride_any_async", exoblas::Call::kAsync, stream);
  if (!(Uplo == 121 || Uplo == 122)) return call.fail_enum("Uplo", Uplo);
  if (!(Trans == 111 || Trans == 112 || Trans == 113)) return call.fail_enum("Trans", Trans);
  if (N < 0) return call.fail_size("N", N);
  if (K < 0) return call.fail_size("K", K);
  if ((Trans == 111) && A.strides[1] != 1) return call.fail_stride("A", A.strides[1]);
  if ((Trans == 111) && N > 1 && A.strides[0] < K) return call.fail_ld("A", A.strides[0]);
  if ((Trans == 111) && A_.strides[1] != 1) return call.fail_stride("A_", A_.strides[1]);
  if ((Trans == 111) && N > 1 && A_.strides[0] < K) return call.fail_ld("A_", A_.strides[0]);
  if ((Trans == 111) && B.strides[1] != 1) return call.fail_stride("B", B.strides[1]);
  if ((Trans == 111) && N > 1 && B.strides[0] < K) return call.fail_ld("B", B.strides[0]);
  if ((Trans == 111) && B_.strides[1] != 1) return call.fail_stride("B_", B_.strides[1]);
  if ((Trans == 111) && N > 1 && B_.strides[0] < K) return call.fail_ld("B_", B_.strides[0]);
  if ((Trans != 111) && AT.strides[1] != 1) return call.fail_stride("AT", AT.strides[1]);
  if ((Trans != 111) && K > 1 && AT.strides[0] < N) return call.fail_ld("AT", AT.strides[0]);
  if ((Trans != 111) && AT_.strides[1] != 1) return call.fail_stride("AT_", AT_.strides[1]);
  if ((Trans != 111) && K > 1 && AT_.strides[0] < N) return call.fail_ld("AT_", AT_.strides[0]);
  if ((Trans != 111) && BT.strides[1] != 1) return call.fail_stride("BT", BT.strides[1]);
  if ((Trans != 111) && K > 1 && BT.strides[0] < N) return call.fail_ld("BT", BT.strides[0]);
  if ((Trans != 111) && BT_.strides[1] != 1) return call.fail_stride("BT_", BT_.strides[1]);
  if ((Trans != 111) && K > 1 && BT_.strides[0] < N) return call.fail_ld("BT_", BT_.strides[0]);
  if ((true) && C.strides[1] != 1) return call.fail_stride("C", C.strides[1]);
  if ((true) && N > 1 && C.strides[0] < N) return call.fail_ld("C", C.strides[0]);
  call.window(0, A.data, 0, exoblas::span2_hi(N, K, A.strides[0]), sizeof(float), false, Trans == 111);
  call.window(1, A_.data, 0, exoblas::span2_hi(N, K, A_.strides[0]), sizeof(float), false, Trans == 111);
  call.window(2, B.data, 0, exoblas::span2_hi(N, K, B.strides[0]), sizeof(float), false, Trans == 111);
  call.window(3, B_.data, 0, exoblas::span2_hi(N, K, B_.strides[0]), sizeof(float), false, Trans == 111);
  call.window(4, AT.data, 0, exoblas::span2_hi(K, N, AT.strides[0]), sizeof(float), false, Trans != 111);
  call.window(5, AT_.data, 0, exoblas::span2_hi(K, N, AT_.strides[0]), sizeof(float), false, Trans != 111);
  call.window(6, BT.data, 0, exoblas::span2_hi(K, N, BT.strides[0]), sizeof(float), false, Trans != 111);
  call.window(7, BT_.data, 0, exoblas::span2_hi(K, N, BT_.strides[0]), sizeof(float), false, Trans != 111);
  call.window(8, C.data, 0, exoblas::span2_hi(N, N, C.strides[0]), sizeof(float), true, true);
  if (!call.stage()) return call.status();
  const float alpha_v = call.scalar(alpha);
  if (!call.ok()) return call.status();
  call.run([&] { exoblas::ops::syr2k_rm<float>(call.ctx(), (int64_t)Uplo, (int64_t)Trans, (int64_t)N, (int64_t)K, alpha_v, call.dev<const float>(0), (int64_t)A.strides[0], call.dev<const float>(1), (int64_t)A_.strides[0], call.dev<const float>(2), (int64_t)B.strides[0], call.dev<const float>(3), (int64_t)B_.strides[0], call.dev<const float>(4), (int64_t)AT.strides[0], call.dev<const float>(5), (int64_t)AT_.strides[0], call.dev<const float>(6), (int64_t)BT.strides[0], call.dev<const float>(7), (int64_t)BT_.strides[0], call.dev<float>(8), (int64_t)C.strides[0]); });
  return call.finish();
}

// src/level3/syr2k.py:14-39
extern "C" void exo_ssyr2k_rm_stride_1(void *ctxt, int_fast32_t Uplo, int_fast32_t Trans, int_fast32_t N, int_fast32_t K, const float *alpha, struct exo_win_2f32c A, struct exo_win_2f32c A_, struct exo_win_2f32c B, struct exo_win_2f32c B_, struct exo_win_2f32c AT, struct exo_win_2f32c AT_, struct exo_win_2f32c BT, struct exo_win_2f32c BT_, struct exo_win_2f32 C) {
  exoblas::Call call("exo_ssyr2k_rm_stride_1", exoblas::Call::kSync, ctxt);
  if (!(Uplo == 121 || Uplo == 122)) return (void)call.fail_enum("Uplo", Uplo);
  if (!(Trans == 111 || Trans == 112 || Trans == 113)) return (void)call.fail_enum("Trans", Trans);
  if (N < 0) return (void)call.fail_size("N", N);
  if (K < 0) return (void)call.fail_size("K", K);
  if ((Trans == 111) && A.strides[1] != 1) return (void)call.fail_stride("A", A.strides[1]);
  if ((Trans == 111) && N > 1 && A.strides[0] < K) return (void)call.fail_ld("A", A.strides[0]);
  if ((Trans == 111) && A_.strides[1] != 1) return (void)call.fail_stride("A_", A_.strides[1]);
  if ((Trans == 111) && N > 1 && A_.strides[0] < K) return (void)call.fail_ld("A_", A_.strides[0]);
  if ((Trans == 111) && B.strides[1] != 1) return (void)call.fail_stride("B", B.strides[1]);
  if ((Trans == 111) && N > 1 && B.strides[0] < K) return (void)call.fail_ld("B", B.strides[0]);
  if ((Trans == 111) && B_.strides[1] != 1) return (void)call.fail_stride("B_", B_.strides[1]);
  if ((Trans == 111) && N > 1 && B_.strides[0] < K) return (void)call.fail_ld("B_", B_.strides[0]);
  if ((Trans != 111) && AT.strides[1] != 1) return (void)call.fail_stride("AT", AT.strides[1]);
  if ((Trans != 111) && K > 1 && AT.strides[0] < N) return (void)call.fail_ld("AT", AT.strides[0]);
  if ((Trans != 111) && AT_.strides[1] != 1) return (void)call.fail_stride("AT_", AT_.strides[1]);
  if ((Trans != 111) && K > 1 && AT_.strides[0] < N) return (void)call.fail_ld("AT_", AT_.strides[0]);
  if ((Trans != 111) && BT.strides[1] != 1) return (void)call.fail_stride("BT", BT.strides[1]);
  if ((Trans != 111) && K > 1 && BT.strides[0] < N) return (void)call.fail_ld("BT", BT.strides[0]);
  if ((Trans != 111) && BT_.strides[1] != 1) return (void)call.fail_stride("BT_", BT_.strides[1]);
  if ((Trans != 111) && K > 1 && BT_.strides[0] < N) return (void)call.fail_ld("BT_", BT_.strides[0]);
  if ((true) && C.strides[1] != 1) return (void)call.fail_stride("C", C.strides[1]);
  if ((true) && N > 1 && C.strides[0] < N) return (void)call.fail_ld("C", C.strides[0]);
  call.window(0, A.data, 0, exoblas::span2_hi(N, K, A.strides[0]), sizeof(float), false, Trans == 111);
  call.window(1, A_.data, 0, exoblas::span2_hi(N, K, A_.strides[0]), sizeof(float), false, Trans == 111);
  call.window(2, B.data, 0, exoblas::span2_hi(N, K, B.strides[0]), sizeof(float), false, Trans == 111);
  call.window(3, B_.data, 0, exoblas::span2_hi(N, K, B_.strides[0]), sizeof(float), false, Trans == 111);
  call.window(4, AT.data, 0, exoblas::span2_hi(K, N, AT.strides[0]), sizeof(float), false, Trans != 111);
  call.window(5, AT_.data, 0, exoblas::span2_hi(K, N, AT_.strides[0]), sizeof(float), false, Trans != 111);
  call.window(6, BT.data, 0, exoblas::span2_hi(K, N, BT.strides[0]), sizeof(float), false, Trans != 111);
  call.window(7, BT_.data, 0, exoblas::span2_hi(K, N, BT_.strides[0]), sizeof(float), false, Trans != 111);
  call.window(8, C.data, 0, exoblas::span2_hi(N, N, C.strides[0]), sizeof(float), true, true);
  if (!call.stage()) return (void)call.status();
  const float alpha_v = call.scalar(alpha);
  if (!call.ok()) return (void)call.status();
  call.run([&] { exoblas::ops::syr2k_rm<float>(call.ctx(), (int64_t)Uplo, (int64_t)Trans, (int64_t)N, (int64_t)K, alpha_v, call.dev<const float>(0), (int64_t)A.strides[0], call.dev<const float>(1), (int64_t)A_.strides[0], call.dev<const float>(2), (int64_t)B.strides[0], call.dev<const float>(3), (int64_t)B_.strides[0], call.dev<const float>(4), (int64_t)AT.strides[0], call.dev<const float>(5), (int64_t)AT_.strides[0], call.dev<const float>(6), (int64_t)BT.strides[0], call.dev<const float>(7), (int64_t)BT_.strides[0], call.dev<float>(8), (int64_t)C.strides[0]); });
  return (void)call.finish();
}
extern "C" int exoblas_ssyr2k_rm_stride_1_async(hipStream_t stream, int_fast32_t Uplo, int_fast32_t Trans, int_fast32_t N, int_fast32_t K, const float *alpha, struct exo_win_2f32c A, struct exo_win_2f32c A_, struct exo_win_2f32c B, struct exo_win_2f32c B_, struct exo_win_2f32c AT, struct exo_win_2f32c AT_, struct exo_win_2f32c BT, struct exo_win_2f32c BT_, struct exo_win_2f32 C) {
  exoblas::Call call("exoblas_ssyr2k_rm_stride_1_async", exoblas::Call::kAsync, stream);
  if (!(Uplo == 121 || Uplo == 122)) return call.fail_enum("Uplo", Uplo);
  if (!(Trans == 111 || Trans == 112 || Trans == 113)) return call.fail_enum("Trans", Trans);
  if (N < 0) return call.fail_size("N", N);
  if (K < 0) return call.fail_size("K", K);
  if ((Trans == 111) && A.strides[1] != 1) return call.fail_stride("A", A.strides[1]);
  if ((Trans == 111) && N > 1 && A.strides[0] < K) return call.fail_ld("A", A.strides[0]);
  if ((Trans == 111) && A_.strides[1] != 1) return call.fail_stride("A_", A_.strides[1]);
  if ((Trans == 111) && N > 1 && A_.strides[0] < K) return call.fail_ld("A_", A_.strides[0]);
  if ((Trans == 111) && B.strides[1] != 1) return call.fail_stride("B", B.strides[1]);
  if ((Trans == 111) && N > 1 && B.strides[0] < K) return call.fail_ld("B", B.strides[0]);
  if ((Trans == 111) && B_.strides[1] != 1) return call.fail_stride("B_", B_.strides[1]);
  if ((Trans == 111) && N > 1 && B_.strides[0] < K) return call.fail_ld("B_", B_.strides[0]);
  if ((Trans != 111) && AT.strides[1] != 1) return call.fail_stride("AT", AT.strides[1]);
  if ((Trans != 111) && K > 1 && AT.strides[0] < N) return call.fail_ld("AT", AT.strides[0]);
  if ((Trans != 111) && AT_.strides[1] != 1) return call.fail_stride("AT_", AT_.strides[1]);
  if ((Trans != 111) && K > 1 && AT_.strides[0] < N) return call.fail_ld("AT_", AT_.strides[0]);
  if ((Trans != 111) && BT.strides[1] != 1) return call.fail_stride("BT", BT.strides[1]);
  if ((Trans != 111) && K > 1 && BT.strides[0] < N) return call.fail_ld("BT", BT.strides[0]);
  if ((Trans != 111) && BT_.strides[1] != 1) return call.fail_stride("BT_", BT_.strides[1]);
  if ((Trans != 111) && K > 1 && BT_.strides[0] < N) return call.fail_ld("BT_", BT_.strides[0]);
  if ((true) && C.strides[1] != 1) return call.fail_stride("C", C.strides[1]);
  if ((true) && N > 1 && C.strides[0] < N) return call.fail_ld("C", C.strides[0]);
  call.window(0, A.data, 0, exoblas::span2_hi(N, K, A.strides[0]), sizeof(float), false, Trans == 111);
  call.window(1, A_.data, 0, exoblas::span2_hi(N, K, A_.strides[0]), sizeof(float), false, Trans == 111);
  call.window(2, B.data, 0, exoblas::span2_hi(N, K, B.strides[0]), sizeof(float), false, Trans == 111);
  call.window(3, B_.data, 0, exoblas::span2_hi(N, K, B_.strides[0]), sizeof(float), false, Trans == 111);
  call.window(4, AT.data, 0, exoblas::span2_hi(K, N, AT.strides[0]), sizeof(float), false, Trans != 111);
  call.window(5, AT_.data, 0, exoblas::span2_hi(K, N, AT_.strides[0]), sizeof(float), false, Trans != 111);
  call.window(6, BT.data, 0, exoblas::span2_hi(K, N, BT.strides[0]), sizeof(float), false, Trans != 111);
  call.window(7, BT_.data, 0, exoblas::span2_hi(K, N, BT_.strides[0]), sizeof(float), false, Trans != 111);
  call.window(8, C.data, 0, exoblas::span2_hi(N, N, C.strides[0]), sizeof(float), true, true);
  if (!call.stage()) return call.status();
  const float alpha_v = call.scalar(alpha);
  if (!call.ok()) return call.status();
  call.run([&] { exoblas::ops::syr2k_rm<float>(call.ctx(), (int64_t)Uplo, (int64_t)Trans, (int64_t)N, (int64_t)K, alpha_v, call.dev<const float>(0), (int64_t)A.strides[0], call.dev<const float>(1), (int64_t)A_.strides[0], call.dev<const float>(2), (int64_t)B.strides[0], call.dev<const float>(3), (int64_t)B_.strides[0], call.dev<const float>(4), (int64_t)AT.strides[0], call.dev<const float>(5), (int64_t)AT_.strides[0], call.dev<const float>(6), (int64_t)BT.strides[0], call.dev<const float>(7), (int64_t)BT_.strides[0], call.dev<float>(8), (int64_t)C.strides[0]); });
  return call.finish();
}

// src/level3/syr2k.py:14-39
extern "C" void exo_dsyr2k_rm_stride_any(void *ctxt, int_fast32_t Uplo, int_fast32_t Trans, int_fast32_t N, int_fast32_t K, const double *alpha, struct exo_win_2f64c A, struct exo_win_2f64c A_, struct exo_win_2f64c B, struct exo_win_2f64c B_, struct exo_win_2f64c AT, struct exo_win_2f64c AT_, struct exo_win_2f64c BT, struct exo_win_2f64c BT_, struct exo_win_2f64 C) {
  exoblas::Call call("exo_dsyr2k_rm_stride_any", exoblas::Call::kSync, ctxt);
  if (!(Uplo == 121 || Uplo == 122)) return (void)call.fail_enum("Uplo", Uplo);
  if (!(Trans == 111 || Trans == 112 || Trans == 113)) return (void)call.fail_enum("Trans", Trans);
  if (N < 0) return (void)call.fail_size("N", N);
  if (K < 0) return (void)call.fail_size("K", K);
  if ((Trans == 111) && A.strides[1] != 1) return (void)call.fail_stride("A", A.strides[1]);
  if ((Trans == 111) && N > 1 && A.strides[0] < K) return (void)call.fail_ld("A", A.strides[0]);
  if ((Trans == 111) && A_.strides[1] != 1) return (void)call.fail_stride("A_", A_.strides[1]);
  if ((Trans == 111) && N > 1 && A_.strides[0] < K) return (void)call.fail_ld("A_", A_.strides[0]);
  if ((Trans == 111) && B.strides[1] != 1) return (void)call.fail_stride("B", B.strides[1]);
  if ((Trans == 111) && N > 1 && B.strides[0] < K) return (void)call.fail_ld("B", B.strides[0]);
  if ((Trans == 111) && B_.strides[1] != 1) return (void)call.fail_stride("B_", B_.strides[1]);
  if ((Trans == 111) && N > 1 && B_.strides[0] < K) return (void)call.fail_ld("B_", B_.strides[0]);
  if ((Trans != 111) && AT.strides[1] != 1) return (void)call.fail_stride("AT", AT.strides[1]);
  if ((Trans != 111) && K > 1 && AT.strides[0] < N) return (void)call.fail_ld("AT", AT.strides[0]);
  if ((Trans != 111) && AT_.strides[1] != 1) return (void)call.fail_stride("AT_", AT_.strides[1]);
  if ((Trans != 111) && K > 1 && AT_.strides[0] < N) return (void)call.fail_ld("AT_", AT_.strides[0]);
  if ((Trans != 111) && BT.strides[1] != 1) return (void)call.fail_stride("BT", BT.strides[1]);
  if ((Trans != 111) && K > 1 && BT.strides[0] < N) return (void)call.fail_ld("BT", BT.strides[0]);
  if ((Trans != 111) && BT_.strides[1] != 1) return (void)call.fail_stride("BT_", BT_.strides[1]);
  if ((Trans != 111) && K > 1 && BT_.strides[0] < N) return (void)call.fail_ld("BT_", BT_.strides[0]);
  if ((true) && C.strides[1] != 1) return (void)call.fail_stride("C", C.strides[1]);
  if ((true) && N > 1 && C.strides[0] < N) return (void)call.fail_ld("C", C.strides[0]);
  call.window(0, A.data, 0, exoblas::span2_hi(N, K, A.strides[0]), sizeof(double), false, Trans == 111);
  call.window(1, A_.data, 0, exoblas::span2_hi(N, K, A_.strides[0]), sizeof(double), false, Trans == 111);
  call.window(2, B.data, 0, exoblas::span2_hi(N, K, B.strides[0]), sizeof(double), false, Trans == 111);
  call.window(3, B_.data, 0, exoblas::span2_hi(N, K, B_.strides[0]), sizeof(double), false, Trans == 111);
  call.window(4, AT.data, 0, exoblas::span2_hi(K, N, AT.strides[0]), sizeof(double), false, Trans != 111);
  call.window(5, AT_.data, 0, exoblas::span2_hi(K, N, AT_.strides[0]), sizeof(double), false, Trans != 111);
  call.window(6, BT.data, 0, exoblas::span2_hi(K, N, BT.strides[0]), sizeof(double), false, Trans != 111);
  call.window(7, BT_.data, 0, exoblas::span2_hi(K, N, BT_.strides[0]), sizeof(double), false, Trans != 111);
  call.window(8, C.data, 0, exoblas::span2_hi(N, N, C.strides[0]), sizeof(double), true, true);
  if (!call.stage()) return (void)call.status();
  const double alpha_v = call.scalar(alpha);
  if (!call.ok()) return (void)call.status();
  call.run([&] { exoblas::ops::syr2k_rm<double>(call.ctx(), (int64_t)Uplo, (int64_t)Trans, (int64_t)N, (int64_t)K, alpha_v, call.dev<const double>(0), (int64_t)A.strides[0], call.dev<const double>(1), (int64_t)A_.strides[0], call.dev<const double>(2), (int64_t)B.strides[0], call.dev<const double>(3), (int64_t)B_.strides[0], call.dev<const double>(4), (int64_t)AT.strides[0], call.dev<const double>(5), (int64_t)AT_.strides[0], call.dev<const double>(6), (int64_t)BT.strides[0], call.dev<const double>(7), (int64_t)BT_.strides[0], call.dev<double>(8), (int64_t)C.strides[0]); });
  return (void)call.finish();
}
extern "C" int exoblas_dsyr2k_rm_stride_any_async(hipStream_t stream, int_fast32_t Uplo, int_fast32_t Trans, int_fast32_t N, int_fast32_t K, const double *alpha, struct exo_win_2f64c A, struct exo_win_2f64c A_, struct exo_win_2f64c B, struct exo_win_2f64c B_, struct exo_win_2f64c AT, struct exo_win_2f64c AT_, struct exo_win_2f64c BT, struct exo_win_2f64c BT_, struct exo_win_2f64 C) {
  exoblas::Call call("exoblas_dsyr2k_rm_stride_any_async", exoblas::Call::kAsync, stream);
  if (!(Uplo == 121 || Uplo == 122)) return call.fail_enum("Uplo", Uplo);
  if (!(Trans == 111 || Trans == 112 || Trans == 113)) return call.fail_enum("Trans", Trans);
  if (N < 0) return call.fail_size("N", N);
  if (K < 0) return call.fail_size("K", K);
  if ((Trans == 111) && A.strides[1] != 1) return call.fail_stride("A", A.strides[1]);
  if ((Trans == 111) && N > 1 && A.strides[0] < K) return call.fail_ld("A", A.strides[0]);
  if ((Trans == 111) && A_.strides[1] != 1) return call.fail_stride("A_", A_.strides[1]);
  if ((Trans == 111) && N > 1 && A_.strides[0] < K) return call.fail_ld("A_", A_.strides[0]);
  if ((Trans == 111) && B.strides[1] != 1) return call.fail_stride("B", B.strides[1]);
  if ((Trans == 111) && N > 1 && B.strides[0] < K) return call.fail_ld("B", B.strides[0]);
  if ((Trans == 111) && B_.strides[1] != 1) return call.fail_stride("B_", B_.strides[1]);
  if ((Trans == 111) && N > 1 && B_.strides[0] < K) return call.fail_ld("B_", B_.strides[0]);
  if ((Trans != 111) && AT.strides[1] != 1) return call.fail_stride("AT", AT.strides[1]);
  if ((Trans != 111) && K > 1 && AT.strides[0] < N) return call.fail_ld("AT", AT.strides[0]);
  if ((Trans != 111) && AT_.strides[1] != 1) return call.fail_stride("AT_", AT_.strides[1]);
  if ((Trans != 111) && K > 1 && AT_.strides[0] < N) return call.fail_ld("AT_", AT_.strides[0]);
  if ((Trans != 111) && BT.strides[1] != 1) return call.fail_stride("BT", BT.strides[1]);
  if ((Trans != 111) && K > 1 && BT.strides[0] < N) return call.fail_ld("BT", BT.strides[0]);
  if ((Trans != 111) && BT_.strides[1] != 1) return call.fail_stride("BT_", BT_.strides[1]);
  if ((Trans != 111) && K > 1 && BT_.strides[0] < N) return call.fail_ld("BT_", BT_.strides[0]);
  if ((true) && C.strides[1] != 1) return call.fail_stride("C", C.strides[1]);
  if ((true) && N > 1 && C.strides[0] < N) return call.fail_ld("C", C.strides[0]);
  call.window(0, A.data, 0, exoblas::span2_hi(N, K, A.strides[0]), sizeof(double), false, Trans == 111);
  call.window(1, A_.data, 0, exoblas::span2_hi(N, K, A_.strides[0]), sizeof(double), false, Trans == 111);
  call.window(2, B.data, 0, exoblas::span2_hi(N, K, B.strides[0]), sizeof(double), false, Trans == 111);
  call.window(3, B_.data, 0, exoblas::span2_hi(N, K, B_.strides[0]), sizeof(double), false, Trans == 111);
  call.window(4, AT.data, 0, exoblas::span2_hi(K, N, AT.strides[0]), sizeof(double), false, Trans != 111);
  call.window(5, AT_.data, 0, exoblas::span2_hi(K, N, AT_.strides[0]), sizeof(double), false, Trans != 111);
  call.window(6, BT.data, 0, exoblas::span2_hi(K, N, BT.strides[0]), sizeof(double), false, Trans != 111);
  call.window(7, BT_.data, 0, exoblas::span2_hi(K, N, BT_.strides[0]), sizeof(double), false, Trans != 111);
  call.window(8, C.data, 0, exoblas::span2_hi(N, N, C.strides[0]), sizeof(double), true, true);
  if (!call.stage()) return call.status();
  const double alpha_v = call.scalar(alpha);
  if (!call.ok()) return call.status();
  call.run([&] { exoblas::ops::syr2k_rm<double>(call.ctx(), (int64_t)Uplo, (int64_t)Trans, (int64_t)N, (int64_t)K, alpha_v, call.dev<const double>(0), (int64_t)A.strides[0], call.dev<const double>(1), (int64_t)A_.strides[0], call.dev<const double>(2), (int64_t)B.strides[0], call.dev<const double>(3), (int64_t)B_.strides[0], call.dev<const double>(4), (int64_t)AT.strides[0], call.dev<const double>(5), (int64_t)AT_.strides[0], call.dev<const double>(6), (int64_t)BT.strides[0], call.dev<const double>(7), (int64_t)BT_.strides[0], call.dev<double>(8), (int64_t)C.strides[0]); });
  return call.finish();
}

// src/level3/syr2k.py:14-39
extern "C" void exo_dsyr2k_rm_stride_1(void *ctxt, int_fast32_t Uplo, int_fast32_t Trans, int_fast32_t N, int_fast32_t K, const double *alpha, struct exo_win_2f64c A, struct exo_win_2f64c A_, struct exo_win_2f64c B, struct exo_win_2f64c B_, struct exo_win_2f64c AT, struct exo_win_2f64c AT_, struct exo_win_2f64c BT, struct exo_win_2f64c BT_, struct exo_win_2f64 C) {
  exoblas::Call call("exo_dsyr2k_rm_stride_1", exoblas::Call::kSync, ctxt);
  if (!(Uplo == 121 || Uplo == 122)) return (void)call.fail_enum("Uplo", Uplo);
  if (!(Trans == 111 || Trans == 112 || Trans == 113)) return (void)call.fail_enum("Trans", Trans);
  if (N < 0) return (void)call.fail_size("N", N);
  if (K < 0) return (void)call.fail_size("K", K);
  if ((Trans == 111) && A.strides[1] != 1) return (void)call.fail_stride("A", A.strides[1]);
  if ((Trans == 111) && N > 1 && A.strides[0] < K) return (void)call.fail_ld("A", A.strides[0]);
  if ((Trans == 111) && A_.strides[1] != 1) return (void)call.fail_stride("A_", A_.strides[1]);
  if ((Trans == 111) && N > 1 && A_.strides[0] < K) return (void)call.fail_ld("A_", A_.strides[0]);
  if ((Trans == 111) && B.strides[1] != 1) return (void)call.fail_stride("B", B.strides[1]);
  if ((Trans == 111) && N > 1 && B.strides[0] < K) return (void)call.fail_ld("B", B.strides[0]);
  if ((Trans == 111) && B_.strides[1] != 1) return (void)call.fail_stride("B_", B_.strides[1]);
  if ((Trans == 111) && N > 1 && B_.strides[0] < K) return (void)call.fail_ld("B_", B_.strides[0]);
  if ((Trans != 111) && AT.strides[1] != 1) return (void)call.fail_stride("AT", AT.strides[1]);
  if ((Trans != 111) && K > 1 && AT.strides[0] < N) return (void)call.fail_ld("AT", AT.strides[0]);
  if ((Trans != 111) && AT_.strides[1] != 1) return (void)call.fail_stride("AT_", AT_.strides[1]);
  if ((Trans != 111) && K > 1 && AT_.strides[0] < N) return (void)call.fail_ld("AT_", AT_.strides[0]);
  if ((Trans != 111) && BT.strides[1] != 1) return (void)call.fail_stride("BT", BT.strides[1]);
  if ((Trans != 111) && K > 1 && BT.strides[0] < N) return (void)call.fail_ld("BT", BT.strides[0]);
  if ((Trans != 111) && BT_.strides[1] != 1) return (void)call.fail_stride("BT_", BT_.strides[1]);
  if ((Trans != 111) && K > 1 && BT_.strides[0] < N) return (void)call.fail_ld("BT_", BT_.strides[0]);
  if ((true) && C.strides[1] != 1) return (void)call.fail_stride("C", C.strides[1]);
  if ((true) && N > 1 && C.strides[0] < N) return (void)call.fail_ld("C", C.strides[0]);
  call.window(0, A.data, 0, exoblas::span2_hi(N, K, A.strides[0]), sizeof(double), false, Trans == 111);
  call.window(1, A_.data, 0, exoblas::span2_hi(N, K, A_.strides[0]), sizeof(double), false, Trans == 111);
  call.window(2, B.data, 0, exoblas::span2_hi(N, K, B.strides[0]), sizeof(double), false, Trans == 111);
  call.window(3, B_.data, 0, exoblas::span2_hi(N, K, B_.strides[0]), sizeof(double), false, Trans == 111);
  call.window(4, AT.data, 0, exoblas::span2_hi(K, N, AT.strides[0]), sizeof(double), false, Trans != 111);
  call.window(5, AT_.data, 0, exoblas::span2_hi(K, N, AT_.strides[0]), sizeof(double), false, Trans != 111);
  call.window(6, BT.data, 0, exoblas::span2_hi(K, N, BT.strides[0]), sizeof(double), false, Trans != 111);
  call.window(7, BT_.data, 0, exoblas::span2_hi(K, N, BT_.strides[0]), sizeof(double), false, Trans != 111);
  call.window(8, C.data, 0, exoblas::span2_hi(N, N, C.strides[0]), sizeof(double), true, true);
  if (!call.stage()) return (void)call.status();
  const double alpha_v = call.scalar(alpha);
  if (!call.ok()) return (void)call.status();
  call.run([&] { exoblas::ops::syr2k_rm<double>(call.ctx(), (int64_t)Uplo, (int64_t)Trans, (int64_t)N, (int64_t)K, alpha_v, call.dev<const double>(0), (int64_t)A.strides[0], call.dev<const double>(1), (int64_t)A_.strides[0], call.dev<const double>(2), (int64_t)B.strides[0], call.dev<const double>(3), (int64_t)B_.strides[0], call.dev<const double>(4), (int64_t)AT.strides[0], call.dev<const double>(5), (int64_t)AT_.strides[0], call.dev<const double>(6), (int64_t)BT.strides[0], call.dev<const double>(7), (int64_t)BT_.strides[0], call.dev<double>(8), (int64_t)C.strides[0]); });
  return (void)call.finish();
}
extern "C" int exoblas_dsyr2k_rm_stride_1_async(hipStream_t stream, int_fast32_t Uplo, int_fast32_t Trans, int_fast32_t N, int_fast32_t K, const double *alpha, struct exo_win_2f64c A, struct exo_win_2f64c A_, struct exo_win_2f64c B, struct exo_win_2f64c B_, struct exo_win_2f64c AT, struct exo_win_2f64c AT_, struct exo_win_2f64c BT, struct exo_win_2f64c BT_, struct exo_win_2f64 C) {
  exoblas::Call call("exoblas_dsyr2k_rm_stride_1_async", exoblas::Call::kAsync, stream);
  if (!(Uplo == 121 || Uplo == 122)) return call.fail_enum("Uplo", Uplo);
  if (!(Trans == 111 || Trans == 112 || Trans == 113)) return call.fail_enum("Trans", Trans);
  if (N < 0) return call.fail_size("N", N);
  if (K < 0) return call.fail_size("K", K);
  if ((Trans == 111) && A.strides[1] != 1) return call.fail_stride("A", A.strides[1]);
  if ((Trans == 111) && N > 1 && A.strides[0] < K) return call.fail_ld("A", A.strides[0]);
  if ((Trans == 111) && A_.strides[1] != 1) return call.fail_stride("A_", A_.strides[1]);
  if ((Trans == 111) && N > 1 && A_.strides[0] < K) return call.fail_ld("A_", A_.strides[0]);
  if ((Trans == 111) && B.strides[1] != 1) return call.fail_stride("B", B.strides[1]);
  if ((Trans == 111) && N > 1 && B.strides[0] < K) return call.fail_ld("B", B.strides[0]);
  if ((Trans == 111) && B_.strides[1] != 1) return call.fail_stride("B_", B_.strides[1]);
  if ((Trans == 111) && N > 1 && B_.strides[0] < K) return call.fail_ld("B_", B_.strides[0]);
  if ((Trans != 111) && AT.strides[1] != 1) return call.fail_stride("AT", AT.strides[1]);
  if ((Trans != 111) && K > 1 && AT.strides[0] < N) return call.fail_ld("AT", AT.strides[0]);
  if ((Trans != 111) && AT_.strides[1] != 1) return call.fail_stride("AT_", AT_.strides[1]);
  if ((Trans != 111) && K > 1 && AT_.strides[0] < N) return call.fail_ld("AT_", AT_.strides[0]);
  if ((Trans != 111) && BT.strides[1] != 1) return call.fail_stride("BT", BT.strides[1]);
  if ((Trans != 111) && K > 1 && BT.strides[0] < N) return call.fail_ld("BT", BT.strides[0]);
  if ((Trans != 111) && BT_.strides[1] != 1) return call.fail_stride("BT_", BT_.strides[1]);
  if ((Trans != 111) && K > 1 && BT_.strides[0] < N) return call.fail_ld("BT_", BT_.strides[0]);
  if ((true) && C.strides[1] != 1) return call.fail_stride("C", C.strides[1]);
  if ((true) && N > 1 && C.strides[0] < N) return call.fail_ld("C", C.strides[0]);
  call.window(0, A.data, 0, exoblas::span2_hi(N, K, A.strides[0]), sizeof(double), false, Trans == 111);
  call.window(1, A_.data, 0, exoblas::span2_hi(N, K, A_.strides[0]), sizeof(double), false, Trans == 111);
  call.window(2, B.data, 0, exoblas::span2_hi(N, K, B.strides[0]), sizeof(double), false, Trans == 111);
  call.window(3, B_.data, 0, exoblas::span2_hi(N, K, B_.strides[0]), sizeof(double), false, Trans == 111);
  call.window(4, AT.data, 0, exoblas::span2_hi(K, N, AT.strides[0]), sizeof(double), false, Trans != 111);
  call.window(5, AT_.data, 0, exoblas::span2_hi(K, N, AT_.strides[0]), sizeof(double), false, Trans != 111);
  call.window(6, BT.data, 0, exoblas::span2_hi(K, N, BT.strides[0]), sizeof(double), false, Trans != 111);
  call.window(7, BT_.data, 0, exoblas::span2_hi(K, N, BT_.strides[0]), sizeof(double), false, Trans != 111);
  call.window(8, C.data, 0, exoblas::span2_hi(N, N, C.strides[0]), sizeof(double), true, true);
  if (!call.stage()) return call.status();
  const double alpha_v = call.scalar(alpha);
  if (!call.ok()) return call.status();
  call.run([&] { exoblas::ops::syr2k_rm<double>(call.ctx(), (int64_t)Uplo, (int64_t)Trans, (int64_t)N, (int64_t)K, alpha_v, call.dev<const double>(0), (int64_t)A.strides[0], call.dev<const double>(1), (int64_t)A_.strides[0], call.dev<const double>(2), (int64_t)B.strides[0], call.dev<const double>(3), (int64_t)B_.strides[0], call.dev<const double>(4), (int64_t)AT.strides[0], call.dev<const double>(5), (int64_t)AT_.strides[0], call.dev<const double>(6), (int64_t)BT.strides[0], call.dev<const double>(7), (int64_t)BT_.strides[0], call.dev<double>(8), (int64_t)C.strides[0]); });
  return call.finish();
}

// src/level3/trmm.py:12-23
extern "C" void exo_strmm_row_major_Left_Upper_NonTrans(void *ctxt, int_fast32_t m, int_fast32_t n, const float *alpha, struct exo_win_2f32c A, struct exo_win_2f32 B, int_fast32_t Diag) {
  exoblas::Call call("exo_strmm_row_major_Left_Upper_NonTrans", exoblas::Call::kSync, ctxt);
  if (m < 0) return (void)call.fail_size("m", m);
  if (n < 0) return (void)call.fail_size("n", n);
  if (Diag < 0) return (void)call.fail_size("Diag", Diag);
  if ((true) && A.strides[1] != 1) return (void)call.fail_stride("A", A.strides[1]);
  if ((true) && m > 1 && A.strides[0] < m) return (void)call.fail_ld("A", A.strides[0]);
  if ((true) && B.strides[1] != 1) return (void)call.fail_stride("B", B.strides[1]);
  if ((true) && m > 1 && B.strides[0] < n) return (void)call.fail_ld("B", B.strides[0]);
  call.window(0, A.data, 0, exoblas::span2_hi(m, m, A.strides[0]), sizeof(float), false, true);
  call.window(1, B.data, 0, exoblas::span2_hi(m, n, B.strides[0]), sizeof(float), true, true);
  if (!call.stage()) return (void)call.status();
  const float alpha_v = call.scalar(alpha);
  if (!call.ok()) return (void)call.status();
  call.run([&] { exoblas::ops::trmm_row_major_Left_Upper_NonTrans<float>(call.ctx(), (int64_t)m, (int64_t)n, alpha_v, call.dev<const float>(0), (int64_t)A.strides[0], call.dev<float>(1), (int64_t)B.strides[0], (int64_t)Diag); });
  return (void)call.finish();
}
extern "C" int exoblas_strmm_row_major_Left_Upper_NonTrans_async(hipStream_t stream, int_fast32_t m, int_fast32_t n, const float *alpha, struct exo_win_2f32c A, struct exo_win_2f32 B, int_fast32_t Diag) {
  exoblas::Call call("exoblas_strmm_row_major_Left_Upper_NonTrans_async", exoblas::Call::kAsync, stream);
  if (m < 0) return call.fail_size("m", m);
  if (n < 0) return call.fail_size("n", n);
  if (Diag < 0) return call.fail_size("Diag", Diag);
  if ((true) && A.strides[1] != 1) return call.fail_stride("A", A.strides[1]);
  if ((true) && m > 1 && A.strides[0] < m) return call.fail_ld("A", A.strides[0]);
  if ((true) && B.strides[1] != 1) return call.fail_stride("B", B.strides[1]);
  if ((true) && m > 1 && B.strides[0] < n) return call.fail_ld("B", B.strides[0]);
  call.window(0, A.data, 0, exoblas::span2_hi(m, m, A.strides[0]), sizeof(float), false, true);
  call.window(1, B.data, 0, exoblas::span2_hi(m, n, B.strides[0]), sizeof(float), true, true);
  if (!call.stage()) return call.status();
  const float alpha_v = call.scalar(alpha);
  if (!call.ok()) return call.status();
  call.run([&] { exoblas::ops::trmm_row_major_Left_Upper_NonTrans<float>(call.ctx(), (int64_t)m, (int64_t)n, alpha_v, call.dev<const float>(0), (int64_t)A.strides[0], call.dev<float>(1), (int64_t)B.strides[0], (int64_t)Diag); });
  return call.finish();
}

// src/level3/trmm.py:26-46
extern "C" void exo_strmm_row_major_Left_Lower_NonTrans(void *ctxt, int_fast32_t m, int_fast32_t n, const float *alpha, struct exo_win_2f32c A, struct exo_win_2f32 B, int_fast32_t Diag) {
  exoblas::Call call("exo_strmm_row_major_Left_Lower_NonTrans", exoblas::Call::kSync, ctxt);
  if (m < 0) return (void)call.fail_size("m", m);
  if (n < 0) return (void)call.fail_size("n", n);
  if (Diag < 0) return (void)call.fail_size("Diag", Diag);
  if ((true) && A.strides[1] != 1) return (void)call.fail_stride("A", A.strides[1]);
  if ((true) && m > 1 && A.strides[0] < m) return (void)call.fail_ld("A", A.strides[0]);
  if ((true) && B.strides[1] != 1) return (void)call.fail_stride("B", B.strides[1]);
  if ((true) && m > 1 && B.strides[0] < n) return (void)call.fail_ld("B", B.strides[0]);
  call.window(0, A.data, 0, exoblas::span2_hi(m, m, A.strides[0]), sizeof(float), false, true);
  call.window(1, B.data, 0, exoblas::span2_hi(m, n, B.strides[0]), sizeof(float), true, true);
  if (!call.stage()) return (void)call.status();
  const float alpha_v = call.scalar(alpha);
  if (!call.ok()) return (void)call.status();
  call.run([&] { exoblas::ops::trmm_row_major_Left_Lower_NonTrans<float>(call.ctx(), (int64_t)m, (int64_t)n, alpha_v, call.dev<const float>(0), (int64_t)A.strides[0], call.dev<float>(1), (int64_t)B.strides[0], (int64_t)Diag); });
  return (void)call.finish();
}
extern "C" int exoblas_strmm_row_major_Left_Lower_NonTrans_async(hipStream_t stream, int_fast32_t m, int_fast32_t n, const float *alpha, struct exo_win_2f32c A, struct exo_win_2f32 B, int_fast32_t Diag) {
  exoblas::Call call("exoblas_strmm_row_major_Left_Lower_NonTrans_async", exoblas::Call::kAsync, stream);
  if (m < 0) return call.fail_size("m", m);
  if (n < 0) return call.fail_size("n", n);
  if (Diag < 0) return call.fail_size("Diag", Diag);
  if ((true) && A.strides[1] != 1) return call.fail_stride("A", A.strides[1]);
  if ((true) && m > 1 && A.strides[0] < m) return call.fail_ld("A", A.strides[0]);
  if ((true) && B.strides[1] != 1) return call.fail_stride("B", B.strides[1]);
  if ((true) && m > 1 && B.strides[0] < n) return call.fail_ld("B", B.strides[0]);
  call.window(0, A.data, 0, exoblas::span2_hi(m, m, A.strides[0]), sizeof(float), false, true);
  call.window(1, B.data, 0, exoblas::span2_hi(m, n, B.strides[0]), sizeof(float), true, true);
  if (!call.stage()) return call.status();
  const float alpha_v = call.scalar(alpha);
  if (!call.ok()) return call.status();
  call.run([&] { exoblas::ops::trmm_row_major_Left_Lower_NonTrans<float>(call.ctx(), (int64_t)m, (int64_t)n, alpha_v, call.dev<const float>(0), (int64_t)A.strides[0], call.dev<float>(1), (int64_t)B.strides[0], (int64_t)Diag); });
  return call.finish();
}

// src/level3/trmm.py:49-69
extern "C" void exo_strmm_row_major_Left_Upper_Trans(void *ctxt, int_fast32_t m, int_fast32_t n, const float *alpha, struct exo_win_2f32c A, struct exo_win_2f32 B, int_fast32_t Diag) {
  exoblas::Call call("exo_strmm_row_major_Left_Upper_Trans", exoblas::Call::kSync, ctxt);
  if (m < 0) return (void)call.fail_size("m", m);
  if (n < 0) return (void)call.fail_size("n", n);
  if (Diag < 0) return (void)call.fail_size("Diag", Diag);
  if ((true) && A.strides[1] != 1) return (void)call.fail_stride("A", A.strides[1]);
  if ((true) && m > 1 && A.strides[0] < m) return (void)call.fail_ld("A", A.strides[0]);
  if ((true) && B.strides[1] != 1) return (void)call.fail_stride("B", B.strides[1]);
  if ((true) && m > 1 && B.strides[0] < n) return (void)call.fail_ld("B", B.strides[0]);
  call.window(0, A.data, 0, exoblas::span2_hi(m, m, A.strides[0]), sizeof(float), false, true);
  call.window(1, B.data, 0, exoblas::span2_hi(m, n, B.strides[0]), sizeof(float), true, true);
  if (!call.stage()) return (void)call.status();
  const float alpha_v = call.scalar(alpha);
  if (!call.ok()) return (void)call.status();
  call.run([&] { exoblas::ops::trmm_row_major_Left_Upper_Trans<float>(call.ctx(), (int64_t)m, (int64_t)n, alpha_v, call.dev<const float>(0), (int64_t)A.strides[0], call.dev<float>(1), (int64_t)B.strides[0], (int64_t)Diag); });
  return (void)call.finish();
}
extern "C" int exoblas_strmm_row_major_Left_Upper_Trans_async(hipStream_t stream, int_fast32_t m, int_fast32_t n, const float *alpha, struct exo_win_2f32c A, struct exo_win_2f32 B, int_fast32_t Diag) {
  exoblas::Call call("exoblas_strmm_row_major_Left_Upper_Trans_async", exoblas::Call::kAsync, stream);
  if (m < 0) return call.fail_size("m", m);
  if (n < 0) return call.fail_size("n", n);
  if (Diag < 0) return call.fail_size("Diag", Diag);
  if ((true) && A.strides[1] != 1) return call.fail_stride("A", A.strides[1]);
  if ((true) && m > 1 && A.strides[0] < m) return call.fail_ld("A", A.strides[0]);
  if ((true) && B.strides[1] != 1) return call.fail_stride("B", B.strides[1]);
  if ((true) && m > 1 && B.strides[0] < n) return call.fail_ld("B", B.strides[0]);
  call.window(0, A.data, 0, exoblas::span2_hi(m, m, A.strides[0]), sizeof(float), false, true);
  call.window(1, B.data, 0, exoblas::span2_hi(m, n, B.strides[0]), sizeof(float), true, true);
  if (!call.stage()) return call.status();
  const float alpha_v = call.scalar(alpha);
  if (!call.ok()) return call.status();
  call.run([&] { exoblas::ops::trmm_row_major_Left_Upper_Trans<float>(call.ctx(), (int64_t)m, (int64_t)n, alpha_v, call.dev<const float>(0), (int64_t)A.strides[0], call.dev<float>(1), (int64_t)B.strides[0], (int64_t)Diag); });
  return call.finish();
}

// src/level3/trmm.py:72-83
extern "C" void exo_strmm_row_major_Left_Lower_Trans(void *ctxt, int_fast32_t m, int_fast32_t n, const float *alpha, struct exo_win_2f32c A, struct exo_win_2f32 B, int_fast32_t Diag) {
  exoblas::Call call("exo_strmm_row_major_Left_Lower_Trans", exoblas::Call::kSync, ctxt);
  if (m < 0) return (void)call.fail_size("m", m);
  if (n < 0) return (void)call.fail_size("n", n);
  if (Diag < 0) return (void)call.fail_size("Diag", Diag);
  if ((true) && A.strides[1] != 1) return (void)call.fail_stride("A", A.strides[1]);
  if ((true) && m > 1 && A.strides[0] < m) return (void)call.fail_ld("A", A.strides[0]);
  if ((true) && B.strides[1] != 1) return (void)call.fail_stride("B", B.strides[1]);
  if ((true) && m > 1 && B.strides[0] < n) return (void)call.fail_ld("B", B.strides[0]);
  call.window(0, A.data, 0, exoblas::span2_hi(m, m, A.strides[0]), sizeof(float), false, true);
  call.window(1, B.data, 0, exoblas::span2_hi(m, n, B.strides[0]), sizeof(float), true, true);
  if (!call.stage()) return (void)call.status();
  const float alpha_v = call.scalar(alpha);
  if (!call.ok()) return (void)call.status();
  call.run([&] { exoblas::ops::trmm_row_major_Left_Lower_Trans<float>(call.ctx(), (int64_t)m, (int64_t)n, alpha_v, call.dev<const float>(0), (int64_t)A.strides[0], call.dev<float>(1), (int64_t)B.strides[0], (int64_t)Diag); });
  return (void)call.finish();
}
extern "C" int exoblas_strmm_row_major_Left_Lower_Trans_async(hipStream_t stream, int_fast32_t m, int_fast32_t n, const float *alpha, struct exo_win_2f32c A, struct exo_win_2f32 B, int_fast32_t Diag) {
  exoblas::Call call("exoblas_strmm_row_major_Left_Lower_Trans_async", exoblas::Call::kAsync, stream);
  if (m < 0) return call.fail_size("m", m);
  if (n < 0) return call.fail_size("n", n);
  if (Diag < 0) return call.fail_size("Diag", Diag);
  if ((true) && A.strides[1] != 1) return call.fail_stride("A", A.strides[1]);
  if ((true) && m > 1 && A.strides[0] < m) return call.fail_ld("A", A.strides[0]);
  if ((true) && B.strides[1] != 1) return call.fail_stride("B", B.strides[1]);
  if ((true) && m > 1 && B.strides[0] < n) return call.fail_ld("B", B.strides[0]);
  call.window(0, A.data, 0, exoblas::span2_hi(m, m, A.strides[0]), sizeof(float), false, true);
  call.window(1, B.data, 0, exoblas::span2_hi(m, n, B.strides[0]), sizeof(float), true, true);
  if (!call.stage()) return call.status();
  const float alpha_v = call.scalar(alpha);
  if (!call.ok()) return call.status();
  call.run([&] { exoblas::ops::trmm_row_major_Left_Lower_Trans<float>(call.ctx(), (int64_t)m, (int64_t)n, alpha_v, call.dev<const float>(0), (int64_t)A.strides[0], call.dev<float>(1), (int64_t)B.strides[0], (int64_t)Diag); });
  return call.finish();
}

// C += alpha op(A) op(B) with FP32 operands split into BF16 terms on the BF16 MFMA; arguments as exo_sgemm_stride_1 (src/level3/gemm.py:17-47) plus the number of BF16 products
extern "C" int exoblas_sgemm_split(void *ctxt, int_fast32_t TransA, int_fast32_t TransB, int_fast32_t M, int_fast32_t N, int_fast32_t K, const float *alpha, struct exo_win_2f32c A, struct exo_win_2f32c B, struct exo_win_2f32c AT, struct exo_win_2f32c BT, struct exo_win_2f32 C, int_fast32_t terms) {
  exoblas::Call call("exoblas_sgemm_split", exoblas::Call::kSync, ctxt);
  if (!(TransA == 111 || TransA == 112 || TransA == 113)) return call.fail_enum("TransA", TransA);
  if (!(TransB == 111 || TransB == 112 || TransB == 113)) return call.fail_enum("TransB", TransB);
  if (M < 0) return call.fail_size("M", M);
  if (N < 0) return call.fail_size("N", N);
  if (K < 0) return call.fail_size("K", K);
  if (!(terms == 3 || terms == 4 || terms == 6)) return call.fail_enum("terms", terms);
  if ((TransA == 111) && A.strides[1] != 1) return call.fail_stride("A", A.strides[1]);
  if ((TransA == 111) && M > 1 && A.strides[0] < K) return call.fail_ld("A", A.strides[0]);
  if ((TransB == 111) && B.strides[1] != 1) return call.fail_stride("B", B.strides[1]);
  if ((TransB == 111) && K > 1 && B.strides[0] < N) return call.fail_ld("B", B.strides[0]);
  if ((TransA != 111) && AT.strides[1] != 1) return call.fail_stride("AT", AT.strides[1]);
  if ((TransA != 111) && K > 1 && AT.strides[0] < M) return call.fail_ld("AT", AT.strides[0]);
  if ((TransB != 111) && BT.strides[1] != 1) return call.fail_stride("BT", BT.strides[1]);
  if ((TransB != 111) && N > 1 && BT.strides[0] < K) return call.fail_ld("BT", BT.strides[0]);
  if ((true) && C.strides[1] != 1) return call.fail_stride("C", C.strides[1]);
  if ((true) && M > 1 && C.strides[0] < N) return call.fail_ld("C", C.strides[0]);
  call.window(0, A.data, 0, exoblas::span2_hi(M, K, A.strides[0]), sizeof(float), false, TransA == 111);
  call.window(1, B.data, 0, exoblas::span2_hi(K, N, B.strides[0]), sizeof(float), false, TransB == 111);
  call.window(2, AT.data, 0, exoblas::span2_hi(K, M, AT.strides[0]), sizeof(float), false, TransA != 111);
  call.window(3, BT.data, 0, exoblas::span2_hi(N, K, BT.strides[0]), sizeof(float), false, TransB != 111);
  call.window(4, C.data, 0, exoblas::span2_hi(M, N, C.strides[0]), sizeof(float), true, true);
  if (!call.stage()) return call.status();
  const float alpha_v = call.scalar(alpha);
  if (!call.ok()) return call.status();
  call.run([&] { exoblas::ops::gemm_split(call.ctx(), (int64_t)TransA, (int64_t)TransB, (int64_t)M, (int64_t)N, (int64_t)K, alpha_v, call.dev<const float>(0), (int64_t)A.strides[0], call.dev<const float>(1), (int64_t)B.strides[0], call.dev<const float>(2), (int64_t)AT.strides[0], call.dev<const float>(3), (int64_t)BT.strides[0], call.dev<float>(4), (int64_t)C.strides[0], (int64_t)terms); });
  return call.finish();
}
extern "C" int exoblas_sgemm_split_async(hipStream_t stream, int_fast32_t TransA, int_fast32_t TransB, int_fast32_t M, int_fast32_t N, int_fast32_t K, const float *alpha, struct exo_win_2f32c A, struct exo_win_2f32c B, struct exo_win_2f32c AT, struct exo_win_2f32c BT, struct exo_win_2f32 C, int_fast32_t terms) {
  exoblas::Call call("exoblas_sgemm_split_async", exoblas::Call::kAsync, stream);
  if (!(TransA == 111 || TransA == 112 || TransA == 113)) return call.fail_enum("TransA", TransA);
  if (!(TransB == 111 || TransB == 112 || TransB == 113)) return call.fail_enum("TransB", TransB);
  if (M < 0) return call.fail_size("M", M);
  if (N < 0) return call.fail_size("N", N);
  if (K < 0) return call.fail_size("K", K);
  if (!(terms == 3 || terms == 4 || terms == 6)) return call.fail_enum("terms", terms);
  if ((TransA == 111) && A.strides[1] != 1) return call.fail_stride("A", A.strides[1]);
  if ((TransA == 111) && M > 1 && A.strides[0] < K) return call.fail_ld("A", A.strides[0]);
  if ((TransB == 111) && B.strides[1] != 1) return call.fail_stride("B", B.strides[1]);
  if ((TransB == 111) && K > 1 && B.strides[0] < N) return call.fail_ld("B", B.strides[0]);
  if ((TransA != 111) && AT.strides[1] != 1) return call.fail_stride("AT", AT.strides[1]);
  if ((TransA != 111) && K > 1 && AT.strides[0] < M) return call.fail_ld("AT", AT.strides[0]);
  if ((TransB != 111) && BT.strides[1] != 1) return call.fail_stride("BT", BT.strides[1]);
  if ((TransB != 111) && N > 1 && BT.strides[0] < K) return call.fail_ld("BT", BT.strides[0]);
  if ((true) && C.strides[1] != 1) return call.fail_stride("C", C.strides[1]);
  if ((true) && M > 1 && C.strides[0] < N) return call.fail_ld("C", C.strides[0]);
  call.window(0, A.data, 0, exoblas::span2_hi(M, K, A.strides[0]), sizeof(float), false, TransA == 111);
  call.window(1, B.data, 0, exoblas::span2_hi(K, N, B.strides[0]), sizeof(float), false, TransB == 111);
  call.window(2, AT.data, 0, exoblas::span2_hi(K, M, AT.strides[0]), sizeof(float), false, TransA != 111);
  call.window(3, BT.data, 0, exoblas::span2_hi(N, K, BT.strides[0]), sizeof(float), false, TransB != 111);
  call.window(4, C.data, 0, exoblas::span2_hi(M, N, C.strides[0]), sizeof(float), true, true);
  if (!call.stage()) return call.status();
  const float alpha_v = call.scalar(alpha);
  if (!call.ok()) return call.status();
  call.run([&] { exoblas::ops::gemm_split(call.ctx(), (int64_t)TransA, (int64_t)TransB, (int64_t)M, (int64_t)N, (int64_t)K, alpha_v, call.dev<const float>(0), (int64_t)A.strides[0], call.dev<const float>(1), (int64_t)B.strides[0], call.dev<const float>(2), (int64_t)AT.strides[0], call.dev<const float>(3), (int64_t)BT.strides[0], call.dev<float>(4), (int64_t)C.strides[0], (int64_t)terms); });
  return call.finish();
}

// one triangle of C += alpha A A^T (or A^T A) in the split BF16 form; arguments as exo_ssyrk_rm_stride_1 (src/level3/syrk.py:17-32) without the duplicate windows
extern "C" int exoblas_ssyrk_split(void *ctxt, int_fast32_t Uplo, int_fast32_t Trans, int_fast32_t N, int_fast32_t K, const float *alpha, struct exo_win_2f32c A, struct exo_win_2f32c AT, struct exo_win_2f32 C, int_fast32_t terms) {
  exoblas::Call call("exoblas_ssyrk_split", exoblas::Call::kSync, ctxt);
  if (!(Uplo == 121 || Uplo == 122)) return call.fail_enum("Uplo", Uplo);
  if (!(Trans == 111 || Trans == 112 || Trans == 113)) return call.fail_enum("Trans", Trans);
  if (N < 0) return call.fail_size("N", N);
  if (K < 0) return call.fail_size("K", K);
  if (!(terms == 3 || terms == 4 || terms == 6)) return call.fail_enum("terms", terms);
  if ((Trans == 111) && A.strides[1] != 1) return call.fail_stride("A", A.strides[1]);
  if ((Trans == 111) && N > 1 && A.strides[0] < K) return call.fail_ld("A", A.strides[0]);
  if ((Trans != 111) && AT.strides[1] != 1) return call.fail_stride("AT", AT.strides[1]);
  if ((Trans != 111) && K > 1 && AT.strides[0] < N) return call.fail_ld("AT", AT.strides[0]);
  if ((true) && C.strides[1] != 1) return call.fail_stride("C", C.strides[1]);
  if ((true) && N > 1 && C.strides[0] < N) return call.fail_ld("C", C.strides[0]);
  call.window(0, A.data, 0, exoblas::span2_hi(N, K, A.strides[0]), sizeof(float), false, Trans == 111);
  call.window(1, AT.data, 0, exoblas::span2_hi(K, N, AT.strides[0]), sizeof(float), false, Trans != 111);
  call.window(2, C.data, 0, exoblas::span2_hi(N, N, C.strides[0]), sizeof(float), true, true);
  if (!call.stage()) return call.status();
  const float alpha_v = call.scalar(alpha);
  if (!call.ok()) return call.status();
  call.run([&] { exoblas::ops::syrk_split(call.ctx(), (int64_t)Uplo, (int64_t)Trans, (int64_t)N, (int64_t)K, alpha_v, call.dev<const float>(0), (int64_t)A.strides[0], call.dev<const float>(1), (int64_t)AT.strides[0], call.dev<float>(2), (int64_t)C.strides[0], (int64_t)terms); });
  return call.finish();
}
extern "C" int exoblas_ssyrk_split_async(hipStream_t stream, int_fast32_t Uplo, int_fast32_t Trans, int_fast32_t N, int_fast32_t K, const float *alpha, struct exo_win_2f32c A, struct exo_win_2f32c AT, struct exo_win_2f32 C, int_fast32_t terms) {
  exoblas::Call call("exoblas_ssyrk_split_async", exoblas::Call::kAsync, stream);
  if (!(Uplo == 121 || Uplo == 122)) return call.fail_enum("Uplo", Uplo);
  if (!(Trans == 111 || Trans == 112 || Trans == 113)) return call.fail_enum("Trans", Trans);
  if (N < 0) return call.fail_size("N", N);
  if (K < 0) return call.fail_size("K", K);
  if (!(terms == 3 || terms == 4 || terms == 6)) return call.fail_enum("terms", terms);
  if ((Trans == 111) && A.strides[1] != 1) return call.fail_stride("A", A.strides[1]);
  if ((Trans == 111) && N > 1 && A.strides[0] < K) return call.fail_ld("A", A.strides[0]);
  if ((Trans != 111) && AT.strides[1] != 1) return call.fail_stride("AT", AT.strides[1]);
  if ((Trans != 111) && K > 1 && AT.strides[0] < N) return call.fail_ld("AT", AT.strides[0]);
  if ((true) && C.strides[1] != 1) return call.fail_stride("C", C.strides[1]);
  if ((true) && N > 1 && C.strides[0] < N) return call.fail_ld("C", C.strides[0]);
  call.window(0, A.data, 0, exoblas::span2_hi(N, K, A.strides[0]), sizeof(float), false, Trans == 111);
  call.window(1, AT.data, 0, exoblas::span2_hi(K, N, AT.strides[0]), sizeof(float), false, Trans != 111);
  call.window(2, C.data, 0, exoblas::span2_hi(N, N, C.strides[0]), sizeof(float), true, true);
  if (!call.stage()) return call.status();
  const float alpha_v = call.scalar(alpha);
  if (!call.ok()) return call.status();
  call.run([&] { exoblas::ops::syrk_split(call.ctx(), (int64_t)Uplo, (int64_t)Trans, (int64_t)N, (int64_t)K, alpha_v, call.dev<const float>(0), (int64_t)A.strides[0], call.dev<const float>(1), (int64_t)AT.strides[0], call.dev<float>(2), (int64_t)C.strides[0], (int64_t)terms); });
  return call.finish();
}

// C = alpha op(A) op(B) + beta C with CBLAS dimensions in row-major (Order 101) or column-major (102) storage, C not read when beta == 0, beta applied in the product's epilogue; each window is the stored array (rows x leading dimension); the product as exo_<p>gemm_stride_1 (src/level3/gemm.py:17-47)
extern "C" int exoblas_sgemm_ex(void *ctxt, int_fast32_t Order, int_fast32_t TransA, int_fast32_t TransB, int_fast32_t M, int_fast32_t N, int_fast32_t K, const float *alpha, struct exo_win_2f32c A, struct exo_win_2f32c B, const float *beta, struct exo_win_2f32 C) {
  exoblas::Call call("exoblas_sgemm_ex", exoblas::Call::kSync, ctxt);
  if (!(Order == 101 || Order == 102)) return call.fail_enum("Order", Order);
  if (!(TransA == 111 || TransA == 112 || TransA == 113)) return call.fail_enum("TransA", TransA);
  if (!(TransB == 111 || TransB == 112 || TransB == 113)) return call.fail_enum("TransB", TransB);
  if (M < 0) return call.fail_size("M", M);
  if (N < 0) return call.fail_size("N", N);
  if (K < 0) return call.fail_size("K", K);
  if ((true) && A.strides[1] != 1) return call.fail_stride("A", A.strides[1]);
  if ((true) && ((Order == 101) == (TransA == 111) ? M : K) > 1 && A.strides[0] < ((Order == 101) == (TransA == 111) ? K : M)) return call.fail_ld("A", A.strides[0]);
  if ((true) && B.strides[1] != 1) return call.fail_stride("B", B.strides[1]);
  if ((true) && ((Order == 101) == (TransB == 111) ? K : N) > 1 && B.strides[0] < ((Order == 101) == (TransB == 111) ? N : K)) return call.fail_ld("B", B.strides[0]);
  if ((true) && C.strides[1] != 1) return call.fail_stride("C", C.strides[1]);
  if ((true) && (Order == 101 ? M : N) > 1 && C.strides[0] < (Order == 101 ? N : M)) return call.fail_ld("C", C.strides[0]);
  call.window(0, A.data, 0, exoblas::span2_hi(((Order == 101) == (TransA == 111) ? M : K), ((Order == 101) == (TransA == 111) ? K : M), A.strides[0]), sizeof(float), false, true);
  call.window(1, B.data, 0, exoblas::span2_hi(((Order == 101) == (TransB == 111) ? K : N), ((Order == 101) == (TransB == 111) ? N : K), B.strides[0]), sizeof(float), false, true);
  call.window(2, C.data, 0, exoblas::span2_hi((Order == 101 ? M : N), (Order == 101 ? N : M), C.strides[0]), sizeof(float), true, true);
  if (!call.stage()) return call.status();
  const float alpha_v = call.scalar(alpha);
  const float beta_v = call.scalar(beta);
  if (!call.ok()) return call.status();
  call.run([&] { exoblas::ops::gemm_ex(call.ctx(), (int64_t)Order, (int64_t)TransA, (int64_t)TransB, (int64_t)M, (int64_t)N, (int64_t)K, alpha_v, call.dev<const float>(0), (int64_t)A.strides[0], call.dev<const float>(1), (int64_t)B.strides[0], beta_v, call.dev<float>(2), (int64_t)C.strides[0]); });
  return call.finish();
}
extern "C" int exoblas_sgemm_ex_async(hipStream_t stream, int_fast32_t Order, int_fast32_t TransA, int_fast32_t TransB, int_fast32_t M, int_fast32_t N, int_fast32_t K, const float *alpha, struct exo_win_2f32c A, struct exo_win_2f32c B, const float *beta, struct exo_win_2f32 C) {
  exoblas::Call call("exoblas_sgemm_ex_async", exoblas::Call::kAsync, stream);
  if (!(Order == 101 || Order == 102)) return call.fail_enum("Order", Order);
  if (!(TransA == 111 || TransA == 112 || TransA == 113)) return call.fail_enum("TransA", TransA);
  if (!(TransB == 111 || TransB == 112 || TransB == 113)) return call.fail_enum("TransB", TransB);
  if (M < 0) return call.fail_size("M", M);
  if (N < 0) return call.fail_size("N", N);
  if (K < 0) return call.fail_size("K", K);
  if ((true) && A.strides[1] != 1) return call.fail_stride("A", A.strides[1]);
  if ((true) && ((Order == 101) == (TransA == 111) ? M : K) > 1 && A.strides[0] < ((Order == 101) == (TransA == 111) ? K : M)) return call.fail_ld("A", A.strides[0]);
  if ((true) && B.strides[1] != 1) return call.fail_stride("B", B.strides[1]);
  if ((true) && ((Order == 101) == (TransB == 111) ? K : N) > 1 && B.strides[0] < ((Order == 101) == (TransB == 111) ? N : K)) return call.fail_ld("B", B.strides[0]);
  if ((true) && C.strides[1] != 1) return call.fail_stride("C", C.strides[1]);
  if ((true) && (Order == 101 ? M : N) > 1 && C.strides[0] < (Order == 101 ? N : M)) return call.fail_ld("C", C.strides[0]);
  call.window(0, A.data, 0, exoblas::span2_hi(((Order == 101) == (TransA == 111) ? M : K), ((Order == 101) == (TransA == 111) ? K : M), A.strides[0]), sizeof(float), false, true);
  call.window(1, B.data, 0, exoblas::span2_hi(((Order == 101) == (TransB == 111) ? K : N), ((Order == 101) == (TransB == 111) ? N : K), B.strides[0]), sizeof(float), false, true);
  call.window(2, C.data, 0, exoblas::span2_hi((Order == 101 ? M : N), (Order == 101 ? N : M), C.strides[0]), sizeof(float), true, true);
  if (!call.stage()) return call.status();
  const float alpha_v = call.scalar(alpha);
  const float beta_v = call.scalar(beta);
  if (!call.ok()) return call.status();
  call.run([&] { exoblas::ops::gemm_ex(call.ctx(), (int64_t)Order, (int64_t)TransA, (int64_t)TransB, (int64_t)M, (int64_t)N, (int64_t)K, alpha_v, call.dev<const float>(0), (int64_t)A.strides[0], call.dev<const float>(1), (int64_t)B.strides[0], beta_v, call.dev<float>(2), (int64_t)C.strides[0]); });
  return call.finish();
}

// C = alpha op(A) op(B) + beta C with CBLAS dimensions in row-major (Order 101) or column-major (102) storage, C not read when beta == 0, beta applied in the product's epilogue; each window is the stored array (rows x leading dimension); the product as exo_<p>gemm_stride_1 (src/level3/gemm.py:17-47)
extern "C" int exoblas_dgemm_ex(void *ctxt, int_fast32_t Order, int_fast32_t TransA, int_fast32_t TransB, int_fast32_t M, int_fast32_t N, int_fast32_t K, const double *alpha, struct exo_win_2f64c A, struct exo_win_2f64c B, const double *beta, struct exo_win_2f64 C) {
  exoblas::Call call("exoblas_dgemm_ex", exoblas::Call::kSync, ctxt);
  if (!(Order == 101 || Order == 102)) return call.fail_enum("Order", Order);
  if (!(TransA == 111 || TransA == 112 || TransA == 113)) return call.fail_enum("TransA", TransA);
  if (!(TransB == 111 || TransB == 112 || TransB == 113)) return call.fail_enum("TransB", TransB);
  if (M < 0) return call.fail_size("M", M);
  if (N < 0) return call.fail_size("N", N);
  if (K < 0) return call.fail_size("K", K);
  if ((true) && A.strides[1] != 1) return call.fail_stride("A", A.strides[1]);
  if ((true) && ((Order == 101) == (TransA == 111) ? M : K) > 1 && A.strides[0] < ((Order == 101) == (TransA == 111) ? K : M)) return call.fail_ld("A", A.strides[0]);
  if ((true) && B.strides[1] != 1) return call.fail_stride("B", B.strides[1]);
  if ((true) && ((Order == 101) == (TransB == 111) ? K : N) > 1 && B.strides[0] < ((Order == 101) == (TransB == 111) ? N : K)) return call.fail_ld("B", B.strides[0]);
  if ((true) && C.strides[1] != 1) return call.fail_stride("C", C.strides[1]);
  if ((true) && (Order == 101 ? M : N) > 1 && C.strides[0] < (Order == 101 ? N : M)) return call.fail_ld("C", C.strides[0]);
  call.window(0, A.data, 0, exoblas::span2_hi(((Order == 101) == (TransA == 111) ? M : K), ((Order == 101) == (TransA == 111) ? K : M), A.strides[0]), sizeof(double), false, true);
  call.window(1, B.data, 0, exoblas::span2_hi(((Order == 101) == (TransB == 111) ? K : N), ((Order == 101) == (TransB == 111) ? N : K), B.strides[0]), sizeof(double), false, true);
  call.window(2, C.data, 0, exoblas::span2_hi((Order == 101 ? M : N), (Order == 101 ? N : M), C.strides[0]), sizeof(double), true, true);
  if (!call.stage()) return call.status();
  const double alpha_v = call.scalar(alpha);
  const double beta_v = call.scalar(beta);
  if (!call.ok()) return call.status();
  call.run([&] { exoblas::ops::gemm_ex(call.ctx(), (int64_t)Order, (int64_t)TransA, (int64_t)TransB, (int64_t)M, (int64_t)N, (int64_t)K, alpha_v, call.dev<const double>(0), (int64_t)A.strides[0], call.dev<const double>(1), (int64_t)B.strides[0], beta_v, call.dev<double>(2), (int64_t)C.strides[0]); });
  return call.finish();
}
extern "C" int exoblas_dgemm_ex_async(hipStream_t stream, int_fast32_t Order, int_fast32_t TransA, int_fast32_t TransB, int_fast32_t M, int_fast32_t N, int_fast32_t K, const double *alpha, struct exo_win_2f64c A, struct exo_win_2f64c B, const double *beta, struct exo_win_2f64 C) {
  exoblas::Call call("exoblas_dgemm_ex_async", exoblas::Call::kAsync, stream);
  if (!(Order == 101 || Order == 102)) return call.fail_enum("Order", Order);
  if (!(TransA == 111 || TransA == 112 || TransA == 113)) return call.fail_enum("TransA", TransA);
  if (!(TransB == 111 || TransB == 112 || TransB == 113)) return call.fail_enum("TransB", TransB);
  if (M < 0) return call.fail_size("M", M);
  if (N < 0) return call.fail_size("N", N);
  if (K < 0) return call.fail_size("K", K);
  if ((true) && A.strides[1] != 1) return call.fail_stride("A", A.strides[1]);
  if ((true) && ((Order == 101) == (TransA == 111) ? M : K) > 1 && A.strides[0] < ((Order == 101) == (TransA == 111) ? K : M)) return call.fail_ld("A", A.strides[0]);
  if ((true) && B.strides[1] != 1) return call.fail_stride("B", B.strides[1]);
  if ((true) && ((Order == 101) == (TransB == 111) ? K : N) > 1 && B.strides[0] < ((Order == 101) == (TransB == 111) ? N : K)) return call.fail_ld("B", B.strides[0]);
  if ((true) && C.strides[1] != 1) return call.fail_stride("C", C.strides[1]);
  if ((true) && (Order == 101 ? M : N) > 1 && C.strides[0] < (Order == 101 ? N : M)) return call.fail_ld("C", C.strides[0]);
  call.window(0, A.data, 0, exoblas::span2_hi(((Order == 101) == (TransA == 111) ? M : K), ((Order == 101) == (TransA == 111) ? K : M), A.strides[0]), sizeof(double), false, true);
  call.window(1, B.data, 0, exoblas::span2_hi(((Order == 101) == (TransB == 111) ? K : N), ((Order == 101) == (TransB == 111) ? N : K), B.strides[0]), sizeof(double), false, true);
  call.window(2, C.data, 0, exoblas::span2_hi((Order == 101 ? M : N), (Order == 101 ? N : M), C.strides[0]), sizeof(double), true, true);
  if (!call.stage()) return call.status();
  const double alpha_v = call.scalar(alpha);
  const double beta_v = call.scalar(beta);
  if (!call.ok()) return call.status();
  call.run([&] { exoblas::ops::gemm_ex(call.ctx(), (int64_t)Order, (int64_t)TransA, (int64_t)TransB, (int64_t)M, (int64_t)N, (int64_t)K, alpha_v, call.dev<const double>(0), (int64_t)A.strides[0], call.dev<const double>(1), (int64_t)B.strides[0], beta_v, call.dev<double>(2), (int64_t)C.strides[0]); });
  return call.finish();
}

// y = alpha op(A) x + beta y for an M x N matrix A in row-major (Order 101) or column-major (102) storage, y not read when beta == 0; element i of x at x.data + i * x.strides[0]; the product as exo_<p>gemv_rm_stride_any (src/level2/gemv.py:10-22)
extern "C" int exoblas_sgemv_ex(void *ctxt, int_fast32_t Order, int_fast32_t TransA, int_fast32_t M, int_fast32_t N, const float *alpha, struct exo_win_2f32c A, struct exo_win_1f32c x, const float *beta, struct exo_win_1f32 y) {
  exoblas::Call call("exoblas_sgemv_ex", exoblas::Call::kSync, ctxt);
  if (!(Order == 101 || Order == 102)) return call.fail_enum("Order", Order);
  if (!(TransA == 111 || TransA == 112 || TransA == 113)) return call.fail_enum("TransA", TransA);
  if (M < 0) return call.fail_size("M", M);
  if (N < 0) return call.fail_size("N", N);
  if ((true) && A.strides[1] != 1) return call.fail_stride("A", A.strides[1]);
  if ((true) && (Order == 101 ? M : N) > 1 && A.strides[0] < (Order == 101 ? N : M)) return call.fail_ld("A", A.strides[0]);
  call.window(0, A.data, 0, exoblas::span2_hi((Order == 101 ? M : N), (Order == 101 ? N : M), A.strides[0]), sizeof(float), false, true);
  call.window(1, x.data, exoblas::span_lo((TransA == 111 ? N : M), x.strides[0]), exoblas::span_hi((TransA == 111 ? N : M), x.strides[0]), sizeof(float), false, true);
  call.window(2, y.data, exoblas::span_lo((TransA == 111 ? M : N), y.strides[0]), exoblas::span_hi((TransA == 111 ? M : N), y.strides[0]), sizeof(float), true, true);
  if (!call.stage()) return call.status();
  const float alpha_v = call.scalar(alpha);
  const float beta_v = call.scalar(beta);
  if (!call.ok()) return call.status();
  call.run([&] { exoblas::ops::gemv_ex(call.ctx(), (int64_t)Order, (int64_t)TransA, (int64_t)M, (int64_t)N, alpha_v, call.dev<const float>(0), (int64_t)A.strides[0], call.dev<const float>(1), (int64_t)x.strides[0], beta_v, call.dev<float>(2), (int64_t)y.strides[0]); });
  return call.finish();
}
extern "C" int exoblas_sgemv_ex_async(hipStream_t stream, int_fast32_t Order, int_fast32_t TransA, int_fast32_t M, int_fast32_t N, const float *alpha, struct exo_win_2f32c A, struct exo_win_1f32c x, const float *beta, struct exo_win_1f32 y) {
  exoblas::Call call("exoblas_sgemv_ex_async", exoblas::Call::kAsync, stream);
  if (!(Order == 101 || Order == 102)) return call.fail_enum("Order", Order);
  if (!(TransA == 111 || TransA == 112 || TransA == 113)) return call.fail_enum("TransA", TransA);
  if (M < 0) return call.fail_size("M", M);
  if (N < 0) return call.fail_size("N", N);
  if ((true) && A.strides[1] != 1) return call.fail_stride("A", A.strides[1]);
  if ((true) && (Order == 101 ? M : N) > 1 && A.strides[0] < (Order == 101 ? N : M)) return call.fail_ld("A", A.strides[0]);
  call.window(0, A.data, 0, exoblas::span2_hi((Order == 101 ? M : N), (Order == 101 ? N : M), A.strides[0]), sizeof(float), false, true);
  call.window(1, x.data, exoblas::span_lo((TransA == 111 ? N : M), x.strides[0]), exoblas::span_hi((TransA == 111 ? N : M), x.strides[0]), sizeof(float), false, true);
  call.window(2, y.data, exoblas::span_lo((TransA == 111 ? M : N), y.strides[0]), exoblas::span_hi((TransA == 111 ? M : N), y.strides[0]), sizeof(float), true, true);
  if (!call.stage()) return call.status();
  const float alpha_v = call.scalar(alpha);
  const float beta_v = call.scalar(beta);
  if (!call.ok()) return call.status();
  call.run([&] { exoblas::ops::gemv_ex(call.ctx(), (int64_t)Order, (int64_t)TransA, (int64_t)M, (int64_t)N, alpha_v, call.dev<const float>(0), (int64_t)A.strides[0], call.dev<const float>(1), (int64_t)x.strides[0], beta_v, call.dev<float>(2), (int64_t)y.strides[0]); });
  return call.finish();
}

// y = alpha op(A) x + beta y for an M x N matrix A in row-major (Order 101) or column-major (102) storage, y not read when beta == 0; element i of x at x.data + i * x.strides[0]; the product as exo_<p>gemv_rm_stride_any (src/level2/gemv.py:10-22)
extern "C" int exoblas_dgemv_ex(void *ctxt, int_fast32_t Order, int_fast32_t TransA, int_fast32_t M, int_fast32_t N, const double *alpha, struct exo_win_2f64c A, struct exo_win_1f64c x, const double *beta, struct exo_win_1f64 y) {
  exoblas::Call call("exoblas_dgemv_ex", exoblas::Call::kSync, ctxt);
  if (!(Order == 101 || Order == 102)) return call.fail_enum("Order", Order);
  if (!(TransA == 111 || TransA == 112 || TransA == 113)) return call.fail_enum("TransA", TransA);
  if (M < 0) return call.fail_size("M", M);
  if (N < 0) return call.fail_size("N", N);
  if ((true) && A.strides[1] != 1) return call.fail_stride("A", A.strides[1]);
  if ((true) && (Order == 101 ? M : N) > 1 && A.strides[0] < (Order == 101 ? N : M)) return call.fail_ld("A", A.strides[0]);
  call.window(0, A.data, 0, exoblas::span2_hi((Order == 101 ? M : N), (Order == 101 ? N : M), A.strides[0]), sizeof(double), false, true);
  call.window(1, x.data, exoblas::span_lo((TransA == 111 ? N : M), x.strides[0]), exoblas::span_hi((TransA == 111 ? N : M), x.strides[0]), sizeof(double), false, true);
  call.window(2, y.data, exoblas::span_lo((TransA == 111 ? M : N), y.strides[0]), exoblas::span_hi((TransA == 111 ? M : N), y.strides[0]), sizeof(double), true, true);
  if (!call.stage()) return call.status();
  const double alpha_v = call.scalar(alpha);
  const double beta_v = call.scalar(beta);
  if (!call.ok()) return call.status();
  call.run([&] { exoblas::ops::gemv_ex(call.ctx(), (int64_t)Order, (int64_t)TransA, (int64_t)M, (int64_t)N, alpha_v, call.dev<const double>(0), (int64_t)A.strides[0], call.dev<const double>(1), (int64_t)x.strides[0], beta_v, call.dev<double>(2), (int64_t)y.strides[0]); });
  return call.finish();
}
extern "C" int exoblas_dgemv_ex_async(hipStream_t stream, int_fast32_t Order, int_fast32_t TransA, int_fast32_t M, int_fast32_t N, const double *alpha, struct exo_win_2f64c A, struct exo_win_1f64c x, const double *beta, struct exo_win_1f64 y) {
  exoblas::Call call("exoblas_dgemv_ex_async", exoblas::Call::kAsync, stream);
  if (!(Order == 101 || Order == 102)) return call.fail_enum("Order", Order);
  if (!(TransA == 111 || TransA == 112 || TransA == 113)) return call.fail_enum("TransA", TransA);
  if (M < 0) return call.fail_size("M", M);
  if (N < 0) return call.fail_size("N", N);
  if ((true) && A.strides[1] != 1) return call.fail_stride("A", A.strides[1]);
  if ((true) && (Order == 101 ? M : N) > 1 && A.strides[0] < (Order == 101 ? N : M)) return call.fail_ld("A", A.strides[0]);
  call.window(0, A.data, 0, exoblas::span2_hi((Order == 101 ? M : N), (Order == 101 ? N : M), A.strides[0]), sizeof(double), false, true);
  call.window(1, x.data, exoblas::span_lo((TransA == 111 ? N : M), x.strides[0]), exoblas::span_hi((TransA == 111 ? N : M), x.strides[0]), sizeof(double), false, true);
  call.window(2, y.data, exoblas::span_lo((TransA == 111 ? M : N), y.strides[0]), exoblas::span_hi((TransA == 111 ? M : N), y.strides[0]), sizeof(double), true, true);
  if (!call.stage()) return call.status();
  const double alpha_v = call.scalar(alpha);
  const double beta_v = call.scalar(beta);
  if (!call.ok()) return call.status();
  call.run([&] { exoblas::ops::gemv_ex(call.ctx(), (int64_t)Order, (int64_t)TransA, (int64_t)M, (int64_t)N, alpha_v, call.dev<const double>(0), (int64_t)A.strides[0], call.dev<const double>(1), (int64_t)x.strides[0], beta_v, call.dev<double>(2), (int64_t)y.strides[0]); });
  return call.finish();
}

// op(A) x = b with BF16 (update 16) or FP32 (update 32) off-diagonal updates, FP64 diagonal-block solves and `refine` steps of iterative refinement on an FP64 residual; arguments as exo_strsv_rm_stride_any (src/level2/trsv.py:10-52) plus update and refine
extern "C" int exoblas_strsv_mixed(void *ctxt, int_fast32_t Uplo, int_fast32_t TransA, int_fast32_t Diag, int_fast32_t n, struct exo_win_1f32 x, struct exo_win_2f32c A, int_fast32_t update, int_fast32_t refine) {
  exoblas::Call call("exoblas_strsv_mixed", exoblas::Call::kSync, ctxt);
  if (!(Uplo == 121 || Uplo == 122)) return call.fail_enum("Uplo", Uplo);
  if (!(TransA == 111 || TransA == 112 || TransA == 113)) return call.fail_enum("TransA", TransA);
  if (!(Diag == 131 || Diag == 132)) return call.fail_enum("Diag", Diag);
  if (n < 0) return call.fail_size("n", n);
  if (!(update == 16 || update == 32)) return call.fail_enum("update", update);
  if (refine < 0) return call.fail_size("refine", refine);
  if ((true) && A.strides[1] != 1) return call.fail_stride("A", A.strides[1]);
  if ((true) && n > 1 && A.strides[0] < n) return call.fail_ld("A", A.strides[0]);
  call.window(0, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(float), true, true);
  call.window(1, A.data, 0, exoblas::span2_hi(n, n, A.strides[0]), sizeof(float), false, true);
  if (!call.stage()) return call.status();
  if (!call.ok()) return call.status();
  call.run([&] { exoblas::ops::trsv_mixed(call.ctx(), (int64_t)Uplo, (int64_t)TransA, (int64_t)Diag, (int64_t)n, call.dev<float>(0), (int64_t)x.strides[0], call.dev<const float>(1), (int64_t)A.strides[0], (int64_t)update, (int64_t)refine); });
  return call.finish();
}
extern "C" int exoblas_strsv_mixed_async(hipStream_t stream, int_fast32_t Uplo, int_fast32_t TransA, int_fast32_t Diag, int_fast32_t n, struct exo_win_1f32 x, struct exo_win_2f32c A, int_fast32_t update, int_fast32_t refine) {
  exoblas::Call call("exoblas_strsv_mixed_async", exoblas::Call::kAsync, stream);
  if (!(Uplo == 121 || Uplo == 122)) return call.fail_enum("Uplo", Uplo);
  if (!(TransA == 111 || TransA == 112 || TransA == 113)) return call.fail_enum("TransA", TransA);
  if (!(Diag == 131 || Diag == 132)) return call.fail_enum("Diag", Diag);
  if (n < 0) return call.fail_size("n", n);
  if (!(update == 16 || update == 32)) return call.fail_enum("update", update);
  if (refine < 0) return call.fail_size("refine", refine);
  if ((true) && A.strides[1] != 1) return call.fail_stride("A", A.strides[1]);
  if ((true) && n > 1 && A.strides[0] < n) return call.fail_ld("A", A.strides[0]);
  call.window(0, x.data, exoblas::span_lo(n, x.strides[0]), exoblas::span_hi(n, x.strides[0]), sizeof(float), true, true);
  call.window(1, A.data, 0, exoblas::span2_hi(n, n, A.strides[0]), sizeof(float), false, true);
  if (!call.stage()) return call.status();
  if (!call.ok()) return call.status();
  call.run([&] { exoblas::ops::trsv_mixed(call.ctx(), (int64_t)Uplo, (int64_t)TransA, (int64_t)Diag, (int64_t)n, call.dev<float>(0), (int64_t)x.strides[0], call.dev<const float>(1), (int64_t)A.strides[0], (int64_t)update, (int64_t)refine); });
  return call.finish();
}

