// Device-side helpers shared by the gfx950 kernels: 16-byte vector types,
// wave64 and workgroup reductions, launch helpers.
#pragma once

#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdint>

#include "runtime.hpp"

namespace exoblas {

constexpr int kWave = 64;

// 16-byte vector of T: the unit of a global_load_dwordx4 / global_store_dwordx4.
template <class T> struct Vec16;
template <> struct Vec16<float> {
  using type = float4;
  static constexpr int n = 4;
};
template <> struct Vec16<double> {
  using type = double2;
  static constexpr int n = 2;
};

template <class T>
__device__ __forceinline__ T& vget(typename Vec16<T>::type& v, int i) {
  return reinterpret_cast<T*>(&v)[i];
}
template <class T>
__device__ __forceinline__ const T& vget(const typename Vec16<T>::type& v, int i) {
  return reinterpret_cast<const T*>(&v)[i];
}

// Streaming 16-byte access (global_load/store_dwordx4 with the non-temporal
// cache policy) for data a call touches once: it keeps the lines from
// displacing reused ones in L2 and MALL, and measured on MI355X it moves
// read streams from 6.45 to 7.15 TB/s and axpy from 5.8 to 6.65 TB/s.
template <class T> struct Native16;
template <> struct Native16<float> { typedef float type __attribute__((ext_vector_type(4))); };
template <> struct Native16<double> { typedef double type __attribute__((ext_vector_type(2))); };
template <class T>
__device__ __forceinline__ typename Vec16<T>::type ld_stream(const T* p) {
  using N = typename Native16<T>::type;
  return __builtin_bit_cast(typename Vec16<T>::type, __builtin_nontemporal_load(reinterpret_cast<const N*>(p)));
}
template <class T>
__device__ __forceinline__ void st_stream(T* p, const typename Vec16<T>::type& v) {
  using N = typename Native16<T>::type;
  __builtin_nontemporal_store(__builtin_bit_cast(N, v), reinterpret_cast<N*>(p));
}

template <class T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
  for (int off = kWave / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
  return v;
}

// Sum over a workgroup of NT threads; the result is valid in every thread.
// `red` must hold NT / 64 elements of LDS.
template <int NT, class T>
__device__ __forceinline__ T block_sum(T v, T* red) {
  v = wave_sum(v);
  const int w = threadIdx.x / kWave;
  if ((threadIdx.x & (kWave - 1)) == 0) red[w] = v;
  __syncthreads();
  T s = 0;
#pragma unroll
  for (int i = 0; i < NT / kWave; ++i) s += red[i];
  __syncthreads();
  return s;
}

inline bool aligned16(const void* p) { return (reinterpret_cast<uintptr_t>(p) & 15) == 0; }

// Raises kernel fn's dynamic-LDS limit to `bytes` (needed above 64 KiB) once
// per device; `done` is the kernel's own mask of devices already set.  Safe
// to call from several host threads (a repeated set is harmless).
inline void ensure_lds_limit(const void* fn, size_t bytes, int device, std::atomic<uint64_t>& done) {
  const uint64_t bit = 1ull << (device & 63);
  if (done.load(std::memory_order_acquire) & bit) return;
  EXO_HIP_CHECK(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
  done.fetch_or(bit, std::memory_order_acq_rel);
}

__host__ __device__ inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

}  // namespace exoblas
