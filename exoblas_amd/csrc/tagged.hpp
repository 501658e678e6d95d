// Tagged publication between the workgroups of one launch.
//
// Every 32 bits of a published value travel with the call's marker in one
// 64-bit word (marker << 32 | bits), stored and polled with relaxed
// device-scope atomics: a reader that sees the marker has those bits, so no
// fence or separate ready flag is needed.  One word per FP32 value, two (low
// and high halves) per FP64 value.  Markers are the per-stream call epochs
// of Ctx::epoch_flags, so words left by earlier calls never match.
//
// Waits are bounded: past the spin limit (kSpinLimit polls unless the host
// set another through set_spin_limit) the reader records its marker in *err,
// the stream's timeout word in pinned host memory, and goes on, so no grid
// can hang the device and the host learns that the results are not valid.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "runtime.hpp"

namespace exoblas {
namespace tagged {

constexpr long long kSpinLimit = 1ll << 22;

// This unit's copy of the spin limit, and its setter, registered with the
// runtime at load time.
static __constant__ long long g_spin_limit = kSpinLimit;
namespace {
int set_unit_spin_limit(long long polls) {
  const long long v = polls < 0 ? kSpinLimit : polls;
  return hipMemcpyToSymbol(HIP_SYMBOL(g_spin_limit), &v, sizeof(v)) == hipSuccess ? 0 : 1;
}
[[maybe_unused]] const bool g_spin_registered = (register_spin_setter(&set_unit_spin_limit), true);
}  // namespace

// 32-bit units of publication space for n values of T.
template <class T>
__host__ __device__ inline int64_t words(int64_t n) {
  return 2 * n * (int64_t)(sizeof(T) / 4);
}

// Polls one tagged word until it carries `marker` (bounded) and returns its
// low 32 bits.
__device__ __forceinline__ unsigned poll_word(uint64_t* w64, unsigned marker, unsigned* err) {
  uint64_t w;
  long long it = 0;
  const long long limit = g_spin_limit;
  while (((w = __hip_atomic_load(w64, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) >> 32) != marker) {
    if (++it > limit) {
      __hip_atomic_store(err, marker ? marker : 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
      break;
    }
    __builtin_amdgcn_s_sleep(1);
  }
  return (unsigned)w;
}

// Polls the two words of one FP64 value (published together) until both
// carry `marker` (bounded), reloading both on every poll: polling one and
// then checking the other's earlier load would cost a second round trip
// whenever the value was not yet published at the first load.
__device__ __forceinline__ void poll_pair(uint64_t* w64, unsigned marker, unsigned* err, unsigned& lo,
                                          unsigned& hi) {
  uint64_t w0, w1;
  long long it = 0;
  const long long limit = g_spin_limit;
  while (true) {
    w0 = __hip_atomic_load(w64, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    w1 = __hip_atomic_load(w64 + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if ((w0 >> 32) == marker && (w1 >> 32) == marker) break;
    if (++it > limit) {
      __hip_atomic_store(err, marker ? marker : 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
      break;
    }
    __builtin_amdgcn_s_sleep(1);
  }
  lo = (unsigned)w0;
  hi = (unsigned)w1;
}

// The low 32 bits of the W words at w64 (one value: W = 1 FP32, 2 FP64),
// given their first loads w[]: those that already carry `marker` are used
// as loaded, otherwise the value is polled until it is whole.
template <int W>
__device__ __forceinline__ void settle(uint64_t* w64, const uint64_t (&w)[W], unsigned marker, unsigned* err,
                                       unsigned (&lo)[W]) {
  if constexpr (W == 1) {
    lo[0] = (w[0] >> 32) == marker ? (unsigned)w[0] : poll_word(w64, marker, err);
  } else {
    if ((w[0] >> 32) == marker && (w[1] >> 32) == marker) {
      lo[0] = (unsigned)w[0];
      lo[1] = (unsigned)w[1];
    } else {
      poll_pair(w64, marker, err, lo[0], lo[1]);
    }
  }
}

// The value whose 32-bit pieces (low first) are lo[].
template <class T, int W>
__device__ __forceinline__ T compose(const unsigned (&lo)[W]) {
  if constexpr (W == 1)
    return __uint_as_float(lo[0]);
  else
    return __longlong_as_double((long long)(((uint64_t)lo[1] << 32) | lo[0]));
}

// Value g of the publication area `pub` (8-byte aligned), once published.
template <class T>
__device__ __forceinline__ T read(unsigned* pub, int64_t g, unsigned marker, unsigned* err) {
  uint64_t* w64 = reinterpret_cast<uint64_t*>(pub);
  if constexpr (sizeof(T) == 4) {
    return __uint_as_float(poll_word(w64 + g, marker, err));
  } else {
    // both halves loaded before either is checked: one round trip, not two
    const uint64_t w[2] = {__hip_atomic_load(w64 + 2 * g, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT),
                           __hip_atomic_load(w64 + 2 * g + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)};
    unsigned lo[2];
    settle<2>(w64 + 2 * g, w, marker, err, lo);
    return compose<T, 2>(lo);
  }
}

// Value g0 of pub0 and g1 of pub1 (either may be skipped: v0 / v1 false
// reads as 0): every word is loaded before any is checked, so when both are
// already published the pair costs one memory round trip, not two.
template <class T>
__device__ __forceinline__ void read2(unsigned* pub0, int64_t g0, bool v0, unsigned* pub1, int64_t g1, bool v1,
                                      unsigned marker, unsigned* err, T& x0, T& x1) {
  constexpr int W = sizeof(T) / 4;
  uint64_t* a = reinterpret_cast<uint64_t*>(pub0) + W * g0;
  uint64_t* b = reinterpret_cast<uint64_t*>(pub1) + W * g1;
  uint64_t wa[W], wb[W];
#pragma unroll
  for (int h = 0; h < W; ++h) {
    if (v0) wa[h] = __hip_atomic_load(a + h, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (v1) wb[h] = __hip_atomic_load(b + h, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  unsigned la[W] = {}, lb[W] = {};
  if (v0) settle<W>(a, wa, marker, err, la);
  if (v1) settle<W>(b, wb, marker, err, lb);
  x0 = compose<T, W>(la);
  x1 = compose<T, W>(lb);
}

// Values g0, g0 + stride, ..., g0 + (K-1) stride of `pub` (those with index
// k < count) into v[k]: all words are loaded before any is checked, so the
// K reads overlap instead of forming a chain of memory round trips.
template <class T, int K>
__device__ __forceinline__ void read_n(unsigned* pub, int64_t g0, int64_t stride, int count, T (&v)[K],
                                       unsigned marker, unsigned* err) {
  uint64_t* w64 = reinterpret_cast<uint64_t*>(pub);
  constexpr int W = sizeof(T) / 4;  // words per value
  uint64_t w[K][W];
#pragma unroll
  for (int k = 0; k < K; ++k)
#pragma unroll
    for (int h = 0; h < W; ++h)
      if (k < count) w[k][h] = __hip_atomic_load(w64 + W * (g0 + k * stride) + h, __ATOMIC_RELAXED,
                                                 __HIP_MEMORY_SCOPE_AGENT);
#pragma unroll
  for (int k = 0; k < K; ++k) {
    if (k >= count) continue;
    unsigned lo[W];
    settle<W>(w64 + W * (g0 + k * stride), w[k], marker, err, lo);
    v[k] = compose<T, W>(lo);
  }
}

// Publishes value g of `pub`.
template <class T>
__device__ __forceinline__ void publish(unsigned* pub, int64_t g, T v, unsigned marker) {
  uint64_t* w64 = reinterpret_cast<uint64_t*>(pub);
  const uint64_t tag = (uint64_t)marker << 32;
  if constexpr (sizeof(T) == 4) {
    __hip_atomic_store(w64 + g, tag | __float_as_uint(v), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  } else {
    const uint64_t bits = (uint64_t)__double_as_longlong(v);
    __hip_atomic_store(w64 + 2 * g, tag | (bits & 0xffffffffu), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(w64 + 2 * g + 1, tag | (bits >> 32), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

// Marker of a launch drawn from the stream's launch generation *gen (top
// bit set, so it never equals a host epoch).  Every workgroup of the launch
// reads it and takes its launch-wide ticket only after a store that uses the
// value (its own publication), so the read is complete before the ticket;
// the workgroup that takes the last ticket advances the generation, so all
// workgroups see one value and the next launch (or graph replay) another.
__device__ __forceinline__ unsigned read_generation(unsigned* gen) {
  return 0x80000000u | (__hip_atomic_load(gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) & 0x7fffffffu);
}
__device__ __forceinline__ void advance_generation(unsigned* gen) { atomicAdd(gen, 1u); }

// Reads the launch generation, then takes a ticket from counter *t shared
// by `count` workgroups (call from one thread).  The empty asm needs the
// loaded value in its register and orders memory operations, so the read is
// complete before the ticket is taken; the workgroup with the last ticket
// resets *t and advances the generation (every workgroup has read it).
// Returns the ticket; *marker receives the launch's marker.
__device__ __forceinline__ unsigned generation_ticket(unsigned* t, unsigned count, unsigned* gen, unsigned* marker) {
  unsigned g = __hip_atomic_load(gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  asm volatile("" : "+v"(g)::"memory");
  *marker = 0x80000000u | (g & 0x7fffffffu);
  const unsigned p = atomicAdd(t, 1u);
  if (p == count - 1) {
    atomicExch(t, 0u);
    advance_generation(gen);
  }
  return p;
}

// Takes a ticket from counter *t shared by `count` workgroups; true for the
// last one, which also resets the counter for the next call on the stream.
// Call from one thread.
__device__ __forceinline__ bool last_ticket(unsigned* t, unsigned count) {
  const unsigned p = atomicAdd(t, 1u);
  if (p == count - 1) {
    atomicExch(t, 0u);
    return true;
  }
  return false;
}

}  // namespace tagged
}  // namespace exoblas
