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

// Value g of the publication area `pub` (8-byte aligned), once published.
template <class T>
__device__ __forceinline__ T read(unsigned* pub, int64_t g, unsigned marker, unsigned* err) {
  uint64_t* w64 = reinterpret_cast<uint64_t*>(pub);
  if constexpr (sizeof(T) == 4) {
    return __uint_as_float(poll_word(w64 + g, marker, err));
  } else {
    // both halves loaded before either is checked: one round trip, not two
    const uint64_t w0 = __hip_atomic_load(w64 + 2 * g, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const uint64_t w1 = __hip_atomic_load(w64 + 2 * g + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const unsigned lo = (w0 >> 32) == marker ? (unsigned)w0 : poll_word(w64 + 2 * g, marker, err);
    const unsigned hi = (w1 >> 32) == marker ? (unsigned)w1 : poll_word(w64 + 2 * g + 1, marker, err);
    return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
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
  unsigned la[W], lb[W];
#pragma unroll
  for (int h = 0; h < W; ++h) {
    la[h] = v0 ? ((wa[h] >> 32) == marker ? (unsigned)wa[h] : poll_word(a + h, marker, err)) : 0u;
    lb[h] = v1 ? ((wb[h] >> 32) == marker ? (unsigned)wb[h] : poll_word(b + h, marker, err)) : 0u;
  }
  if constexpr (W == 1) {
    x0 = __uint_as_float(la[0]);
    x1 = __uint_as_float(lb[0]);
  } else {
    x0 = __longlong_as_double((long long)(((uint64_t)la[1] << 32) | la[0]));
    x1 = __longlong_as_double((long long)(((uint64_t)lb[1] << 32) | lb[0]));
  }
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
#pragma unroll
    for (int h = 0; h < W; ++h)
      lo[h] = (w[k][h] >> 32) == marker ? (unsigned)w[k][h]
                                        : poll_word(w64 + W * (g0 + k * stride) + h, marker, err);
    if constexpr (W == 1)
      v[k] = __uint_as_float(lo[0]);
    else
      v[k] = __longlong_as_double((long long)(((uint64_t)lo[1] << 32) | lo[0]));
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
