/*
 * handoff.h -- data hand-off between wavefronts of one launch without flags
 * or fences (triangular solves, ILU(0)).
 *
 * Every slot a consumer reads is preset to a signalling NaN that names the
 * call (the sentinel); the producer stores its results with agent-scope
 * atomic stores, the consumer loads them with agent-scope atomic loads and
 * retries the ones that still hold the sentinel.  Only single-location
 * coherence is relied on -- the value read is the value used -- so no
 * release/acquire (L2 write-back and invalidate on the 8 XCDs) is needed,
 * and the data arrives in the same round trip that signals it.  Arithmetic
 * never yields a signalling NaN and every published result is canonicalised,
 * so a result never equals a sentinel.
 */
#ifndef KF_HANDOFF_H
#define KF_HANDOFF_H

#include "kf_dev.h"

/* sentinels: signalling NaNs (quiet bit clear, top payload bit set so the
 * payload is never zero) carrying the low bits of the sweep number k */
template <typename TV> struct sentinel;
template <> struct sentinel<double> {
  typedef unsigned long long bits;
  static __host__ __device__ bits of(uint64_t k) { return 0x7ff4000000000000ull | (k & 0x0003ffffffffffffull); }
  static __device__ bits raw(double v) { return (bits)__double_as_longlong(v); }
  static __device__ double val(bits b) { return __longlong_as_double((long long)b); }
};
template <> struct sentinel<float> {
  typedef unsigned int bits;
  static __host__ __device__ bits of(uint64_t k) { return 0x7fa00000u | (uint32_t)(k & 0x1fffffu); }
  static __device__ bits raw(float v) { return __float_as_uint(v); }
  static __device__ float val(bits b) { return __uint_as_float(b); }
};

template <typename TV>
__device__ __forceinline__ TV ld_coherent(const TV *p)
{
  return __hip_atomic_load(const_cast<TV *>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

template <typename TV>
__device__ __forceinline__ void st_coherent(TV *p, TV v)
{
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

/* canonical value (quiets a signalling NaN, identity otherwise) */
__device__ __forceinline__ double canon(double v) { return __builtin_canonicalize(v); }
__device__ __forceinline__ float canon(float v) { return __builtin_canonicalizef(v); }

template <typename TV>
__global__ void k_fill_bits(TV *v, size_t n, typename sentinel<TV>::bits s)
{
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) v[i] = sentinel<TV>::val(s);
}

#endif
