// common.h -- device helpers shared by every HIP translation unit.
//
// Ordered keys: a floating-point depth is mapped to an unsigned integer whose
// unsigned order equals the float order, so a z-buffer can be a single
// unsigned atomicMax.  The "empty" key is 0, below every mapped value.
#pragma once

#include <hip/hip_runtime.h>
#include <limits.h>
#include <stdint.h>
#include <stdio.h>

#define SHP_CHECK(call)                                                                   \
  do {                                                                                    \
    hipError_t _e = (call);                                                               \
    if (_e != hipSuccess) {                                                               \
      fprintf(stderr, "HIP error %s at %s:%d: %s\n", #call, __FILE__, __LINE__,           \
              hipGetErrorString(_e));                                                     \
      return shp_fail(_e);                                                                \
    }                                                                                     \
  } while (0)

namespace shp {

__host__ __device__ inline uint64_t ord64(double d) {
  uint64_t b = __builtin_bit_cast(uint64_t, d);
  return (b & 0x8000000000000000ULL) ? ~b : (b | 0x8000000000000000ULL);
}
__host__ __device__ inline double unord64(uint64_t k) {
  uint64_t b = (k & 0x8000000000000000ULL) ? (k & 0x7FFFFFFFFFFFFFFFULL) : ~k;
  return __builtin_bit_cast(double, b);
}
__host__ __device__ inline uint32_t ord32(float f) {
  uint32_t b = __builtin_bit_cast(uint32_t, f);
  return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}
__host__ __device__ inline float unord32(uint32_t k) {
  uint32_t b = (k & 0x80000000u) ? (k & 0x7FFFFFFFu) : ~k;
  return __builtin_bit_cast(float, b);
}

// Packed z-buffer key of the mixed path: depth in the high word, inverted
// facet index in the low word, so max(key) = nearest facet, ties -> lowest f.
__device__ inline uint64_t pack_key(float z, int f) {
  return ((uint64_t)ord32(z) << 32) | (uint64_t)(0xFFFFFFFFu - (uint32_t)f);
}
__device__ inline int key_facet(uint64_t k) { return (int)(0xFFFFFFFFu - (uint32_t)(k & 0xFFFFFFFFu)); }
__device__ inline float key_z(uint64_t k) { return unord32((uint32_t)(k >> 32)); }

// atomic min/max on doubles through their ordered integer image
__device__ inline void atomic_min_d(unsigned long long *a, double v) { atomicMin(a, (unsigned long long)ord64(v)); }
__device__ inline void atomic_max_d(unsigned long long *a, double v) { atomicMax(a, (unsigned long long)ord64(v)); }

}  // namespace shp

extern "C" int shp_fail(hipError_t e);
