// Wavefront broadcast helpers shared by the HIP kernels (device code only).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace vmps {

// 64-bit wave-uniform broadcast of lane `src`'s value.
__device__ inline int64_t readlane64(int64_t x, int src) {
  const int lo = __builtin_amdgcn_readlane((int)(uint32_t)(uint64_t)x, src);
  const int hi = __builtin_amdgcn_readlane((int)(uint32_t)((uint64_t)x >> 32), src);
  return (int64_t)(((uint64_t)(uint32_t)hi << 32) | (uint64_t)(uint32_t)lo);
}
__device__ inline int readlane32(int x, int src) { return __builtin_amdgcn_readlane(x, src); }
__device__ inline double readlane_d(double x, int src) { return __longlong_as_double(readlane64(__double_as_longlong(x), src)); }
__device__ inline float readlane_f(float x, int src) { return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), src)); }
template <class T>
__device__ inline T readlane_t(T x, int src);
template <>
__device__ inline double readlane_t<double>(double x, int src) { return readlane_d(x, src); }
template <>
__device__ inline float readlane_t<float>(float x, int src) { return readlane_f(x, src); }

}  // namespace vmps
