// Explicit stage B kernel shared by the FP64 / FP32 and split-BF16 paths:
// U_u = sum_j c_j T_j elementwise over blocks of equal shape and layout.
#pragma once
#include "pivot.hpp"
#include "wave.hpp"

namespace vmps {
namespace {

// One workgroup per task (at most kCombTaskElems elements of one U block).
// Loads and stores are 16 bytes per lane (W = 16 / sizeof(T) elements):
// vector q = threadIdx.x + 256 v of the task covers elements [W q, W q + W),
// so a wave moves 1 KB per instruction; the planner keeps every T and U
// block offset a multiple of 4 elements, which makes these accesses aligned.
// A task's last vector may be partial and is then handled element-wise.  The
// task's terms come 64 at a time, one descriptor per lane, broadcast with
// readlane; two terms per round, their loads in flight together.
template <class T, class TermT>
__global__ __launch_bounds__(256) void k_combine(const CombTask* __restrict__ tasks, const TermT* __restrict__ terms,
                                                 const T* __restrict__ Tbuf, T* __restrict__ Ubuf) {
  constexpr int W = 16 / (int)sizeof(T);
  constexpr int NV = kCombTaskElems / 256 / W;
  typedef T vec_t __attribute__((ext_vector_type(W)));
  const CombTask t = tasks[blockIdx.x];
  const int lane = threadIdx.x & 63;
  auto load = [&](const T* __restrict__ src, int e) -> vec_t {
    if (e + W <= t.n) return *reinterpret_cast<const vec_t*>(src + e);
    vec_t x;
#pragma unroll
    for (int q = 0; q < W; ++q) x[q] = e + q < t.n ? src[e + q] : T(0);
    return x;
  };
  vec_t acc[NV];
#pragma unroll
  for (int v = 0; v < NV; ++v) acc[v] = vec_t(T(0));
  for (int g = t.term_begin; g < t.term_end; g += 64) {
    const int nt = min(64, t.term_end - g);
    int64_t my_off = 0;
    T my_c = T(0);
    if (lane < nt) {
      const TermT tm = terms[g + lane];
      my_off = tm.t_off;
      my_c = (T)tm.c;
    }
    int j = 0;
    for (; j + 1 < nt; j += 2) {
      const T* __restrict__ s0 = Tbuf + readlane64(my_off, j) + t.e0;
      const T* __restrict__ s1 = Tbuf + readlane64(my_off, j + 1) + t.e0;
      const T c0 = readlane_t<T>(my_c, j), c1 = readlane_t<T>(my_c, j + 1);
      vec_t x0[NV], x1[NV];
#pragma unroll
      for (int v = 0; v < NV; ++v) {
        const int e = W * ((int)threadIdx.x + 256 * v);
        if (e < t.n) {
          x0[v] = load(s0, e);
          x1[v] = load(s1, e);
        } else {
          x0[v] = x1[v] = vec_t(T(0));
        }
      }
#pragma unroll
      for (int v = 0; v < NV; ++v) acc[v] += c0 * x0[v] + c1 * x1[v];
    }
    if (j < nt) {
      const T* __restrict__ s0 = Tbuf + readlane64(my_off, j) + t.e0;
      const T c0 = readlane_t<T>(my_c, j);
#pragma unroll
      for (int v = 0; v < NV; ++v) {
        const int e = W * ((int)threadIdx.x + 256 * v);
        if (e < t.n) acc[v] += c0 * load(s0, e);
      }
    }
  }
  T* __restrict__ dst = Ubuf + t.u_off + t.e0;
#pragma unroll
  for (int v = 0; v < NV; ++v) {
    const int e = W * ((int)threadIdx.x + 256 * v);
    if (e + W <= t.n) {
      *reinterpret_cast<vec_t*>(dst + e) = acc[v];
    } else {
#pragma unroll
      for (int q = 0; q < W; ++q)
        if (e + q < t.n) dst[e + q] = acc[v][q];
    }
  }
}

}  // namespace
}  // namespace vmps
