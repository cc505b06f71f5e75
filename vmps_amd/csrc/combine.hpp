// Explicit stage B kernel shared by the FP64 / FP32 and split-BF16 paths:
// U_u = sum_j c_j T_j elementwise over blocks of equal shape and layout.
#pragma once
#include "pivot.hpp"
#include "wave.hpp"

namespace vmps {
namespace {

// One workgroup per task (at most kCombTaskElems elements of one U block,
// kCombTaskElems / 256 per thread, strided so that the loads coalesce).  The
// task's terms come 64 at a time, one descriptor per lane, broadcast with
// readlane; a term's loads are independent of the previous term's.
template <class T, class TermT>
__global__ __launch_bounds__(256) void k_combine(const CombTask* __restrict__ tasks, const TermT* __restrict__ terms,
                                                 const T* __restrict__ Tbuf, T* __restrict__ Ubuf) {
  constexpr int P = kCombTaskElems / 256;
  const CombTask t = tasks[blockIdx.x];
  const int lane = threadIdx.x & 63;
  T acc[P];
#pragma unroll
  for (int p = 0; p < P; ++p) acc[p] = T(0);
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
    for (; j + 1 < nt; j += 2) {  // two terms per round: their loads are in flight together
      const T* __restrict__ s0 = Tbuf + readlane64(my_off, j) + t.e0;
      const T* __restrict__ s1 = Tbuf + readlane64(my_off, j + 1) + t.e0;
      const T c0 = readlane_t<T>(my_c, j), c1 = readlane_t<T>(my_c, j + 1);
      T x0[P], x1[P];
#pragma unroll
      for (int p = 0; p < P; ++p) {
        const int e = (int)threadIdx.x + 256 * p;
        x0[p] = e < t.n ? s0[e] : T(0);
        x1[p] = e < t.n ? s1[e] : T(0);
      }
#pragma unroll
      for (int p = 0; p < P; ++p) acc[p] += c0 * x0[p] + c1 * x1[p];
    }
    if (j < nt) {
      const T* __restrict__ s0 = Tbuf + readlane64(my_off, j) + t.e0;
      const T c0 = readlane_t<T>(my_c, j);
#pragma unroll
      for (int p = 0; p < P; ++p) {
        const int e = (int)threadIdx.x + 256 * p;
        if (e < t.n) acc[p] += c0 * s0[e];
      }
    }
  }
  T* __restrict__ dst = Ubuf + t.u_off + t.e0;
#pragma unroll
  for (int p = 0; p < P; ++p) {
    const int e = (int)threadIdx.x + 256 * p;
    if (e < t.n) dst[e] = acc[p];
  }
}

}  // namespace
}  // namespace vmps
