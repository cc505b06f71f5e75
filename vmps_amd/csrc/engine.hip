// Device engine for the staged effective-Hamiltonian application on gfx950.
//
//   stage A   T_t   = L_l . V_v                     one problem per distinct (l, v)
//   stage C   out_o = sum_r (sum_j c_j T_j) . R_r    one problem per output block;
//             the scalar combine U = sum c T is done while loading the MFMA operand
//
// Work unit = one 64-lane wavefront computing one E x E output tile (E = 16
// for pivots whose blocks are all small, else 32) with (E/16)^2 16x16 MFMA
// accumulators.  Operands go from global memory straight to
// registers (every element of a wave's A and B tiles is used by exactly one
// MFMA, so staging in LDS would buy no reuse); the tile descriptor is a single
// load.  Every output element is owned by exactly one wave, so there are no
// atomics and results are bitwise reproducible run to run.
#include "engine.hpp"
#include "combine.hpp"
#include "wave.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>

namespace vmps {

typedef double v4d __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

namespace {

template <class T>
struct Mfma;

// The GEMM C = A.B is evaluated as C^T = B^T . A^T: the MFMA A operand is
// B[k][j] (lane l: j = l%16, k = l/16) and the MFMA B operand is A[i][k]
// (i = l%16, k = l/16).  Accumulator element v of lane l then holds
// C[i = l%16][j = row(l, v)], so a store instruction covers 16 consecutive
// rows of a column-major block.  row(l, v) as measured on gfx950:
// f64 16x16x4 -> l/16 + 4v, f32 16x16x4 -> 4(l/16) + v.
template <>
struct Mfma<double> {
  typedef v4d acc_t;
  __device__ static inline void mac(double b, double a, acc_t& c) {
    c = __builtin_amdgcn_mfma_f64_16x16x4f64(b, a, c, 0, 0, 0);
  }
  __device__ static inline int col(int lane, int v) { return (lane >> 4) + 4 * v; }
};
template <>
struct Mfma<float> {
  typedef v4f acc_t;
  __device__ static inline void mac(float b, float a, acc_t& c) {
    c = __builtin_amdgcn_mfma_f32_16x16x4f32(b, a, c, 0, 0, 0);
  }
  __device__ static inline int col(int lane, int v) { return 4 * (lane >> 4) + v; }
};

// A wave computes an E x E output tile (E = 16 or 32) as S x S MFMA
// accumulators of 16 x 16 (S = E / 16); sub-tiles past the block edge are
// skipped with wave-uniform tests.
template <class T, int S>
struct WaveTile {
  typedef typename Mfma<T>::acc_t acc_t;
  acc_t c[S][S];
  __device__ void zero() {
#pragma unroll
    for (int x = 0; x < S; ++x)
#pragma unroll
      for (int y = 0; y < S; ++y) c[x][y] = acc_t{0, 0, 0, 0};
  }
  // a[x][u]: A[i = m0 + 16x + lane%16][k = 4u + lane/16], b[y][u]: B[k][j = n0 + 16y + lane%16]
  __device__ void mac(const T (&a)[S][4], const T (&b)[S][4], bool has_m1, bool has_n1) {
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
      for (int y = 0; y < S; ++y)
#pragma unroll
        for (int x = 0; x < S; ++x)
          if ((x == 0 || has_m1) && (y == 0 || has_n1)) Mfma<T>::mac(b[y][u], a[x][u], c[x][y]);
  }
  template <class OutT>
  __device__ void store(OutT* __restrict__ C, int M, int N, int m0, int n0, int lane) const {
#pragma unroll
    for (int x = 0; x < S; ++x) {
      const int i = m0 + 16 * x + (lane & 15);
      if (i >= M) continue;
#pragma unroll
      for (int y = 0; y < S; ++y)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int j = n0 + 16 * y + Mfma<T>::col(lane, v);
          if (j < N) C[(int64_t)j * M + i] = (OutT)c[x][y][v];
        }
    }
  }
};

// Stage A: T (M x N, ld M) = L (M x K, ld M) . V (K x N, ld K), one E x E tile
// per wave.  The loads of the next 16-deep K block are issued before the
// MFMAs of the current one (two register buffers), so a wave waits for memory
// once per tile rather than once per K block.  The wave index goes through
// readfirstlane so the compiler knows it is wave-uniform: the descriptor is
// then one scalar load and the K loop branches on scalars instead of masks.
template <class T, int E>
__global__ __launch_bounds__(256) void k_stage_a(const TileA* __restrict__ tiles, int ntiles, const T* __restrict__ L,
                                                 const T* __restrict__ V, T* __restrict__ Tout) {
  constexpr int S = E / 16;
  const int w = blockIdx.x * 4 + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  if (w >= ntiles) return;
  const int lane = threadIdx.x & 63;
  const TileA t = tiles[w];
  const int m0 = t.mt * E, n0 = t.nt * E;
  const int li = lane & 15, lk = lane >> 4;
  const bool has_m1 = S > 1 && m0 + 16 < t.M, has_n1 = S > 1 && n0 + 16 < t.N;  // wave-uniform
  const T* __restrict__ A = L + t.a_off;
  const T* __restrict__ B = V + t.b_off;
  WaveTile<T, S> acc;
  acc.zero();
  auto load = [&](int kb, T (&a)[S][4], T (&b)[S][4]) {
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int k = kb + 4 * u + lk;
      const bool kin = k < t.K;
#pragma unroll
      for (int x = 0; x < S; ++x) {
        const int i = m0 + 16 * x + li, j = n0 + 16 * x + li;
        a[x][u] = (kin && i < t.M) ? A[(int64_t)k * t.M + i] : T(0);
        b[x][u] = (kin && j < t.N) ? B[(int64_t)j * t.K + k] : T(0);
      }
    }
  };
  T a0[S][4], b0[S][4], a1[S][4], b1[S][4];
  load(0, a0, b0);
  for (int kb = 0; kb < t.K; kb += 32) {
    if (kb + 16 < t.K) load(kb + 16, a1, b1);
    acc.mac(a0, b0, has_m1, has_n1);
    if (kb + 16 >= t.K) break;
    if (kb + 32 < t.K) load(kb + 32, a0, b0);
    acc.mac(a1, b1, has_m1, has_n1);
  }
  acc.store(Tout + t.c_off, t.M, t.N, m0, n0, lane);
}

// Stage C: Vout (M x N, ld M) = sum over chunks (sum_j c_j T_j)(M x K, ld M) . R (K x N, ld K).
// One workgroup per E x E output tile.  The iteration space (chunk, 16-deep K
// block) is dealt round-robin to the 4 waves (split K).  Chunk descriptors and
// each chunk's term list (at most kMaxTermsPerChunk) are fetched with one
// load per lane and broadcast with readlane, so within a K block all loads
// (R and every T_j) are independent and in flight together.  The 4 partial
// accumulators are summed through LDS in wave order, so the result does not
// depend on timing.
template <class T, class OutT, int E>
__global__ __launch_bounds__(256) void k_stage_c(const TileC* __restrict__ tiles, int ntiles,
                                                 const ChunkC* __restrict__ chunks, const TermC* __restrict__ terms,
                                                 const T* __restrict__ Tbuf, const T* __restrict__ R,
                                                 OutT* __restrict__ Vout) {
  constexpr int S = E / 16;
  typedef typename Mfma<T>::acc_t acc_t;
  __shared__ acc_t red[3][S * S][64];
  if ((int)blockIdx.x >= ntiles) return;  // workgroup-uniform
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  const TileC t = tiles[blockIdx.x];
  const int m0 = t.mt * E, n0 = t.nt * E;
  const int li = lane & 15, lk = lane >> 4;
  const bool has_m1 = S > 1 && m0 + 16 < t.M, has_n1 = S > 1 && n0 + 16 < t.N;
  WaveTile<T, S> acc;
  acc.zero();
  int it = 0;
  for (int g = t.chunk_begin; g < t.chunk_end; g += 64) {
    const int nch = min(64, t.chunk_end - g);
    int64_t my_r = 0;
    int my_K = 0, my_tb = 0, my_te = 0;
    if (lane < nch) {
      const ChunkC cd = chunks[g + lane];
      my_r = cd.r_off;
      my_K = cd.K;
      my_tb = cd.term_begin;
      my_te = cd.term_end;
    }
    for (int ci = 0; ci < nch; ++ci) {
      const int K = __builtin_amdgcn_readlane(my_K, ci);
      const int nkb = (K + 15) >> 4;
      // K blocks of this chunk handled by this wave: it + q with (it + q) % 4 == wave
      const int first = (wave - (it & 3)) & 3;
      if (first >= nkb) {
        it += nkb;
        continue;
      }
      const T* __restrict__ B = R + readlane64(my_r, ci);
      const int tb = __builtin_amdgcn_readlane(my_tb, ci);
      const int nt = __builtin_amdgcn_readlane(my_te, ci) - tb;
      int64_t my_toff = 0;
      T my_c = T(0);
      if (lane < nt) {
        const TermC tm = terms[tb + lane];
        my_toff = tm.t_off;
        my_c = (T)tm.c;
      }
      for (int q = first; q < nkb; q += 4) {
        const int kb = q * 16;
        T a[S][4], b[S][4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const int k = kb + 4 * u + lk;
          const bool kin = k < K;
#pragma unroll
          for (int y = 0; y < S; ++y) {
            const int j = n0 + 16 * y + li;
            b[y][u] = (kin && j < t.N) ? B[(int64_t)j * K + k] : T(0);
          }
#pragma unroll
          for (int x = 0; x < S; ++x) a[x][u] = T(0);
        }
        int jt = 0;
        for (; jt + 1 < nt; jt += 2) {  // two terms per round, all loads first
          const T* __restrict__ T0 = Tbuf + readlane64(my_toff, jt);
          const T* __restrict__ T1 = Tbuf + readlane64(my_toff, jt + 1);
          const T c0 = readlane_t<T>(my_c, jt), c1 = readlane_t<T>(my_c, jt + 1);
          T x0[S][4], x1[S][4];
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            const int k = kb + 4 * u + lk;
            const bool kin = k < K;
#pragma unroll
            for (int x = 0; x < S; ++x) {
              const int i = m0 + 16 * x + li;
              const bool ok = kin && i < t.M;
              x0[x][u] = ok ? T0[(int64_t)k * t.M + i] : T(0);
              x1[x][u] = ok ? T1[(int64_t)k * t.M + i] : T(0);
            }
          }
#pragma unroll
          for (int u = 0; u < 4; ++u)
#pragma unroll
            for (int x = 0; x < S; ++x) a[x][u] += c0 * x0[x][u] + c1 * x1[x][u];
        }
        if (jt < nt) {
          const T* __restrict__ T0 = Tbuf + readlane64(my_toff, jt);
          const T c0 = readlane_t<T>(my_c, jt);
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            const int k = kb + 4 * u + lk;
            const bool kin = k < K;
#pragma unroll
            for (int x = 0; x < S; ++x) {
              const int i = m0 + 16 * x + li;
              a[x][u] += c0 * ((kin && i < t.M) ? T0[(int64_t)k * t.M + i] : T(0));
            }
          }
        }
        acc.mac(a, b, has_m1, has_n1);
      }
      it += nkb;
    }
  }
  if (wave > 0) {
#pragma unroll
    for (int x = 0; x < S; ++x)
#pragma unroll
      for (int y = 0; y < S; ++y) red[wave - 1][x * S + y][lane] = acc.c[x][y];
  }
  __syncthreads();
  if (wave == 0) {
#pragma unroll
    for (int w = 0; w < 3; ++w)
#pragma unroll
      for (int x = 0; x < S; ++x)
#pragma unroll
        for (int y = 0; y < S; ++y) acc.c[x][y] += red[w][x * S + y][lane];
    acc.store(Vout + t.c_off, t.M, t.N, m0, n0, lane);
  }
}

// ---- LDS-staged 128 x 128 workgroup tiles for large blocks ------------------
// The big-tile kernels ask for two waves per SIMD (amdgpu_waves_per_eu): the
// compiler then keeps accumulators and operands within 256 registers (no
// spills) instead of using up to 384, and two workgroups share a CU.
// Problems with both dimensions >= H.big_min64 (a tile configuration choice,
// off by default).  4 waves in a 2 x 2 grid, each owning 64 x 64 = 4 x 4 MFMA
// accumulators.  Per 16-deep K step the workgroup
// stages A[128 x 16] and B[16 x 128] k-major in LDS (the layout the MFMA
// lanes read: 16 consecutive rows / columns per k) with coalesced global
// loads, so every operand element is fetched from HBM once per tile instead
// of once per 16- or 32-wide wave tile.  The next K step's global loads are
// issued before the MFMAs of the current one.
constexpr int kBT = kBigTile;
constexpr int kBKd = 16;
constexpr int kBQ = kBKd / 2;  // K elements each thread stages per operand and step

template <class T>
struct BigTileSmem {
  T A[kBKd][kBT];
  T B[kBKd][kBT];
};

template <class T>
__device__ inline void big_tile_mac(const BigTileSmem<T>& s, int wm, int wn, int li, int lk,
                                    typename Mfma<T>::acc_t (&acc)[4][4]) {
#pragma unroll
  for (int ks = 0; ks < kBKd; ks += 4) {
    T a[4], b[4];
#pragma unroll
    for (int x = 0; x < 4; ++x) {
      a[x] = s.A[ks + lk][wm + 16 * x + li];
      b[x] = s.B[ks + lk][wn + 16 * x + li];
    }
#pragma unroll
    for (int y = 0; y < 4; ++y)
#pragma unroll
      for (int x = 0; x < 4; ++x) Mfma<T>::mac(b[y], a[x], acc[x][y]);
  }
}

template <class T, class OutT>
__device__ inline void big_tile_store(OutT* __restrict__ C, int M, int N, int m0, int n0, int lane,
                                      const typename Mfma<T>::acc_t (&acc)[4][4]) {
#pragma unroll
  for (int x = 0; x < 4; ++x) {
    const int i = m0 + 16 * x + (lane & 15);
    if (i >= M) continue;
#pragma unroll
    for (int y = 0; y < 4; ++y)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int j = n0 + 16 * y + Mfma<T>::col(lane, v);
        if (j < N) C[(int64_t)j * M + i] = (OutT)acc[x][y][v];
      }
  }
}

// Thread roles in the staging loads (256 threads, 2048 elements per operand):
//   A (column-major, ld M): row i = tid % 128, k = kb + tid/128 + 2q, q < kBQ
//   B (column-major K x N, ld K): column j = tid / 2, k = kb + kBQ (tid % 2) + q
template <class T>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void k_stage_a_big(const TileA* __restrict__ tiles, int ntiles,
                                                     const T* __restrict__ L, const T* __restrict__ V,
                                                     T* __restrict__ Tout) {
  __shared__ BigTileSmem<T> s;
  if ((int)blockIdx.x >= ntiles) return;  // workgroup-uniform
  const int tid = threadIdx.x, wave = __builtin_amdgcn_readfirstlane(tid >> 6), lane = tid & 63;
  const int li = lane & 15, lk = lane >> 4;
  const int wm = (wave & 1) * 64, wn = (wave >> 1) * 64;
  const TileA t = tiles[blockIdx.x];
  const int m0 = t.mt * kBT, n0 = t.nt * kBT;
  const T* __restrict__ A = L + t.a_off;
  const T* __restrict__ B = V + t.b_off;
  const int ai = tid & 127, ak = tid >> 7, bj = tid >> 1, bk = (tid & 1) * kBQ;
  const bool arow = m0 + ai < t.M, bcol = n0 + bj < t.N;
  typename Mfma<T>::acc_t acc[4][4];
#pragma unroll
  for (int x = 0; x < 4; ++x)
#pragma unroll
    for (int y = 0; y < 4; ++y) acc[x][y] = typename Mfma<T>::acc_t{0, 0, 0, 0};
  T ra[kBQ], rb[kBQ];
  auto gload = [&](int kb) {
#pragma unroll
    for (int q = 0; q < kBQ; ++q) {
      const int ka = kb + ak + 2 * q, kbb = kb + bk + q;
      ra[q] = (arow && ka < t.K) ? A[(int64_t)ka * t.M + m0 + ai] : T(0);
      rb[q] = (bcol && kbb < t.K) ? B[(int64_t)(n0 + bj) * t.K + kbb] : T(0);
    }
  };
  gload(0);
  for (int kb = 0; kb < t.K; kb += kBKd) {
    __syncthreads();  // the previous step's LDS reads are done
#pragma unroll
    for (int q = 0; q < kBQ; ++q) {
      s.A[ak + 2 * q][ai] = ra[q];
      s.B[bk + q][bj] = rb[q];
    }
    __syncthreads();
    if (kb + kBKd < t.K) gload(kb + kBKd);
    big_tile_mac<T>(s, wm, wn, li, lk, acc);
  }
  big_tile_store<T, T>(Tout + t.c_off, t.M, t.N, m0 + wm, n0 + wn, lane, acc);
}

// Stage C, large output blocks: Vout = sum over chunks (sum_j c_j T_j) . R_r.
// The combine U = sum c_j T_j is formed while staging the A operand: the
// chunk's term list is loaded once per lane and broadcast with readlane.
// kOneTerm: every chunk has exactly one term (explicit stage B), which the
// kernel then streams with the next K step's loads in flight; a separate
// instantiation, so that the general path's registers do not add up with it.
template <class T, class OutT, bool kOneTerm>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void k_stage_c_big(const TileC* __restrict__ tiles, int ntiles,
                                                     const ChunkC* __restrict__ chunks,
                                                     const TermC* __restrict__ terms, const T* __restrict__ Tbuf,
                                                     const T* __restrict__ R, OutT* __restrict__ Vout) {
  __shared__ BigTileSmem<T> s;
  if ((int)blockIdx.x >= ntiles) return;  // workgroup-uniform
  const int tid = threadIdx.x, wave = __builtin_amdgcn_readfirstlane(tid >> 6), lane = tid & 63;
  const int li = lane & 15, lk = lane >> 4;
  const int wm = (wave & 1) * 64, wn = (wave >> 1) * 64;
  const TileC t = tiles[blockIdx.x];
  const int m0 = t.mt * kBT, n0 = t.nt * kBT;
  const int ai = tid & 127, ak = tid >> 7, bj = tid >> 1, bk = (tid & 1) * kBQ;
  const bool arow = m0 + ai < t.M, bcol = n0 + bj < t.N;
  typename Mfma<T>::acc_t acc[4][4];
#pragma unroll
  for (int x = 0; x < 4; ++x)
#pragma unroll
    for (int y = 0; y < 4; ++y) acc[x][y] = typename Mfma<T>::acc_t{0, 0, 0, 0};
  for (int c = t.chunk_begin; c < t.chunk_end; ++c) {
    const ChunkC ch = chunks[c];
    const int K = ch.K;
    const T* __restrict__ B = R + ch.r_off;
    const int nt = ch.term_end - ch.term_begin;
    int64_t my_toff = 0;
    T my_c = T(0);
    if (lane < nt) {
      const TermC tm = terms[ch.term_begin + lane];
      my_toff = tm.t_off;
      my_c = (T)tm.c;
    }
    if (kOneTerm) {  // one term (explicit stage B): a plain product, next step's loads in flight
      const T* __restrict__ T0 = Tbuf + readlane64(my_toff, 0);
      const T c0 = readlane_t<T>(my_c, 0);
      T ra[kBQ], rb[kBQ];
      auto gload = [&](int kb) {
#pragma unroll
        for (int q = 0; q < kBQ; ++q) {
          const int ka = kb + ak + 2 * q, kbb = kb + bk + q;
          ra[q] = (arow && ka < K) ? T0[(int64_t)ka * t.M + m0 + ai] : T(0);
          rb[q] = (bcol && kbb < K) ? B[(int64_t)(n0 + bj) * K + kbb] : T(0);
        }
      };
      gload(0);
      for (int kb = 0; kb < K; kb += kBKd) {
        __syncthreads();  // the previous step's LDS reads are done
#pragma unroll
        for (int q = 0; q < kBQ; ++q) {
          s.A[ak + 2 * q][ai] = c0 * ra[q];
          s.B[bk + q][bj] = rb[q];
        }
        __syncthreads();
        if (kb + kBKd < K) gload(kb + kBKd);
        big_tile_mac<T>(s, wm, wn, li, lk, acc);
      }
      continue;
    }
    for (int kb = 0; kb < K; kb += kBKd) {
      T ra[kBQ], rb[kBQ];
#pragma unroll
      for (int q = 0; q < kBQ; ++q) {
        const int kbb = kb + bk + q;
        rb[q] = (bcol && kbb < K) ? B[(int64_t)(n0 + bj) * K + kbb] : T(0);
        ra[q] = T(0);
      }
      int jt = 0;
      for (; jt + 1 < nt; jt += 2) {
        const T* __restrict__ T0 = Tbuf + readlane64(my_toff, jt);
        const T* __restrict__ T1 = Tbuf + readlane64(my_toff, jt + 1);
        const T c0 = readlane_t<T>(my_c, jt), c1 = readlane_t<T>(my_c, jt + 1);
        T x0[kBQ], x1[kBQ];
#pragma unroll
        for (int q = 0; q < kBQ; ++q) {
          const int ka = kb + ak + 2 * q;
          const bool ok = arow && ka < K;
          x0[q] = ok ? T0[(int64_t)ka * t.M + m0 + ai] : T(0);
          x1[q] = ok ? T1[(int64_t)ka * t.M + m0 + ai] : T(0);
        }
#pragma unroll
        for (int q = 0; q < kBQ; ++q) ra[q] += c0 * x0[q] + c1 * x1[q];
      }
      if (jt < nt) {
        const T* __restrict__ T0 = Tbuf + readlane64(my_toff, jt);
        const T c0 = readlane_t<T>(my_c, jt);
#pragma unroll
        for (int q = 0; q < kBQ; ++q) {
          const int ka = kb + ak + 2 * q;
          ra[q] += c0 * ((arow && ka < K) ? T0[(int64_t)ka * t.M + m0 + ai] : T(0));
        }
      }
      __syncthreads();  // the previous step's LDS reads are done
#pragma unroll
      for (int q = 0; q < kBQ; ++q) {
        s.A[ak + 2 * q][ai] = ra[q];
        s.B[bk + q][bj] = rb[q];
      }
      __syncthreads();
      big_tile_mac<T>(s, wm, wn, li, lk, acc);
    }
  }
  big_tile_store<T, OutT>(Vout + t.c_off, t.M, t.N, m0 + wm, n0 + wn, lane, acc);
}

// Weighted dot product, pass 1: workgroup b reduces segments [b*spb, (b+1)*spb).
__global__ __launch_bounds__(256) void k_dot_partial(const int64_t* __restrict__ seg_off, const int32_t* __restrict__ seg_len,
                                                     const double* __restrict__ seg_w, int nseg, int spb,
                                                     const double* __restrict__ x, const double* __restrict__ y,
                                                     double* __restrict__ partial) {
  __shared__ double red[256];
  double acc = 0.0;
  const int s0 = blockIdx.x * spb;
  const int s1 = min(nseg, s0 + spb);
  for (int s = s0; s < s1; ++s) {
    const int64_t off = seg_off[s];
    const int len = seg_len[s];
    double part = 0.0;
    for (int i = threadIdx.x; i < len; i += 256) part += x[off + i] * y[off + i];
    acc += seg_w[s] * part;
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

// Pass 2: one workgroup sums the partials in a fixed order.
__global__ __launch_bounds__(256) void k_dot_final(const double* __restrict__ partial, int n, double* __restrict__ out) {
  __shared__ double red[256];
  double acc = 0.0;
  for (int i = threadIdx.x; i < n; i += 256) acc += partial[i];
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = red[0];
}

__global__ void k_f64_to_f32(const double* __restrict__ x, float* __restrict__ y, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) y[i] = (float)x[i];
}

__global__ void k_zero(double* __restrict__ y, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) y[i] = 0.0;
}

// y += alpha[0] * scale * x  (projector update; alpha is a device scalar)
__global__ void k_axpy_dev(const double* __restrict__ alpha, double scale, const double* __restrict__ x,
                           double* __restrict__ y, int64_t n) {
  const double a = alpha[0] * scale;
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) y[i] += a * x[i];
}

constexpr int kDotSegPerBlock = 4;

inline unsigned grid_for(size_t waves) { return (unsigned)((waves + 3) / 4); }

}  // namespace

namespace {
std::mutex g_stream_mutex;
std::map<int, std::vector<hipStream_t>> g_free_streams;  // device -> idle streams
}  // namespace

hipStream_t acquire_stream() {
  int dev = 0;
  VMPS_HIP_CHECK(hipGetDevice(&dev));
  {
    std::lock_guard<std::mutex> lock(g_stream_mutex);
    auto& v = g_free_streams[dev];
    if (!v.empty()) {
      hipStream_t s = v.back();
      v.pop_back();
      return s;
    }
  }
  hipStream_t s = nullptr;
  VMPS_HIP_CHECK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  return s;
}

void release_stream(hipStream_t s, int device) {
  if (!s) return;
  std::lock_guard<std::mutex> lock(g_stream_mutex);
  g_free_streams[device].push_back(s);
}

void release_stream(hipStream_t s) {
  if (!s) return;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) {
    (void)hipStreamDestroy(s);
    return;
  }
  release_stream(s, dev);
}

DevPlan::~DevPlan() {
  if (stream) {
    (void)hipStreamSynchronize(stream);
    release_stream(stream, device);
  }
  if (ev0) (void)hipEventDestroy(ev0);
  if (ev1) (void)hipEventDestroy(ev1);
}

void DevPlan::create(const Pivot& P, vmps_precision precision) {
  build_host_plan(P, H);
  upload(precision);
  nproj = (int)P.A0proj.size();
  Epenalty = P.Epenalty;
  if (nproj) {
    std::vector<double> a0;
    for (auto& v : P.A0proj) a0.insert(a0.end(), v.begin(), v.end());
    A0.upload(a0, stream);
    VMPS_HIP_CHECK(hipStreamSynchronize(stream));
  }
}

void DevPlan::create_triples(const TripleSet& S, vmps_precision precision) {
  build_plan(S, H);
  upload(precision);
}

void DevPlan::create_triples(TripleSet&& S, vmps_precision precision) {
  build_plan(std::move(S), H);
  upload(precision);
}

// Environment blocks resident on the device, copied into the packed operand
// (one workgroup per block, strided over the grid).
__global__ void k_gather_blocks(const Gather* g, int n, double* dst) {
  for (int i = blockIdx.x; i < n; i += gridDim.x) {
    const double* src = g[i].src;
    double* d = dst + g[i].dst_off;
    const int64_t len = g[i].n;
    for (int64_t j = threadIdx.x; j < len; j += blockDim.x) d[j] = src[j];
  }
}

namespace {
void upload_env(DevBuf<double>& dst, const std::vector<double>& host, const std::vector<Gather>& g, int64_t n,
                hipStream_t s) {
  if (g.empty()) {
    dst.upload(host, s);
    return;
  }
  dst.alloc((size_t)n);
  DevBuf<Gather> list;
  list.upload(g, s);
  const int nb = (int)g.size();
  hipLaunchKernelGGL(k_gather_blocks, dim3(std::min(nb, 8192)), dim3(256), 0, s, list.p, nb, dst.p);
  VMPS_HIP_CHECK(hipGetLastError());
  VMPS_HIP_CHECK(hipStreamSynchronize(s));  // the list is freed on return
}
}  // namespace

void DevPlan::upload(vmps_precision precision) {
  VMPS_HIP_CHECK(hipGetDevice(&device));
  VMPS_HIP_CHECK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, device));
  stream = acquire_stream();
  VMPS_HIP_CHECK(hipEventCreate(&ev0));
  VMPS_HIP_CHECK(hipEventCreate(&ev1));
  // device-resident environments whose update may still be running
  for (auto& w : H.env_waits) w->wait();
  H.env_waits.clear();
  upload_env(L, H.Ldata, H.Lgather, H.L_size, stream);
  upload_env(R, H.Rdata, H.Rgather, H.R_size, stream);
  T.alloc(H.T_size);
  vin.alloc(H.ysize);
  vout.alloc(H.osize);
  if (H.ysize || H.osize) {
    VMPS_HIP_CHECK(hipMemsetAsync(vin.p, 0, std::max<int64_t>(1, H.ysize) * sizeof(double), stream));
    VMPS_HIP_CHECK(hipMemsetAsync(vout.p, 0, std::max<int64_t>(1, H.osize) * sizeof(double), stream));
  }
  upload_lists();
  seg_off.upload(H.seg_off, stream);
  seg_len.upload(H.seg_len, stream);
  seg_w.upload(H.seg_w, stream);
  const size_t nseg = H.seg_off.size();
  partial.alloc(std::max<size_t>(1, (nseg + kDotSegPerBlock - 1) / kDotSegPerBlock));
  scalars.alloc(64);
  prec = precision;
  ensure_precision_buffers();
  VMPS_HIP_CHECK(hipStreamSynchronize(stream));
}

void DevPlan::upload_lists() {
  tilesA.upload(H.tilesA_w, stream);  // FP64 / FP32 wave kernels (edge H.wave_edge)
  tilesC.upload(H.tilesC_w, stream);
  tilesAb.upload(H.tilesA_b, stream);  // FP64 / FP32 big-tile kernels
  tilesCb.upload(H.tilesC_b, stream);
  chunksC.upload(H.chunksC, stream);
  termsC.upload(H.termsC, stream);
  combTasks.upload(H.comb_tasks, stream);
  combTerms.upload(H.comb_terms, stream);
  if (H.explicit_u) {
    if (Ub.n != (size_t)H.U_size) Ub.alloc(H.U_size);
    if (T32.p && U32.n != (size_t)H.U_size) U32.alloc(H.U_size);
  } else {
    Ub.release();
    U32.release();
  }
  ++gen;
}

void DevPlan::ensure_precision_buffers() {
  if (prec == VMPS_PREC_FP32 && H.explicit_u && T32.p && U32.n != (size_t)H.U_size) {
    U32.alloc(H.U_size);
    ++gen;
  }
  if (prec == VMPS_PREC_FP32 && !T32.p && (H.T_size || H.vsize)) {
    L32.alloc(H.L_size);
    R32.alloc(H.R_size);
    V32.alloc(H.ysize);
    T32.alloc(H.T_size);
    if (H.explicit_u) U32.alloc(H.U_size);
    if (L.n) hipLaunchKernelGGL(k_f64_to_f32, dim3(1024), dim3(256), 0, stream, L.p, L32.p, (int64_t)L.n);
    if (R.n) hipLaunchKernelGGL(k_f64_to_f32, dim3(1024), dim3(256), 0, stream, R.p, R32.p, (int64_t)R.n);
    VMPS_HIP_CHECK(hipGetLastError());
    ++gen;
  }
  if (prec == VMPS_PREC_BF16X3) ensure_bf16_buffers();
}

void DevPlan::launch_stage(int stage, const double* d_vin, double* d_vout, hipStream_t s) {
  const int nA = (int)tilesA.n, nC = (int)tilesC.n, nAb = (int)tilesAb.n, nCb = (int)tilesCb.n;
  if (prec == VMPS_PREC_FP64) {
    if (stage == 0 && nAb)
      hipLaunchKernelGGL(k_stage_a_big<double>, dim3(nAb), dim3(256), 0, s, tilesAb.p, nAb, L.p, d_vin, T.p);
    if (stage == 0 && nA) {
      if (H.wave_edge == 16)
        hipLaunchKernelGGL((k_stage_a<double, 16>), dim3(grid_for(nA)), dim3(256), 0, s, tilesA.p, nA, L.p, d_vin,
                           T.p);
      else
        hipLaunchKernelGGL((k_stage_a<double, 32>), dim3(grid_for(nA)), dim3(256), 0, s, tilesA.p, nA, L.p, d_vin,
                           T.p);
    }
    if (stage == 2 && (nC || nCb)) {
      const double* src = T.p;
      if (H.explicit_u) {  // stage B: U = sum c T, then stage C reads U
        if (combTasks.n)
          hipLaunchKernelGGL((k_combine<double, TermC>), dim3((unsigned)combTasks.n), dim3(256), 0, s, combTasks.p,
                             combTerms.p, T.p, Ub.p);
        src = Ub.p;
      }
      if (nCb && H.explicit_u)
        hipLaunchKernelGGL((k_stage_c_big<double, double, true>), dim3(nCb), dim3(256), 0, s, tilesCb.p, nCb,
                           chunksC.p, termsC.p, src, R.p, d_vout);
      else if (nCb)
        hipLaunchKernelGGL((k_stage_c_big<double, double, false>), dim3(nCb), dim3(256), 0, s, tilesCb.p, nCb,
                           chunksC.p, termsC.p, src, R.p, d_vout);
      if (nC && H.wave_edge == 16)
        hipLaunchKernelGGL((k_stage_c<double, double, 16>), dim3(nC), dim3(256), 0, s, tilesC.p, nC, chunksC.p,
                           termsC.p, src, R.p, d_vout);
      else if (nC)
        hipLaunchKernelGGL((k_stage_c<double, double, 32>), dim3(nC), dim3(256), 0, s, tilesC.p, nC, chunksC.p,
                           termsC.p, src, R.p, d_vout);
    }
  } else if (prec == VMPS_PREC_FP32) {
    if (stage == 0) {
      if (H.ysize) hipLaunchKernelGGL(k_f64_to_f32, dim3(1024), dim3(256), 0, s, d_vin, V32.p, (int64_t)H.ysize);
      if (nAb)
        hipLaunchKernelGGL(k_stage_a_big<float>, dim3(nAb), dim3(256), 0, s, tilesAb.p, nAb, L32.p, V32.p, T32.p);
      if (nA) {
        if (H.wave_edge == 16)
          hipLaunchKernelGGL((k_stage_a<float, 16>), dim3(grid_for(nA)), dim3(256), 0, s, tilesA.p, nA, L32.p, V32.p,
                             T32.p);
        else
          hipLaunchKernelGGL((k_stage_a<float, 32>), dim3(grid_for(nA)), dim3(256), 0, s, tilesA.p, nA, L32.p, V32.p,
                             T32.p);
      }
    }
    if (stage == 2 && (nC || nCb)) {
      const float* src = T32.p;
      if (H.explicit_u) {
        if (combTasks.n)
          hipLaunchKernelGGL((k_combine<float, TermC>), dim3((unsigned)combTasks.n), dim3(256), 0, s, combTasks.p,
                             combTerms.p, T32.p, U32.p);
        src = U32.p;
      }
      if (nCb && H.explicit_u)
        hipLaunchKernelGGL((k_stage_c_big<float, double, true>), dim3(nCb), dim3(256), 0, s, tilesCb.p, nCb,
                           chunksC.p, termsC.p, src, R32.p, d_vout);
      else if (nCb)
        hipLaunchKernelGGL((k_stage_c_big<float, double, false>), dim3(nCb), dim3(256), 0, s, tilesCb.p, nCb,
                           chunksC.p, termsC.p, src, R32.p, d_vout);
      if (nC && H.wave_edge == 16)
        hipLaunchKernelGGL((k_stage_c<float, double, 16>), dim3(nC), dim3(256), 0, s, tilesC.p, nC, chunksC.p,
                           termsC.p, src, R32.p, d_vout);
      else if (nC)
        hipLaunchKernelGGL((k_stage_c<float, double, 32>), dim3(nC), dim3(256), 0, s, tilesC.p, nC, chunksC.p,
                           termsC.p, src, R32.p, d_vout);
    }
  } else if (prec == VMPS_PREC_BF16X3) {
    launch_stage_bf16x3(stage, d_vin, d_vout, s);
  } else {
    throw VmpsError(VMPS_ERR_ARG, "unsupported precision");
  }
  VMPS_HIP_CHECK(hipGetLastError());
}

void DevPlan::hxv(const double* d_vin, double* d_vout, hipStream_t s) {
  if (!s) s = stream;
  launch_stage(0, d_vin, d_vout, s);
  launch_stage(2, d_vin, d_vout, s);
  // excited-state penalty: Vout += Epenalty * <A0_n|Vin> * A0_n, the overlap
  // being A0proj[n][s].adjoint().contract(Vin[s]).trace() = coeff_dot-weighted
  // dot (DmrgPivotMatrix2.h:250-271, tensors/Biped.h:936-944, :976-1010)
  for (int n = 0; n < nproj; ++n) {
    const double* a0 = A0.p + (size_t)n * H.vsize;
    dot_async(a0, d_vin, scalars.p + 1, s);
    hipLaunchKernelGGL(k_axpy_dev, dim3(1024), dim3(256), 0, s, scalars.p + 1, Epenalty, a0, d_vout, (int64_t)H.vsize);
  }
  VMPS_HIP_CHECK(hipGetLastError());
}

void DevPlan::dot_async(const double* d_x, const double* d_y, double* d_result, hipStream_t s) {
  if (!s) s = stream;
  const int nseg = (int)H.seg_off.size();
  if (nseg == 0) {
    hipLaunchKernelGGL(k_zero, dim3(1), dim3(64), 0, s, d_result, (int64_t)1);
    VMPS_HIP_CHECK(hipGetLastError());
    return;
  }
  const int nb = (nseg + kDotSegPerBlock - 1) / kDotSegPerBlock;
  hipLaunchKernelGGL(k_dot_partial, dim3(nb), dim3(256), 0, s, seg_off.p, seg_len.p, seg_w.p, nseg, kDotSegPerBlock,
                     d_x, d_y, partial.p);
  hipLaunchKernelGGL(k_dot_final, dim3(1), dim3(256), 0, s, partial.p, nb, d_result);
  VMPS_HIP_CHECK(hipGetLastError());
}

double DevPlan::dot(const double* d_x, const double* d_y, hipStream_t s) {
  if (!s) s = stream;
  dot_async(d_x, d_y, scalars.p, s);
  double r = 0.0;
  VMPS_HIP_CHECK(hipMemcpyAsync(&r, scalars.p, sizeof(double), hipMemcpyDeviceToHost, s));
  VMPS_HIP_CHECK(hipStreamSynchronize(s));
  return r;
}

}  // namespace vmps
