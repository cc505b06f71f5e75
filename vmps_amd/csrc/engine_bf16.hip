// Split-BF16 x3 path of the staged effective-Hamiltonian application.
//
// Every FP64 operand x is carried as hi = bf16(x), lo = bf16(x - hi); a
// product a.b is formed as a_hi.b_hi + a_hi.b_lo + a_lo.b_hi by three
// v_mfma_f32_16x16x32_bf16 with FP32 accumulation (relative error ~1e-5).
// Operands are k-contiguous with k padded to a multiple of 8 (one aligned
// 16-byte load = one MFMA lane's 8 k values; the zero padding ends K):
//   L^T (row-major M x Kp), V and R (column-major Kp x N), T (row-major M x Np).
// Work decomposition as in the FP64 kernels: stage A one wave per 32 x 32
// tile, stage C one workgroup per tile with the combine sum_j c_j T_j done on
// the operand load and a 4-way split of the K iterations reduced through LDS.
#include "engine.hpp"

namespace vmps {

typedef float v4f __attribute__((ext_vector_type(4)));
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));

namespace {

// Operand roles as in the FP64 kernels: the MFMA computes C^T = B^T A^T, so the
// MFMA A operand is B[k..k+7][j] and the MFMA B operand is A[i][k..k+7];
// accumulator element v of lane l then holds C[i = l%16][j = 4(l/16) + v].
__device__ inline v4f mac3(const v8bf& bh, const v8bf& bl, const v8bf& ah, const v8bf& al, v4f c) {
  c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bl, ah, c, 0, 0, 0);
  c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bh, al, c, 0, 0, 0);
  c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bh, ah, c, 0, 0, 0);
  return c;
}

__device__ inline v8bf ld8(const __bf16* p) { return *reinterpret_cast<const v8bf*>(p); }

__device__ inline void split8(const float (&x)[8], v8bf& hi, v8bf& lo) {
#pragma unroll
  for (int e = 0; e < 8; ++e) {
    const __bf16 h = (__bf16)x[e];
    hi[e] = h;
    lo[e] = (__bf16)(x[e] - (float)h);
  }
}

// Packed FP64 V columns -> padded split-BF16 copies (one workgroup per task of
// at most 4096 padded elements).
__global__ __launch_bounds__(256) void k_split_v(const VConv* __restrict__ tasks, const double* __restrict__ vin,
                                                 __bf16* __restrict__ Vhi, __bf16* __restrict__ Vlo) {
  const VConv t = tasks[blockIdx.x];
  const int n = t.Kp * t.N;
  for (int e = threadIdx.x; e < n; e += 256) {
    const int j = e / t.Kp, k = e - j * t.Kp;
    const double x = (k < t.K) ? vin[t.src_off + (int64_t)j * t.K + k] : 0.0;
    const __bf16 h = (__bf16)(float)x;
    Vhi[t.dst_off + e] = h;
    Vlo[t.dst_off + e] = (__bf16)(float)(x - (double)(float)h);
  }
}

// Stage A: Tp (row-major M x Np) = L (Lt row-major M x Kp) . V (col-major Kp x N)
__global__ __launch_bounds__(256) void k_stage_a_bf16x3(const TileA16* __restrict__ tiles, int ntiles,
                                                        const __bf16* __restrict__ Lhi, const __bf16* __restrict__ Llo,
                                                        const __bf16* __restrict__ Vhi, const __bf16* __restrict__ Vlo,
                                                        float* __restrict__ Tp) {
  const int w = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (w >= ntiles) return;
  const int lane = threadIdx.x & 63;
  const TileA16 t = tiles[w];
  const int m0 = t.mt * kTile, n0 = t.nt * kTile;
  const int li = lane & 15, lg = lane >> 4;
  const int i0 = m0 + li, i1 = i0 + 16, j0 = n0 + li, j1 = j0 + 16;
  const bool has_m1 = m0 + 16 < t.M, has_n1 = n0 + 16 < t.N;
  const v8bf z = {};
  v4f c00 = {0, 0, 0, 0}, c01 = c00, c10 = c00, c11 = c00;
  for (int kb = 0; kb < t.Kp; kb += 32) {
    const int k = kb + 8 * lg;
    const bool kin = k < t.Kp;
    const v8bf a0h = (kin && i0 < t.M) ? ld8(Lhi + t.a_off + (int64_t)i0 * t.Kp + k) : z;
    const v8bf a0l = (kin && i0 < t.M) ? ld8(Llo + t.a_off + (int64_t)i0 * t.Kp + k) : z;
    const v8bf a1h = (kin && i1 < t.M) ? ld8(Lhi + t.a_off + (int64_t)i1 * t.Kp + k) : z;
    const v8bf a1l = (kin && i1 < t.M) ? ld8(Llo + t.a_off + (int64_t)i1 * t.Kp + k) : z;
    const v8bf b0h = (kin && j0 < t.N) ? ld8(Vhi + t.b_off + (int64_t)j0 * t.Kp + k) : z;
    const v8bf b0l = (kin && j0 < t.N) ? ld8(Vlo + t.b_off + (int64_t)j0 * t.Kp + k) : z;
    const v8bf b1h = (kin && j1 < t.N) ? ld8(Vhi + t.b_off + (int64_t)j1 * t.Kp + k) : z;
    const v8bf b1l = (kin && j1 < t.N) ? ld8(Vlo + t.b_off + (int64_t)j1 * t.Kp + k) : z;
    c00 = mac3(b0h, b0l, a0h, a0l, c00);
    if (has_m1) c10 = mac3(b0h, b0l, a1h, a1l, c10);
    if (has_n1) {
      c01 = mac3(b1h, b1l, a0h, a0l, c01);
      if (has_m1) c11 = mac3(b1h, b1l, a1h, a1l, c11);
    }
  }
  float* __restrict__ C = Tp + t.c_off;
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    const int ja = n0 + 4 * lg + v, jb = ja + 16;
    if (i0 < t.M) {
      if (ja < t.N) C[(int64_t)i0 * t.Np + ja] = c00[v];
      if (jb < t.N) C[(int64_t)i0 * t.Np + jb] = c01[v];
    }
    if (i1 < t.M) {
      if (ja < t.N) C[(int64_t)i1 * t.Np + ja] = c10[v];
      if (jb < t.N) C[(int64_t)i1 * t.Np + jb] = c11[v];
    }
  }
}

// Stage C: Vout (FP64 col-major M x N) = sum over chunks (sum_j c_j T_j) . R
__global__ __launch_bounds__(256) void k_stage_c_bf16x3(const TileC* __restrict__ tiles, int ntiles,
                                                        const ChunkC16* __restrict__ chunks,
                                                        const TermC16* __restrict__ terms, const float* __restrict__ Tp,
                                                        const __bf16* __restrict__ Rhi, const __bf16* __restrict__ Rlo,
                                                        double* __restrict__ Vout) {
  __shared__ v4f red[3][4][64];
  if ((int)blockIdx.x >= ntiles) return;  // workgroup-uniform
  const int wave = threadIdx.x >> 6;
  const int lane = threadIdx.x & 63;
  const TileC t = tiles[blockIdx.x];
  const int m0 = t.mt * kTile, n0 = t.nt * kTile;
  const int li = lane & 15, lg = lane >> 4;
  const int i0 = m0 + li, i1 = i0 + 16, j0 = n0 + li, j1 = j0 + 16;
  const bool has_m1 = m0 + 16 < t.M, has_n1 = n0 + 16 < t.N;
  const bool r0 = i0 < t.M, r1 = i1 < t.M, q0 = j0 < t.N, q1 = j1 < t.N;
  const v8bf z = {};
  v4f c00 = {0, 0, 0, 0}, c01 = c00, c10 = c00, c11 = c00;
  int it = 0;
  for (int c = t.chunk_begin; c < t.chunk_end; ++c) {
    const ChunkC16 ch = chunks[c];
    for (int kb = 0; kb < ch.Kp; kb += 32, ++it) {
      if ((it & 3) != wave) continue;
      const int k = kb + 8 * lg;
      const bool kin = k < ch.Kp;
      float u0[8], u1[8];
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        u0[e] = 0.f;
        u1[e] = 0.f;
      }
      for (int j = ch.term_begin; j < ch.term_end; ++j) {
        const TermC16 tm = terms[j];
        const float* __restrict__ Tt = Tp + tm.t_off + k;
        if (kin && r0) {
          const v4f x = *reinterpret_cast<const v4f*>(Tt + (int64_t)i0 * ch.Kp);
          const v4f y = *reinterpret_cast<const v4f*>(Tt + (int64_t)i0 * ch.Kp + 4);
#pragma unroll
          for (int e = 0; e < 4; ++e) {
            u0[e] += tm.c * x[e];
            u0[e + 4] += tm.c * y[e];
          }
        }
        if (kin && r1) {
          const v4f x = *reinterpret_cast<const v4f*>(Tt + (int64_t)i1 * ch.Kp);
          const v4f y = *reinterpret_cast<const v4f*>(Tt + (int64_t)i1 * ch.Kp + 4);
#pragma unroll
          for (int e = 0; e < 4; ++e) {
            u1[e] += tm.c * x[e];
            u1[e + 4] += tm.c * y[e];
          }
        }
      }
      v8bf a0h, a0l, a1h, a1l;
      split8(u0, a0h, a0l);
      split8(u1, a1h, a1l);
      const v8bf b0h = (kin && q0) ? ld8(Rhi + ch.r_off + (int64_t)j0 * ch.Kp + k) : z;
      const v8bf b0l = (kin && q0) ? ld8(Rlo + ch.r_off + (int64_t)j0 * ch.Kp + k) : z;
      const v8bf b1h = (kin && q1) ? ld8(Rhi + ch.r_off + (int64_t)j1 * ch.Kp + k) : z;
      const v8bf b1l = (kin && q1) ? ld8(Rlo + ch.r_off + (int64_t)j1 * ch.Kp + k) : z;
      c00 = mac3(b0h, b0l, a0h, a0l, c00);
      if (has_m1) c10 = mac3(b0h, b0l, a1h, a1l, c10);
      if (has_n1) {
        c01 = mac3(b1h, b1l, a0h, a0l, c01);
        if (has_m1) c11 = mac3(b1h, b1l, a1h, a1l, c11);
      }
    }
  }
  if (wave > 0) {
    red[wave - 1][0][lane] = c00;
    red[wave - 1][1][lane] = c01;
    red[wave - 1][2][lane] = c10;
    red[wave - 1][3][lane] = c11;
  }
  __syncthreads();
  if (wave == 0) {
#pragma unroll
    for (int w = 0; w < 3; ++w) {
      c00 += red[w][0][lane];
      c01 += red[w][1][lane];
      c10 += red[w][2][lane];
      c11 += red[w][3][lane];
    }
    double* __restrict__ C = Vout + t.c_off;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int ja = n0 + 4 * lg + v, jb = ja + 16;
      if (ja < t.N) {
        if (r0) C[(int64_t)ja * t.M + i0] = (double)c00[v];
        if (r1) C[(int64_t)ja * t.M + i1] = (double)c10[v];
      }
      if (jb < t.N) {
        if (r0) C[(int64_t)jb * t.M + i0] = (double)c01[v];
        if (r1) C[(int64_t)jb * t.M + i1] = (double)c11[v];
      }
    }
  }
}

// ---- LDS-staged 128 x 128 workgroup tiles for large blocks -------------------
// 4 waves in a 2 x 2 grid, each owning 64 x 64 = 4 x 4 MFMA blocks.  Per 32-deep
// K step the workgroup stages the split operand tiles A[128 x 32] and
// B[32 x 128] (hi and lo) in LDS, k-contiguous with rows padded to 40
// elements, and each wave then issues 4 x 4 x 3 MFMAs from LDS reads.
constexpr int kBT = kBigTile;
constexpr int kBK = 32;
constexpr int kLS = 40;  // LDS row stride in bf16 elements (80 bytes)

struct BigSmem {
  __bf16 Ah[kBT * kLS], Al[kBT * kLS], Bh[kBT * kLS], Bl[kBT * kLS];
};

__device__ inline void big_compute(const BigSmem& s, int wm, int wn, int li, int lg, v4f (&acc)[4][4]) {
  v8bf ah[4], al[4], bh[4], bl[4];
#pragma unroll
  for (int x = 0; x < 4; ++x) {
    ah[x] = *reinterpret_cast<const v8bf*>(&s.Ah[(wm + 16 * x + li) * kLS + 8 * lg]);
    al[x] = *reinterpret_cast<const v8bf*>(&s.Al[(wm + 16 * x + li) * kLS + 8 * lg]);
    bh[x] = *reinterpret_cast<const v8bf*>(&s.Bh[(wn + 16 * x + li) * kLS + 8 * lg]);
    bl[x] = *reinterpret_cast<const v8bf*>(&s.Bl[(wn + 16 * x + li) * kLS + 8 * lg]);
  }
#pragma unroll
  for (int mi = 0; mi < 4; ++mi)
#pragma unroll
    for (int ni = 0; ni < 4; ++ni) acc[mi][ni] = mac3(bh[ni], bl[ni], ah[mi], al[mi], acc[mi][ni]);
}

// Stage A, large problems: Tp (row-major M x Np) = Lt (row-major M x Kp) . V (col-major Kp x N)
__global__ __launch_bounds__(256) void k_stage_a_bf16x3_big(const TileA16* __restrict__ tiles, int ntiles,
                                                            const __bf16* __restrict__ Lhi,
                                                            const __bf16* __restrict__ Llo,
                                                            const __bf16* __restrict__ Vhi,
                                                            const __bf16* __restrict__ Vlo, float* __restrict__ Tp) {
  __shared__ BigSmem s;
  if ((int)blockIdx.x >= ntiles) return;  // workgroup-uniform
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int li = lane & 15, lg = lane >> 4;
  const int wm = (wave & 1) * 64, wn = (wave >> 1) * 64;
  const TileA16 t = tiles[blockIdx.x];
  const int m0 = t.mt * kBT, n0 = t.nt * kBT;
  const v8bf z = {};
  v4f acc[4][4];
#pragma unroll
  for (int mi = 0; mi < 4; ++mi)
#pragma unroll
    for (int ni = 0; ni < 4; ++ni) acc[mi][ni] = v4f{0, 0, 0, 0};
  for (int kb = 0; kb < t.Kp; kb += kBK) {
    v8bf ah[2], al[2], bh[2], bl[2];
#pragma unroll
    for (int h = 0; h < 2; ++h) {  // 512 16-byte chunks per operand tile, 2 per thread
      const int c = tid + 256 * h, row = c >> 2, kc = (c & 3) * 8;
      const int k = kb + kc;
      const bool ka = (m0 + row < t.M) && (k < t.Kp), kbv = (n0 + row < t.N) && (k < t.Kp);
      ah[h] = ka ? ld8(Lhi + t.a_off + (int64_t)(m0 + row) * t.Kp + k) : z;
      al[h] = ka ? ld8(Llo + t.a_off + (int64_t)(m0 + row) * t.Kp + k) : z;
      bh[h] = kbv ? ld8(Vhi + t.b_off + (int64_t)(n0 + row) * t.Kp + k) : z;
      bl[h] = kbv ? ld8(Vlo + t.b_off + (int64_t)(n0 + row) * t.Kp + k) : z;
    }
    __syncthreads();  // the previous step's LDS reads are done
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int c = tid + 256 * h, row = c >> 2, kc = (c & 3) * 8;
      *reinterpret_cast<v8bf*>(&s.Ah[row * kLS + kc]) = ah[h];
      *reinterpret_cast<v8bf*>(&s.Al[row * kLS + kc]) = al[h];
      *reinterpret_cast<v8bf*>(&s.Bh[row * kLS + kc]) = bh[h];
      *reinterpret_cast<v8bf*>(&s.Bl[row * kLS + kc]) = bl[h];
    }
    __syncthreads();
    big_compute(s, wm, wn, li, lg, acc);
  }
  float* __restrict__ C = Tp + t.c_off;
#pragma unroll
  for (int mi = 0; mi < 4; ++mi) {
    const int i = m0 + wm + 16 * mi + li;
    if (i >= t.M) continue;
#pragma unroll
    for (int ni = 0; ni < 4; ++ni)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int j = n0 + wn + 16 * ni + 4 * lg + v;
        if (j < t.N) C[(int64_t)i * t.Np + j] = acc[mi][ni][v];
      }
  }
}

// Stage C, large output blocks: Vout (FP64 col-major M x N) = sum_chunks (sum_j c_j T_j) . R.
// Work item w covers iterations [it_begin, it_end) of its tile's flattened
// (chunk, 32-deep k block) loop; with slot >= 0 it writes an FP32 partial tile
// (column-major 128 x 128) that k_reduce_partials sums in slot order.
__global__ __launch_bounds__(256) void k_stage_c_bf16x3_big(const TileC* __restrict__ tiles,
                                                            const int32_t* __restrict__ w_tile,
                                                            const int32_t* __restrict__ w_begin,
                                                            const int32_t* __restrict__ w_end,
                                                            const int64_t* __restrict__ w_slot, int nwork,
                                                            const ChunkC16* __restrict__ chunks,
                                                            const TermC16* __restrict__ terms,
                                                            const float* __restrict__ Tp,
                                                            const __bf16* __restrict__ Rhi,
                                                            const __bf16* __restrict__ Rlo,
                                                            double* __restrict__ Vout, float* __restrict__ partial) {
  __shared__ BigSmem s;
  if ((int)blockIdx.x >= nwork) return;  // workgroup-uniform
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int li = lane & 15, lg = lane >> 4;
  const int wm = (wave & 1) * 64, wn = (wave >> 1) * 64;
  const TileC t = tiles[w_tile[blockIdx.x]];
  const int it0 = w_begin[blockIdx.x], it1 = w_end[blockIdx.x];
  const int64_t slot = w_slot[blockIdx.x];
  const int m0 = t.mt * kBT, n0 = t.nt * kBT;
  const v8bf z = {};
  v4f acc[4][4];
#pragma unroll
  for (int mi = 0; mi < 4; ++mi)
#pragma unroll
    for (int ni = 0; ni < 4; ++ni) acc[mi][ni] = v4f{0, 0, 0, 0};
  int it = 0;
  for (int c = t.chunk_begin; c < t.chunk_end && it < it1; ++c) {
    const ChunkC16 ch = chunks[c];
    const int nsteps = (ch.Kp + kBK - 1) / kBK;
    if (it + nsteps <= it0) {
      it += nsteps;
      continue;
    }
    for (int kb = 0; kb < ch.Kp; kb += kBK, ++it) {
      if (it < it0 || it >= it1) continue;  // workgroup-uniform
      v8bf ah[2], al[2], bh[2], bl[2];
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const int cc = tid + 256 * h, row = cc >> 2, kc = (cc & 3) * 8;
        const int k = kb + kc;
        const bool ka = (m0 + row < t.M) && (k < ch.Kp), kbv = (n0 + row < t.N) && (k < ch.Kp);
        float u[8];
#pragma unroll
        for (int e = 0; e < 8; ++e) u[e] = 0.f;
        if (ka) {
          for (int j = ch.term_begin; j < ch.term_end; ++j) {
            const TermC16 tm = terms[j];
            const float* __restrict__ Tt = Tp + tm.t_off + (int64_t)(m0 + row) * ch.Kp + k;
            const v4f x = *reinterpret_cast<const v4f*>(Tt);
            const v4f y = *reinterpret_cast<const v4f*>(Tt + 4);
#pragma unroll
            for (int e = 0; e < 4; ++e) {
              u[e] += tm.c * x[e];
              u[e + 4] += tm.c * y[e];
            }
          }
        }
        split8(u, ah[h], al[h]);
        bh[h] = kbv ? ld8(Rhi + ch.r_off + (int64_t)(n0 + row) * ch.Kp + k) : z;
        bl[h] = kbv ? ld8(Rlo + ch.r_off + (int64_t)(n0 + row) * ch.Kp + k) : z;
      }
      __syncthreads();
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const int cc = tid + 256 * h, row = cc >> 2, kc = (cc & 3) * 8;
        *reinterpret_cast<v8bf*>(&s.Ah[row * kLS + kc]) = ah[h];
        *reinterpret_cast<v8bf*>(&s.Al[row * kLS + kc]) = al[h];
        *reinterpret_cast<v8bf*>(&s.Bh[row * kLS + kc]) = bh[h];
        *reinterpret_cast<v8bf*>(&s.Bl[row * kLS + kc]) = bl[h];
      }
      __syncthreads();
      big_compute(s, wm, wn, li, lg, acc);
    }
  }
  double* __restrict__ C = Vout + t.c_off;
  float* __restrict__ P = partial + (slot < 0 ? 0 : slot * (int64_t)(kBT * kBT));
#pragma unroll
  for (int mi = 0; mi < 4; ++mi) {
    const int r = wm + 16 * mi + li, i = m0 + r;
    if (i >= t.M) continue;
#pragma unroll
    for (int ni = 0; ni < 4; ++ni)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int cl = wn + 16 * ni + 4 * lg + v, j = n0 + cl;
        if (j >= t.N) continue;
        if (slot < 0)
          C[(int64_t)j * t.M + i] = (double)acc[mi][ni][v];
        else
          P[cl * kBT + r] = acc[mi][ni][v];
      }
  }
}

// Vout tile = sum over its split partials, in slot order.
__global__ __launch_bounds__(256) void k_reduce_partials(const TileC* __restrict__ tiles,
                                                         const int32_t* __restrict__ r_tile,
                                                         const int64_t* __restrict__ r_slot,
                                                         const int32_t* __restrict__ r_nsplit,
                                                         const float* __restrict__ partial, double* __restrict__ Vout) {
  const TileC t = tiles[r_tile[blockIdx.x]];
  const int64_t s0 = r_slot[blockIdx.x];
  const int ns = r_nsplit[blockIdx.x];
  const int m0 = t.mt * kBT, n0 = t.nt * kBT;
  for (int e = threadIdx.x; e < kBT * kBT; e += 256) {
    const int cl = e / kBT, r = e - cl * kBT;
    const int i = m0 + r, j = n0 + cl;
    if (i >= t.M || j >= t.N) continue;
    float acc = 0.f;
    for (int s = 0; s < ns; ++s) acc += partial[(s0 + s) * (int64_t)(kBT * kBT) + e];
    Vout[t.c_off + (int64_t)j * t.M + i] = (double)acc;
  }
}

}  // namespace

void DevPlan::ensure_bf16_buffers() {
  if (bfA.p || H.tilesA.empty() && H.tilesC.empty()) return;
  Bf16Layout B;
  build_bf16_layout(H, B);
  Lhi.upload(B.Lhi, stream);
  Llo.upload(B.Llo, stream);
  Rhi.upload(B.Rhi, stream);
  Rlo.upload(B.Rlo, stream);
  Vhi.alloc(std::max<int64_t>(1, B.V_size));
  Vlo.alloc(std::max<int64_t>(1, B.V_size));
  Tp.alloc(std::max<int64_t>(1, B.T_size));
  VMPS_HIP_CHECK(hipMemsetAsync(Tp.p, 0, Tp.n * sizeof(float), stream));  // padding columns stay zero
  bfA.upload(B.tilesA, stream);
  bfAbig.upload(B.tilesA_big, stream);
  bfCsmall.upload(B.tilesC_small, stream);
  bfCbig.upload(B.tilesC_big, stream);
  wTile.upload(B.splitC_tile, stream);
  wBegin.upload(B.splitC_begin, stream);
  wEnd.upload(B.splitC_end, stream);
  wSlot.upload(B.splitC_slot, stream);
  rTile.upload(B.redC_tile, stream);
  rSlot.upload(B.redC_slot, stream);
  rNs.upload(B.redC_nsplit, stream);
  Cpart.alloc(std::max<int64_t>(1, B.partial_slots) * (int64_t)kBigTile * kBigTile);
  bfC.upload(B.chunksC, stream);
  bfT.upload(B.termsC, stream);
  vconv.upload(B.vconv, stream);
  VMPS_HIP_CHECK(hipStreamSynchronize(stream));
}

void DevPlan::launch_stage_bf16x3(int stage, const double* d_vin, double* d_vout, hipStream_t s) {
  const int nA = (int)bfA.n, nAb = (int)bfAbig.n, nC = (int)bfCsmall.n, nCb = (int)bfCbig.n;
  if (stage == 0) {
    if (vconv.n)
      hipLaunchKernelGGL(k_split_v, dim3((unsigned)vconv.n), dim3(256), 0, s, vconv.p, d_vin, (__bf16*)Vhi.p,
                         (__bf16*)Vlo.p);
    if (nAb)
      hipLaunchKernelGGL(k_stage_a_bf16x3_big, dim3((unsigned)nAb), dim3(256), 0, s, bfAbig.p, nAb,
                         (const __bf16*)Lhi.p, (const __bf16*)Llo.p, (const __bf16*)Vhi.p, (const __bf16*)Vlo.p, Tp.p);
    if (nA)
      hipLaunchKernelGGL(k_stage_a_bf16x3, dim3((unsigned)((nA + 3) / 4)), dim3(256), 0, s, bfA.p, nA,
                         (const __bf16*)Lhi.p, (const __bf16*)Llo.p, (const __bf16*)Vhi.p, (const __bf16*)Vlo.p, Tp.p);
  } else if (stage == 2) {
    const int nW = (int)wTile.n, nR = (int)rTile.n;
    if (nCb && nW)
      hipLaunchKernelGGL(k_stage_c_bf16x3_big, dim3((unsigned)nW), dim3(256), 0, s, bfCbig.p, wTile.p, wBegin.p,
                         wEnd.p, wSlot.p, nW, bfC.p, bfT.p, Tp.p, (const __bf16*)Rhi.p, (const __bf16*)Rlo.p, d_vout,
                         Cpart.p);
    if (nR)
      hipLaunchKernelGGL(k_reduce_partials, dim3((unsigned)nR), dim3(256), 0, s, bfCbig.p, rTile.p, rSlot.p, rNs.p,
                         Cpart.p, d_vout);
    if (nC)
      hipLaunchKernelGGL(k_stage_c_bf16x3, dim3((unsigned)nC), dim3(256), 0, s, bfCsmall.p, nC, bfC.p, bfT.p, Tp.p,
                         (const __bf16*)Rhi.p, (const __bf16*)Rlo.p, d_vout);
  }
  VMPS_HIP_CHECK(hipGetLastError());
}

}  // namespace vmps
