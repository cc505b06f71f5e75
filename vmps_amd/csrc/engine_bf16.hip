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

// Packed FP64 V blocks -> padded split-BF16 copies (one workgroup per block).
__global__ __launch_bounds__(256) void k_split_v(const VConv* __restrict__ tasks, const double* __restrict__ vin,
                                                 __bf16* __restrict__ Vhi, __bf16* __restrict__ Vlo) {
  const VConv t = tasks[blockIdx.x];
  const int64_t n = (int64_t)t.Kp * t.N;
  for (int64_t e = threadIdx.x; e < n; e += 256) {
    const int j = (int)(e / t.Kp), k = (int)(e % t.Kp);
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
  bfC.upload(B.chunksC, stream);
  bfT.upload(B.termsC, stream);
  vconv.upload(B.vconv, stream);
  VMPS_HIP_CHECK(hipStreamSynchronize(stream));
}

void DevPlan::launch_stage_bf16x3(int stage, const double* d_vin, double* d_vout, hipStream_t s) {
  const int nA = (int)bfA.n, nC = (int)tilesC.n;
  if (stage == 0) {
    if (vconv.n)
      hipLaunchKernelGGL(k_split_v, dim3((unsigned)vconv.n), dim3(256), 0, s, vconv.p, d_vin, (__bf16*)Vhi.p,
                         (__bf16*)Vlo.p);
    if (nA)
      hipLaunchKernelGGL(k_stage_a_bf16x3, dim3((unsigned)((nA + 3) / 4)), dim3(256), 0, s, bfA.p, nA,
                         (const __bf16*)Lhi.p, (const __bf16*)Llo.p, (const __bf16*)Vhi.p, (const __bf16*)Vlo.p, Tp.p);
  } else if (stage == 2 && nC) {
    hipLaunchKernelGGL(k_stage_c_bf16x3, dim3((unsigned)nC), dim3(256), 0, s, tilesC.p, nC, bfC.p, bfT.p, Tp.p,
                       (const __bf16*)Rhi.p, (const __bf16*)Rlo.p, d_vout);
  }
  VMPS_HIP_CHECK(hipGetLastError());
}

}  // namespace vmps
