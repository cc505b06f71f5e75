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
#include "combine.hpp"
#include "wave.hpp"

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

// FP64 L blocks (column-major M x K, ld M) -> row-major M x Kp split-BF16
// copies (L^T in the layout of k_split_v's output); one workgroup per task of
// at most 4096 padded elements, task rows r read src[src_off + k M + r].
__global__ __launch_bounds__(256) void k_split_lt(const VConv* __restrict__ tasks, const double* __restrict__ L,
                                                  __bf16* __restrict__ Lhi, __bf16* __restrict__ Llo) {
  const VConv t = tasks[blockIdx.x];
  const int n = t.Kp * t.N;
  for (int e = threadIdx.x; e < n; e += 256) {
    const int r = e / t.Kp, k = e - r * t.Kp;
    const double x = (k < t.K) ? L[t.src_off + (int64_t)k * t.pad + r] : 0.0;
    const __bf16 h = (__bf16)(float)x;
    Lhi[t.dst_off + e] = h;
    Llo[t.dst_off + e] = (__bf16)(float)(x - (double)(float)h);
  }
}

// The split operands of one 32-deep K block of a 32 x 32 wave tile:
// A rows i0 = m0 + lane%16 and i0 + 16, B columns j0 = n0 + lane%16 and j0 + 16,
// k = kb + 8 (lane/16) .. +7.
struct WaveOps {
  v8bf a0h, a0l, a1h, a1l, b0h, b0l, b1h, b1l;
};

__device__ inline void wave_mac(const WaveOps& o, bool has_m1, bool has_n1, v4f& c00, v4f& c01, v4f& c10,
                                v4f& c11) {
  c00 = mac3(o.b0h, o.b0l, o.a0h, o.a0l, c00);
  if (has_m1) c10 = mac3(o.b0h, o.b0l, o.a1h, o.a1l, c10);
  if (has_n1) {
    c01 = mac3(o.b1h, o.b1l, o.a0h, o.a0l, c01);
    if (has_m1) c11 = mac3(o.b1h, o.b1l, o.a1h, o.a1l, c11);
  }
}

// Stage A, small problems: Tp (row-major M x Np) = L (Lt row-major M x Kp) . V (col-major Kp x N).
// Most of these tiles need one or two 32-deep K blocks, so a wave that took a
// single tile would spend its life waiting: for the descriptor, then for the
// operands.  Instead each wave walks the tiles w, w + W, w + 2W, ... (W = waves
// in the grid): their descriptors arrive 64 at a time, one per lane, and are
// broadcast with readlane, and the operand loads of the next K block -- of the
// same tile or of the wave's next tile -- are issued before the MFMAs of the
// current one.
struct TileCursor {
  int64_t a_off, b_off, c_off;
  int M, N, Kp, Np, m0, n0;
};

__device__ inline TileCursor tile_at(const TileA16& d, int src) {
  TileCursor c;
  c.a_off = readlane64(d.a_off, src);
  c.b_off = readlane64(d.b_off, src);
  c.c_off = readlane64(d.c_off, src);
  c.M = readlane32(d.M, src);
  c.N = readlane32(d.N, src);
  c.Kp = readlane32(d.Kp, src);
  c.Np = readlane32(d.Np, src);
  c.m0 = readlane32((int)d.mt, src) * kTile;
  c.n0 = readlane32((int)d.nt, src) * kTile;
  return c;
}

__global__ __launch_bounds__(256) void k_stage_a_bf16x3(const TileA16* __restrict__ tiles, int ntiles,
                                                        const __bf16* __restrict__ Lhi, const __bf16* __restrict__ Llo,
                                                        const __bf16* __restrict__ Vhi, const __bf16* __restrict__ Vlo,
                                                        float* __restrict__ Tp) {
  const int nwaves = gridDim.x * 4;
  const int w0 = blockIdx.x * 4 + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  if (w0 >= ntiles) return;
  const int nmine = (ntiles - 1 - w0) / nwaves + 1;  // tiles w0 + j * nwaves, j < nmine
  const int lane = threadIdx.x & 63;
  const int li = lane & 15, lg = lane >> 4;
  const v8bf z = {};
  auto load = [&](const TileCursor& t, int kb, WaveOps& o) {
    const int k = kb + 8 * lg;
    const bool kin = k < t.Kp;
    const int i0 = t.m0 + li, i1 = i0 + 16, j0 = t.n0 + li, j1 = j0 + 16;
    const bool ra = kin && i0 < t.M, rb = kin && i1 < t.M, qa = kin && j0 < t.N, qb = kin && j1 < t.N;
    o.a0h = ra ? ld8(Lhi + t.a_off + (int64_t)i0 * t.Kp + k) : z;
    o.a0l = ra ? ld8(Llo + t.a_off + (int64_t)i0 * t.Kp + k) : z;
    o.a1h = rb ? ld8(Lhi + t.a_off + (int64_t)i1 * t.Kp + k) : z;
    o.a1l = rb ? ld8(Llo + t.a_off + (int64_t)i1 * t.Kp + k) : z;
    o.b0h = qa ? ld8(Vhi + t.b_off + (int64_t)j0 * t.Kp + k) : z;
    o.b0l = qa ? ld8(Vlo + t.b_off + (int64_t)j0 * t.Kp + k) : z;
    o.b1h = qb ? ld8(Vhi + t.b_off + (int64_t)j1 * t.Kp + k) : z;
    o.b1l = qb ? ld8(Vlo + t.b_off + (int64_t)j1 * t.Kp + k) : z;
  };
  auto store = [&](const TileCursor& t, const v4f& c00, const v4f& c01, const v4f& c10, const v4f& c11) {
    const int i0 = t.m0 + li, i1 = i0 + 16;
    float* __restrict__ C = Tp + t.c_off;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int ja = t.n0 + 4 * lg + v, jb = ja + 16;
      if (i0 < t.M) {
        if (ja < t.N) C[(int64_t)i0 * t.Np + ja] = c00[v];
        if (jb < t.N) C[(int64_t)i0 * t.Np + jb] = c01[v];
      }
      if (i1 < t.M) {
        if (ja < t.N) C[(int64_t)i1 * t.Np + ja] = c10[v];
        if (jb < t.N) C[(int64_t)i1 * t.Np + jb] = c11[v];
      }
    }
  };
  for (int g = 0; g < nmine; g += 64) {
    const int n = min(64, nmine - g);
    TileA16 d = {};
    if (lane < n) d = tiles[w0 + (int64_t)(g + lane) * nwaves];
    // load cursor (lt, lj, lk) runs one K block ahead of the compute cursor (ct, cj, ck);
    // a tile has max(1, Kp / 32) K blocks (Kp is a multiple of 8)
    TileCursor lt = tile_at(d, 0), ct = lt;
    int lj = 0, lk = 0, cj = 0, ck = 0;
    auto advance_load = [&]() {
      lk += 32;
      if (lk >= max(lt.Kp, 1)) {
        lk = 0;
        if (++lj < n) lt = tile_at(d, lj);
      }
    };
    WaveOps o0, o1;
    load(lt, 0, o0);
    advance_load();
    v4f c00 = {0, 0, 0, 0}, c01 = c00, c10 = c00, c11 = c00;
    // one pipeline step: prefetch into `nx`, multiply `cur`; false when the group is done
    auto step = [&](const WaveOps& cur, WaveOps& nx) -> bool {
      if (lj < n) {
        load(lt, lk, nx);
        advance_load();
      }
      wave_mac(cur, ct.m0 + 16 < ct.M, ct.n0 + 16 < ct.N, c00, c01, c10, c11);
      ck += 32;
      if (ck >= max(ct.Kp, 1)) {
        store(ct, c00, c01, c10, c11);
        c00 = v4f{0, 0, 0, 0};
        c01 = c00;
        c10 = c00;
        c11 = c00;
        ck = 0;
        if (++cj == n) return false;
        ct = tile_at(d, cj);
      }
      return true;
    };
    while (step(o0, o1) && step(o1, o0)) {
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
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
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
  for (int g = t.chunk_begin; g < t.chunk_end; g += 64) {
    // chunk descriptors, one per lane, broadcast with readlane
    const int nch = min(64, t.chunk_end - g);
    int64_t my_r = 0;
    int my_K = 0, my_tb = 0, my_te = 0;
    if (lane < nch) {
      const ChunkC16 cd = chunks[g + lane];
      my_r = cd.r_off;
      my_K = cd.Kp;
      my_tb = cd.term_begin;
      my_te = cd.term_end;
    }
    for (int ci = 0; ci < nch; ++ci) {
      const int Kp = readlane32(my_K, ci);
      const int nkb = (Kp + 31) >> 5;
      // K blocks of this chunk handled by this wave: it + q with (it + q) % 4 == wave
      const int first = (wave - (it & 3)) & 3;
      if (first >= nkb) {
        it += nkb;
        continue;
      }
      const int64_t roff = readlane64(my_r, ci);
      const int tb = readlane32(my_tb, ci);
      const int nt = readlane32(my_te, ci) - tb;
      // the chunk's terms (at most 64), one per lane
      int64_t my_toff = 0;
      float my_c = 0.f;
      if (lane < nt) {
        const TermC16 tm = terms[tb + lane];
        my_toff = tm.t_off;
        my_c = tm.c;
      }
      for (int q = first; q < nkb; q += 4) {
        const int k = q * 32 + 8 * lg;
        const bool kin = k < Kp;
        const bool a0 = kin && r0, a1 = kin && r1;
        const v8bf b0h = (kin && q0) ? ld8(Rhi + roff + (int64_t)j0 * Kp + k) : z;
        const v8bf b0l = (kin && q0) ? ld8(Rlo + roff + (int64_t)j0 * Kp + k) : z;
        const v8bf b1h = (kin && q1) ? ld8(Rhi + roff + (int64_t)j1 * Kp + k) : z;
        const v8bf b1l = (kin && q1) ? ld8(Rlo + roff + (int64_t)j1 * Kp + k) : z;
        float u0[8], u1[8];
#pragma unroll
        for (int e = 0; e < 8; ++e) {
          u0[e] = 0.f;
          u1[e] = 0.f;
        }
        // U rows i0, i1 at k..k+7 = sum_j c_j T_j, two terms per round with all loads first
        int jt = 0;
        for (; jt + 1 < nt; jt += 2) {
          const float* __restrict__ T0 = Tp + readlane64(my_toff, jt) + k;
          const float* __restrict__ T1 = Tp + readlane64(my_toff, jt + 1) + k;
          const float c0 = readlane_f(my_c, jt), c1 = readlane_f(my_c, jt + 1);
          const v4f zf = {0, 0, 0, 0};
          const v4f x00 = a0 ? *reinterpret_cast<const v4f*>(T0 + (int64_t)i0 * Kp) : zf;
          const v4f x01 = a0 ? *reinterpret_cast<const v4f*>(T0 + (int64_t)i0 * Kp + 4) : zf;
          const v4f x10 = a1 ? *reinterpret_cast<const v4f*>(T0 + (int64_t)i1 * Kp) : zf;
          const v4f x11 = a1 ? *reinterpret_cast<const v4f*>(T0 + (int64_t)i1 * Kp + 4) : zf;
          const v4f y00 = a0 ? *reinterpret_cast<const v4f*>(T1 + (int64_t)i0 * Kp) : zf;
          const v4f y01 = a0 ? *reinterpret_cast<const v4f*>(T1 + (int64_t)i0 * Kp + 4) : zf;
          const v4f y10 = a1 ? *reinterpret_cast<const v4f*>(T1 + (int64_t)i1 * Kp) : zf;
          const v4f y11 = a1 ? *reinterpret_cast<const v4f*>(T1 + (int64_t)i1 * Kp + 4) : zf;
#pragma unroll
          for (int e = 0; e < 4; ++e) {
            u0[e] += c0 * x00[e] + c1 * y00[e];
            u0[e + 4] += c0 * x01[e] + c1 * y01[e];
            u1[e] += c0 * x10[e] + c1 * y10[e];
            u1[e + 4] += c0 * x11[e] + c1 * y11[e];
          }
        }
        if (jt < nt) {
          const float* __restrict__ T0 = Tp + readlane64(my_toff, jt) + k;
          const float c0 = readlane_f(my_c, jt);
          const v4f zf = {0, 0, 0, 0};
          const v4f x00 = a0 ? *reinterpret_cast<const v4f*>(T0 + (int64_t)i0 * Kp) : zf;
          const v4f x01 = a0 ? *reinterpret_cast<const v4f*>(T0 + (int64_t)i0 * Kp + 4) : zf;
          const v4f x10 = a1 ? *reinterpret_cast<const v4f*>(T0 + (int64_t)i1 * Kp) : zf;
          const v4f x11 = a1 ? *reinterpret_cast<const v4f*>(T0 + (int64_t)i1 * Kp + 4) : zf;
#pragma unroll
          for (int e = 0; e < 4; ++e) {
            u0[e] += c0 * x00[e];
            u0[e + 4] += c0 * x01[e];
            u1[e] += c0 * x10[e];
            u1[e + 4] += c0 * x11[e];
          }
        }
        WaveOps o;
        split8(u0, o.a0h, o.a0l);
        split8(u1, o.a1h, o.a1l);
        o.b0h = b0h;
        o.b0l = b0l;
        o.b1h = b1h;
        o.b1l = b1l;
        wave_mac(o, has_m1, has_n1, c00, c01, c10, c11);
      }
      it += nkb;
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
// 4 waves in a 2 x 2 grid, each owning 64 x 64 = 4 x 4 MFMA blocks.  Per
// 64-deep K step the workgroup stages the split operand tiles A[128 x 64] and
// B[64 x 128] (hi and lo) in LDS, k-contiguous with rows padded to 80
// elements (16-byte aligned reads), and each wave then issues
// 2 x 4 x 4 x 3 MFMAs from LDS reads.  The global loads of the next step are
// in flight during those MFMAs.
constexpr int kBT = kBigTile;
constexpr int kBK = kBigK;
// LDS row stride in bf16 elements: 160-byte rows (two workgroups of 80 KB
// fill the CU's 160 KB); 3% faster than 144-byte rows on the C60 pivots
constexpr int kLS = kBK + 16;
constexpr int kPieces = kBT * kBK / 8 / 256;  // 16-byte pieces per operand tile and thread

struct BigSmem {
  __bf16 Ah[kBT * kLS], Al[kBT * kLS], Bh[kBT * kLS], Bl[kBT * kLS];
};

// The wave raises its issue priority for its MFMA block (s_setprio), so that
// a wave with operands in registers is served before the other workgroup's
// waves staging their loads.
__device__ inline void big_compute(const BigSmem& s, int wm, int wn, int li, int lg, v4f (&acc)[4][4]) {
  __builtin_amdgcn_s_setprio(1);
#pragma unroll
  for (int ks = 0; ks < kBK; ks += 32) {
    v8bf ah[4], al[4], bh[4], bl[4];
#pragma unroll
    for (int x = 0; x < 4; ++x) {
      ah[x] = *reinterpret_cast<const v8bf*>(&s.Ah[(wm + 16 * x + li) * kLS + ks + 8 * lg]);
      al[x] = *reinterpret_cast<const v8bf*>(&s.Al[(wm + 16 * x + li) * kLS + ks + 8 * lg]);
      bh[x] = *reinterpret_cast<const v8bf*>(&s.Bh[(wn + 16 * x + li) * kLS + ks + 8 * lg]);
      bl[x] = *reinterpret_cast<const v8bf*>(&s.Bl[(wn + 16 * x + li) * kLS + ks + 8 * lg]);
    }
#pragma unroll
    for (int mi = 0; mi < 4; ++mi)
#pragma unroll
      for (int ni = 0; ni < 4; ++ni) acc[mi][ni] = mac3(bh[ni], bl[ni], ah[mi], al[mi], acc[mi][ni]);
  }
  __builtin_amdgcn_s_setprio(0);
}

// thread tid's piece h of a staged operand tile: row (of A) / column (of B), k offset
__device__ inline int piece_row(int tid, int h) { return (tid + 256 * h) / (kBK / 8); }
__device__ inline int piece_k(int tid, int h) { return ((tid + 256 * h) % (kBK / 8)) * 8; }

__device__ inline void big_stage(BigSmem& s, int tid, const v8bf (&ah)[kPieces], const v8bf (&al)[kPieces],
                                 const v8bf (&bh)[kPieces], const v8bf (&bl)[kPieces]) {
#pragma unroll
  for (int h = 0; h < kPieces; ++h) {
    const int o = piece_row(tid, h) * kLS + piece_k(tid, h);
    *reinterpret_cast<v8bf*>(&s.Ah[o]) = ah[h];
    *reinterpret_cast<v8bf*>(&s.Al[o]) = al[h];
    *reinterpret_cast<v8bf*>(&s.Bh[o]) = bh[h];
    *reinterpret_cast<v8bf*>(&s.Bl[o]) = bl[h];
  }
}

// Stage A, large problems: Tp (row-major M x Np) = Lt (row-major M x Kp) . V (col-major Kp x N)
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void k_stage_a_bf16x3_big(const TileA16* __restrict__ tiles, int ntiles,
                                                            const __bf16* __restrict__ Lhi,
                                                            const __bf16* __restrict__ Llo,
                                                            const __bf16* __restrict__ Vhi,
                                                            const __bf16* __restrict__ Vlo, float* __restrict__ Tp) {
  __shared__ BigSmem s;
  if ((int)blockIdx.x >= ntiles) return;  // workgroup-uniform
  const int tid = threadIdx.x, wave = __builtin_amdgcn_readfirstlane(tid >> 6), lane = tid & 63;
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
  v8bf ah[kPieces], al[kPieces], bh[kPieces], bl[kPieces];
  auto gload = [&](int kb) {
#pragma unroll
    for (int h = 0; h < kPieces; ++h) {
      const int row = piece_row(tid, h), k = kb + piece_k(tid, h);
      const bool ka = (m0 + row < t.M) && (k < t.Kp), kbv = (n0 + row < t.N) && (k < t.Kp);
      ah[h] = ka ? ld8(Lhi + t.a_off + (int64_t)(m0 + row) * t.Kp + k) : z;
      al[h] = ka ? ld8(Llo + t.a_off + (int64_t)(m0 + row) * t.Kp + k) : z;
      bh[h] = kbv ? ld8(Vhi + t.b_off + (int64_t)(n0 + row) * t.Kp + k) : z;
      bl[h] = kbv ? ld8(Vlo + t.b_off + (int64_t)(n0 + row) * t.Kp + k) : z;
    }
  };
  gload(0);
  for (int kb = 0; kb < t.Kp; kb += kBK) {
    __syncthreads();  // the previous step's LDS reads are done
    big_stage(s, tid, ah, al, bh, bl);
    __syncthreads();
    if (kb + kBK < t.Kp) gload(kb + kBK);  // in flight during the MFMAs below
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
// (chunk, kBK-deep k block) loop; with slot >= 0 it writes an FP32 partial tile
// (column-major 128 x 128) that k_reduce_partials sums in slot order.
// kOneTerm: every chunk has one term (explicit stage B): streamed with the
// next K step's loads in flight (a separate instantiation, see k_stage_c_big)
template <bool kOneTerm>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void k_stage_c_bf16x3_big(const TileC* __restrict__ tiles,
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
  const int tid = threadIdx.x, wave = __builtin_amdgcn_readfirstlane(tid >> 6), lane = tid & 63;
  const int li = lane & 15, lg = lane >> 4;
  const int wm = (wave & 1) * 64, wn = (wave >> 1) * 64;
  const TileC t = tiles[w_tile[blockIdx.x]];
  const int it0 = w_begin[blockIdx.x], it1 = w_end[blockIdx.x];
  const int64_t slot = w_slot[blockIdx.x];
  const int m0 = t.mt * kBT, n0 = t.nt * kBT;
  const v8bf z = {};
  const v4f zf = {0, 0, 0, 0};
  v4f acc[4][4];
#pragma unroll
  for (int mi = 0; mi < 4; ++mi)
#pragma unroll
    for (int ni = 0; ni < 4; ++ni) acc[mi][ni] = zf;
  int it = 0;
  for (int c = t.chunk_begin; c < t.chunk_end && it < it1; ++c) {
    const ChunkC16 ch = chunks[c];
    const int nsteps = (ch.Kp + kBK - 1) / kBK;
    if (it + nsteps <= it0) {
      it += nsteps;
      continue;
    }
    // the chunk's terms (at most 64), one per lane, broadcast with readlane
    const int nt = ch.term_end - ch.term_begin;
    int64_t my_toff = 0;
    float my_c = 0.f;
    if (lane < nt) {
      const TermC16 tm = terms[ch.term_begin + lane];
      my_toff = tm.t_off;
      my_c = tm.c;
    }
    if (kOneTerm) {  // one term (explicit stage B): a plain product, next step's loads in flight
      const float* __restrict__ T0 = Tp + readlane64(my_toff, 0);
      const float c0 = readlane_f(my_c, 0);
      const int s_lo = max(0, it0 - it), s_hi = min(nsteps, it1 - it);  // this work item's K steps
      v4f xr[kPieces][2];
      v8bf bh[kPieces], bl[kPieces];
      auto gload = [&](int kb) {
#pragma unroll
        for (int h = 0; h < kPieces; ++h) {
          const int row = piece_row(tid, h), k = kb + piece_k(tid, h);
          const bool ka = (m0 + row < t.M) && (k < ch.Kp), kbv = (n0 + row < t.N) && (k < ch.Kp);
          const float* __restrict__ src = T0 + (int64_t)(m0 + row) * ch.Kp + k;
          xr[h][0] = ka ? *reinterpret_cast<const v4f*>(src) : zf;
          xr[h][1] = ka ? *reinterpret_cast<const v4f*>(src + 4) : zf;
          bh[h] = kbv ? ld8(Rhi + ch.r_off + (int64_t)(n0 + row) * ch.Kp + k) : z;
          bl[h] = kbv ? ld8(Rlo + ch.r_off + (int64_t)(n0 + row) * ch.Kp + k) : z;
        }
      };
      if (s_lo < s_hi) gload(s_lo * kBK);
      for (int st = s_lo; st < s_hi; ++st) {
        v8bf ah[kPieces], al[kPieces];
#pragma unroll
        for (int h = 0; h < kPieces; ++h) {
          float u[8];
#pragma unroll
          for (int e = 0; e < 4; ++e) {
            u[e] = c0 * xr[h][0][e];
            u[e + 4] = c0 * xr[h][1][e];
          }
          split8(u, ah[h], al[h]);
        }
        __syncthreads();  // the previous step's LDS reads are done
        big_stage(s, tid, ah, al, bh, bl);
        __syncthreads();
        if (st + 1 < s_hi) gload((st + 1) * kBK);
        big_compute(s, wm, wn, li, lg, acc);
      }
      it += nsteps;
      continue;
    }
    for (int kb = 0; kb < ch.Kp; kb += kBK, ++it) {
      if (it < it0 || it >= it1) continue;  // workgroup-uniform
      v8bf ah[kPieces], al[kPieces], bh[kPieces], bl[kPieces];
      bool ka[kPieces];
      int64_t arow[kPieces];
      float u[kPieces][8];
#pragma unroll
      for (int h = 0; h < kPieces; ++h) {
        const int row = piece_row(tid, h), k = kb + piece_k(tid, h);
        ka[h] = (m0 + row < t.M) && (k < ch.Kp);
        arow[h] = (int64_t)(m0 + row) * ch.Kp + k;
        const bool kbv = (n0 + row < t.N) && (k < ch.Kp);
        bh[h] = kbv ? ld8(Rhi + ch.r_off + (int64_t)(n0 + row) * ch.Kp + k) : z;
        bl[h] = kbv ? ld8(Rlo + ch.r_off + (int64_t)(n0 + row) * ch.Kp + k) : z;
#pragma unroll
        for (int e = 0; e < 8; ++e) u[h][e] = 0.f;
      }
      // U = sum_j c_j T_j on this thread's 8-wide row pieces
      for (int jt = 0; jt < nt; ++jt) {
        const float* __restrict__ T0 = Tp + readlane64(my_toff, jt);
        const float c0 = readlane_f(my_c, jt);
        v4f x[kPieces][2];
#pragma unroll
        for (int h = 0; h < kPieces; ++h) {
          x[h][0] = ka[h] ? *reinterpret_cast<const v4f*>(T0 + arow[h]) : zf;
          x[h][1] = ka[h] ? *reinterpret_cast<const v4f*>(T0 + arow[h] + 4) : zf;
        }
#pragma unroll
        for (int h = 0; h < kPieces; ++h)
#pragma unroll
          for (int e = 0; e < 4; ++e) {
            u[h][e] += c0 * x[h][0][e];
            u[h][e + 4] += c0 * x[h][1][e];
          }
      }
#pragma unroll
      for (int h = 0; h < kPieces; ++h) split8(u[h], ah[h], al[h]);
      __syncthreads();
      big_stage(s, tid, ah, al, bh, bl);
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
  if (bf16_ready || (H.tilesA.empty() && H.tilesC.empty())) return;
  Bf16Layout B;
  build_bf16_layout(H, B);  // work lists; the operands are split on the device
  Lhi.alloc(std::max<int64_t>(1, B.L_size));
  Llo.alloc(std::max<int64_t>(1, B.L_size));
  Rhi.alloc(std::max<int64_t>(1, B.R_size));
  Rlo.alloc(std::max<int64_t>(1, B.R_size));
  {
    DevBuf<VConv> lc, rc;  // freed after the synchronisation below
    lc.upload(B.lconv, stream);
    rc.upload(B.rconv, stream);
    if (lc.n)
      hipLaunchKernelGGL(k_split_lt, dim3((unsigned)lc.n), dim3(256), 0, stream, lc.p, L.p, (__bf16*)Lhi.p,
                         (__bf16*)Llo.p);
    if (rc.n)
      hipLaunchKernelGGL(k_split_v, dim3((unsigned)rc.n), dim3(256), 0, stream, rc.p, R.p, (__bf16*)Rhi.p,
                         (__bf16*)Rlo.p);
    VMPS_HIP_CHECK(hipGetLastError());
    VMPS_HIP_CHECK(hipStreamSynchronize(stream));
  }
  Vhi.alloc(std::max<int64_t>(1, B.V_size));
  Vlo.alloc(std::max<int64_t>(1, B.V_size));
  Tp.alloc(std::max<int64_t>(1, B.T_size));
  VMPS_HIP_CHECK(hipMemsetAsync(Tp.p, 0, Tp.n * sizeof(float), stream));  // padding columns stay zero
  upload_bf16_lists(B);
  VMPS_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ_a16, k_stage_a_bf16x3, 256, 0));
  VMPS_HIP_CHECK(hipStreamSynchronize(stream));
  bf16_ready = true;
  ++gen;
}

void DevPlan::upload_bf16_lists(const Bf16Layout& B) {
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
  bfCombTasks.upload(B.comb_tasks, stream);
  bfCombTerms.upload(B.comb_terms, stream);
  if (H.explicit_u)
    Up.alloc(std::max<int64_t>(1, B.U_size));
  else
    Up.release();
}

void DevPlan::retile(int wave_edge, int big_min, int big_min64, bool explicit_u) {
  if (wave_edge == H.wave_edge && big_min == H.big_min && big_min64 == H.big_min64 && explicit_u == H.explicit_u)
    return;  // the lists on the device are already these
  ++gen;
  const bool bf16_lists = bf16_ready && (big_min != H.big_min || explicit_u != H.explicit_u);
  if (explicit_u != H.explicit_u) {
    H.explicit_u = explicit_u;
    build_tiles(H);  // chunk / term tables and every tile list
  }
  set_wave_edge(H, wave_edge, big_min64);
  H.big_min = big_min;
  upload_lists();
  if (bf16_lists) {  // split-BF16 work lists; the split operands stay
    VMPS_HIP_CHECK(hipStreamSynchronize(stream));
    Bf16Layout B;
    build_bf16_layout(H, B);
    upload_bf16_lists(B);
  }
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
      hipLaunchKernelGGL(k_stage_a_bf16x3, dim3((unsigned)persistent_blocks(nA, occ_a16)), dim3(256), 0, s, bfA.p, nA,
                         (const __bf16*)Lhi.p, (const __bf16*)Llo.p, (const __bf16*)Vhi.p, (const __bf16*)Vlo.p, Tp.p);
  } else if (stage == 2) {
    const int nW = (int)wTile.n, nR = (int)rTile.n;
    const float* src = Tp.p;
    if (H.explicit_u) {  // stage B: U = sum c T, then stage C reads U
      if (bfCombTasks.n)
        hipLaunchKernelGGL((k_combine<float, TermC16>), dim3((unsigned)bfCombTasks.n), dim3(256), 0, s,
                           bfCombTasks.p, bfCombTerms.p, Tp.p, Up.p);
      src = Up.p;
    }
    if (nCb && nW && H.explicit_u)
      hipLaunchKernelGGL(k_stage_c_bf16x3_big<true>, dim3((unsigned)nW), dim3(256), 0, s, bfCbig.p, wTile.p,
                         wBegin.p, wEnd.p, wSlot.p, nW, bfC.p, bfT.p, src, (const __bf16*)Rhi.p, (const __bf16*)Rlo.p,
                         d_vout, Cpart.p);
    else if (nCb && nW)
      hipLaunchKernelGGL(k_stage_c_bf16x3_big<false>, dim3((unsigned)nW), dim3(256), 0, s, bfCbig.p, wTile.p,
                         wBegin.p, wEnd.p, wSlot.p, nW, bfC.p, bfT.p, src, (const __bf16*)Rhi.p, (const __bf16*)Rlo.p,
                         d_vout, Cpart.p);
    if (nR)
      hipLaunchKernelGGL(k_reduce_partials, dim3((unsigned)nR), dim3(256), 0, s, bfCbig.p, rTile.p, rSlot.p, rNs.p,
                         Cpart.p, d_vout);
    if (nC)
      hipLaunchKernelGGL(k_stage_c_bf16x3, dim3((unsigned)nC), dim3(256), 0, s, bfCsmall.p, nC, bfC.p, bfT.p, src,
                         (const __bf16*)Rhi.p, (const __bf16*)Rlo.p, d_vout);
  }
  VMPS_HIP_CHECK(hipGetLastError());
}

}  // namespace vmps
