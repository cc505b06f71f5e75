// Device engine for the staged effective-Hamiltonian application on gfx950.
//
//   stage A   T_t   = L_l . V_v                     one problem per distinct (l, v)
//   stage C   out_o = sum_r (sum_j c_j T_j) . R_r    one problem per output block;
//             the scalar combine U = sum c T is done while loading the MFMA operand
//
// Work unit = one 64-lane wavefront computing one 32 x 32 output tile with four
// 16x16 MFMA accumulators.  Operands go from global memory straight to
// registers (every element of a wave's A and B tiles is used by exactly one
// MFMA, so staging in LDS would buy no reuse); the tile descriptor is a single
// load.  Every output element is owned by exactly one wave, so there are no
// atomics and results are bitwise reproducible run to run.
#include "engine.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>

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

template <class T, class OutT>
__device__ inline void store_tile(OutT* __restrict__ C, int M, int N, int m0, int n0, int lane,
                                  const typename Mfma<T>::acc_t& c00, const typename Mfma<T>::acc_t& c01,
                                  const typename Mfma<T>::acc_t& c10, const typename Mfma<T>::acc_t& c11) {
  const int i0 = m0 + (lane & 15), i1 = i0 + 16;
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    const int j0 = n0 + Mfma<T>::col(lane, v), j1 = j0 + 16;
    if (j0 < N) {
      if (i0 < M) C[(int64_t)j0 * M + i0] = (OutT)c00[v];
      if (i1 < M) C[(int64_t)j0 * M + i1] = (OutT)c10[v];
    }
    if (j1 < N) {
      if (i0 < M) C[(int64_t)j1 * M + i0] = (OutT)c01[v];
      if (i1 < M) C[(int64_t)j1 * M + i1] = (OutT)c11[v];
    }
  }
}

// Stage A: T (M x N, ld M) = L (M x K, ld M) . V (K x N, ld K), one tile per wave.
template <class T>
__global__ __launch_bounds__(256) void k_stage_a(const TileA* __restrict__ tiles, int ntiles, const T* __restrict__ L,
                                                 const T* __restrict__ V, T* __restrict__ Tout) {
  typedef typename Mfma<T>::acc_t acc_t;
  const int w = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (w >= ntiles) return;
  const int lane = threadIdx.x & 63;
  const TileA t = tiles[w];
  const int m0 = t.mt * kTile, n0 = t.nt * kTile;
  const int li = lane & 15, lk = lane >> 4;
  const int i0 = m0 + li, i1 = i0 + 16, j0 = n0 + li, j1 = j0 + 16;
  const bool has_m1 = m0 + 16 < t.M, has_n1 = n0 + 16 < t.N;  // wave-uniform
  const T* __restrict__ A = L + t.a_off;
  const T* __restrict__ B = V + t.b_off;
  acc_t c00 = {0, 0, 0, 0}, c01 = {0, 0, 0, 0}, c10 = {0, 0, 0, 0}, c11 = {0, 0, 0, 0};
  for (int kb = 0; kb < t.K; kb += 16) {
    T a0[4], a1[4], b0[4], b1[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {  // all loads of the 16-deep block first
      const int k = kb + 4 * u + lk;
      const bool kin = k < t.K;
      a0[u] = (kin && i0 < t.M) ? A[(int64_t)k * t.M + i0] : T(0);
      a1[u] = (kin && i1 < t.M) ? A[(int64_t)k * t.M + i1] : T(0);
      b0[u] = (kin && j0 < t.N) ? B[(int64_t)j0 * t.K + k] : T(0);
      b1[u] = (kin && j1 < t.N) ? B[(int64_t)j1 * t.K + k] : T(0);
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      Mfma<T>::mac(b0[u], a0[u], c00);
      if (has_m1) Mfma<T>::mac(b0[u], a1[u], c10);
      if (has_n1) {
        Mfma<T>::mac(b1[u], a0[u], c01);
        if (has_m1) Mfma<T>::mac(b1[u], a1[u], c11);
      }
    }
  }
  store_tile<T, T>(Tout + t.c_off, t.M, t.N, m0, n0, lane, c00, c01, c10, c11);
}

// Stage C: Vout (M x N, ld M) = sum over chunks (sum_j c_j T_j)(M x K, ld M) . R (K x N, ld K).
// One workgroup per output tile.  The iteration space (chunk, 16-deep K block)
// is dealt round-robin to the 4 waves (split K), each K block issues all its
// loads before its 16 MFMAs, and the 4 partial accumulators are summed through
// LDS in wave order, so the result does not depend on timing.
constexpr int kKU = 16;  // K depth per unrolled step (4 MFMA k-steps)

template <class T, class OutT>
__global__ __launch_bounds__(256) void k_stage_c(const TileC* __restrict__ tiles, int ntiles,
                                                 const ChunkC* __restrict__ chunks, const TermC* __restrict__ terms,
                                                 const T* __restrict__ Tbuf, const T* __restrict__ R,
                                                 OutT* __restrict__ Vout) {
  typedef typename Mfma<T>::acc_t acc_t;
  __shared__ acc_t red[3][4][64];
  if ((int)blockIdx.x >= ntiles) return;  // workgroup-uniform
  const int wave = threadIdx.x >> 6;
  const int lane = threadIdx.x & 63;
  const TileC t = tiles[blockIdx.x];
  const int m0 = t.mt * kTile, n0 = t.nt * kTile;
  const int li = lane & 15, lk = lane >> 4;
  const int i0 = m0 + li, i1 = i0 + 16, j0 = n0 + li, j1 = j0 + 16;
  const bool has_m1 = m0 + 16 < t.M, has_n1 = n0 + 16 < t.N;
  const bool r0 = i0 < t.M, r1 = i1 < t.M, q0 = j0 < t.N, q1 = j1 < t.N;
  acc_t c00 = {0, 0, 0, 0}, c01 = {0, 0, 0, 0}, c10 = {0, 0, 0, 0}, c11 = {0, 0, 0, 0};
  int it = 0;
  for (int c = t.chunk_begin; c < t.chunk_end; ++c) {
    const ChunkC ch = chunks[c];
    const T* __restrict__ B = R + ch.r_off;
    for (int kb = 0; kb < ch.K; kb += kKU, ++it) {
      if ((it & 3) != wave) continue;
      T a0[4], a1[4], b0[4], b1[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int k = kb + 4 * u + lk;
        const bool kin = k < ch.K;
        a0[u] = T(0);
        a1[u] = T(0);
        b0[u] = (kin && q0) ? B[(int64_t)j0 * ch.K + k] : T(0);
        b1[u] = (kin && q1) ? B[(int64_t)j1 * ch.K + k] : T(0);
      }
      for (int j = ch.term_begin; j < ch.term_end; ++j) {
        const TermC tm = terms[j];
        const T cj = (T)tm.c;
        const T* __restrict__ Tt = Tbuf + tm.t_off;
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const int k = kb + 4 * u + lk;
          const bool kin = k < ch.K;
          const T x0 = (kin && r0) ? Tt[(int64_t)k * t.M + i0] : T(0);
          const T x1 = (kin && r1) ? Tt[(int64_t)k * t.M + i1] : T(0);
          a0[u] += cj * x0;
          a1[u] += cj * x1;
        }
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        Mfma<T>::mac(b0[u], a0[u], c00);
        if (has_m1) Mfma<T>::mac(b0[u], a1[u], c10);
        if (has_n1) {
          Mfma<T>::mac(b1[u], a0[u], c01);
          if (has_m1) Mfma<T>::mac(b1[u], a1[u], c11);
        }
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
    store_tile<T, OutT>(Vout + t.c_off, t.M, t.N, m0, n0, lane, c00, c01, c10, c11);
  }
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

DevPlan::~DevPlan() {
  if (stream) {
    (void)hipStreamSynchronize(stream);
    (void)hipStreamDestroy(stream);
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

void DevPlan::upload(vmps_precision precision) {
  VMPS_HIP_CHECK(hipGetDevice(&device));
  VMPS_HIP_CHECK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
  VMPS_HIP_CHECK(hipEventCreate(&ev0));
  VMPS_HIP_CHECK(hipEventCreate(&ev1));
  L.upload(H.Ldata, stream);
  R.upload(H.Rdata, stream);
  T.alloc(H.T_size);
  vin.alloc(H.ysize);
  vout.alloc(H.osize);
  if (H.ysize || H.osize) {
    VMPS_HIP_CHECK(hipMemsetAsync(vin.p, 0, std::max<int64_t>(1, H.ysize) * sizeof(double), stream));
    VMPS_HIP_CHECK(hipMemsetAsync(vout.p, 0, std::max<int64_t>(1, H.osize) * sizeof(double), stream));
  }
  tilesA.upload(H.tilesA, stream);
  tilesC.upload(H.tilesC, stream);
  chunksC.upload(H.chunksC, stream);
  termsC.upload(H.termsC, stream);
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

void DevPlan::ensure_precision_buffers() {
  if (prec == VMPS_PREC_FP32 && !T32.p && (H.T_size || H.vsize)) {
    L32.alloc(H.Ldata.size());
    R32.alloc(H.Rdata.size());
    V32.alloc(H.ysize);
    T32.alloc(H.T_size);
    if (L.n) hipLaunchKernelGGL(k_f64_to_f32, dim3(1024), dim3(256), 0, stream, L.p, L32.p, (int64_t)L.n);
    if (R.n) hipLaunchKernelGGL(k_f64_to_f32, dim3(1024), dim3(256), 0, stream, R.p, R32.p, (int64_t)R.n);
    VMPS_HIP_CHECK(hipGetLastError());
  }
  if (prec == VMPS_PREC_BF16X3) ensure_bf16_buffers();
}

void DevPlan::launch_stage(int stage, const double* d_vin, double* d_vout, hipStream_t s) {
  const int nA = (int)tilesA.n, nC = (int)tilesC.n;
  if (prec == VMPS_PREC_FP64) {
    if (stage == 0 && nA)
      hipLaunchKernelGGL(k_stage_a<double>, dim3(grid_for(nA)), dim3(256), 0, s, tilesA.p, nA, L.p, d_vin, T.p);
    if (stage == 2) {
      if (nC)
        hipLaunchKernelGGL((k_stage_c<double, double>), dim3(nC), dim3(256), 0, s, tilesC.p, nC, chunksC.p,
                         termsC.p, T.p, R.p, d_vout);
    }
  } else if (prec == VMPS_PREC_FP32) {
    if (stage == 0) {
      if (H.ysize) hipLaunchKernelGGL(k_f64_to_f32, dim3(1024), dim3(256), 0, s, d_vin, V32.p, (int64_t)H.ysize);
      if (nA)
        hipLaunchKernelGGL(k_stage_a<float>, dim3(grid_for(nA)), dim3(256), 0, s, tilesA.p, nA, L32.p, V32.p, T32.p);
    }
    if (stage == 2 && nC)
      hipLaunchKernelGGL((k_stage_c<float, double>), dim3(nC), dim3(256), 0, s, tilesC.p, nC, chunksC.p, termsC.p,
                         T32.p, R32.p, d_vout);
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
