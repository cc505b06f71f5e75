// Batched thin SVD of dense FP64 blocks (the per-sector SVD of the two-site
// split: Biped::truncateSVD, tensors/Biped.h:754-775, called from split_AA2,
// tensors/DmrgContractions.h:2460-2630).
//
// One-sided Jacobi (Hestenes): the columns of W (m x n, m >= n) are rotated in
// pairs until they are mutually orthogonal; then sigma_i = ||w_i||,
// u_i = w_i / sigma_i, and V accumulates the rotations.  Relative accuracy of
// every singular value is at the level of the rounding unit, also for the
// small ones that a Gram-matrix eigensolver would lose.
//
// One workgroup (4 waves) per block.  A sweep visits all n (n - 1) / 2 pairs in
// n' - 1 rounds of n' / 2 disjoint pairs (round-robin tournament, n' = n
// rounded up to even); a wave takes every fourth pair of a round: three dot
// products reduced across the wave, then the rotation of the two columns of W
// and of V.  Rounds are separated by a workgroup barrier; the sweeps stop when
// a whole sweep rotated nothing (|w_p . w_q| <= eps sqrt(|w_p|^2 |w_q|^2)).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "engine.hpp"
#include "vmps_hxv.h"

namespace vmps {

namespace {

struct SvdTask {
  int64_t w_off, v_off, s_off;  // W (m x n), V (n x n), sigma (n) in the work arrays
  int32_t m, n;
};

constexpr double kJacobiEps = 2.0e-16;
constexpr int kMaxSweeps = 60;

__device__ inline double wave_sum(double x) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
  return x;
}

// player at position i of round r in the circle method (position 0 fixed)
__device__ inline int player(int i, int r, int np) { return i == 0 ? 0 : 1 + (i - 1 + r) % (np - 1); }

__global__ __launch_bounds__(256) void k_jacobi_svd(const SvdTask* __restrict__ tasks, double* __restrict__ W,
                                                    double* __restrict__ V, double* __restrict__ sigma,
                                                    int* __restrict__ sweeps_used) {
  __shared__ int rotated_any;
  const SvdTask t = tasks[blockIdx.x];
  const int m = t.m, n = t.n;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  double* __restrict__ Wb = W + t.w_off;
  double* __restrict__ Vb = V + t.v_off;
  const int64_t nn = (int64_t)n * n;
  for (int64_t e = threadIdx.x; e < nn; e += 256) Vb[e] = (e % (n + 1) == 0) ? 1.0 : 0.0;
  const int np = n + (n & 1);
  // rotation threshold on |w_p . w_q| / (|w_p| |w_q|): a computed dot of m
  // terms carries rounding of order sqrt(m) eps, so a threshold of eps alone
  // can stay unreachable and the sweeps run to the limit (the one-sided
  // Jacobi of LAPACK's dgesvj uses sqrt(m) eps the same way)
  const double tol = kJacobiEps * fmax(1.0, sqrt((double)m));
  int sweep = 0;
  for (; sweep < kMaxSweeps && n > 1; ++sweep) {
    if (threadIdx.x == 0) rotated_any = 0;
    __syncthreads();
    int rotated = 0;
    for (int r = 0; r < np - 1; ++r) {
      for (int j = wave; j < np / 2; j += 4) {
        const int p0 = player(j, r, np), q0 = player(np - 1 - j, r, np);
        if (p0 >= n || q0 >= n) continue;  // the dummy player of an odd n
        const int p = min(p0, q0), q = max(p0, q0);
        double* __restrict__ wp = Wb + (int64_t)p * m;
        double* __restrict__ wq = Wb + (int64_t)q * m;
        double a = 0.0, b = 0.0, g = 0.0;
        for (int i = lane; i < m; i += 64) {
          const double x = wp[i], y = wq[i];
          a += x * x;
          b += y * y;
          g += x * y;
        }
        a = wave_sum(a);
        b = wave_sum(b);
        g = wave_sum(g);
        if (!(fabs(g) > tol * sqrt(a * b)) || g == 0.0) continue;
        // rotation that zeroes the (p, q) entry of W^T W
        const double zeta = (b - a) / (2.0 * g);
        const double tt = (zeta >= 0.0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
        const double c = 1.0 / sqrt(1.0 + tt * tt), s = c * tt;
        for (int i = lane; i < m; i += 64) {
          const double x = wp[i], y = wq[i];
          wp[i] = c * x - s * y;
          wq[i] = s * x + c * y;
        }
        double* __restrict__ vp = Vb + (int64_t)p * n;
        double* __restrict__ vq = Vb + (int64_t)q * n;
        for (int i = lane; i < n; i += 64) {
          const double x = vp[i], y = vq[i];
          vp[i] = c * x - s * y;
          vq[i] = s * x + c * y;
        }
        rotated = 1;
      }
      __syncthreads();  // the next round's pairs read columns this round wrote
    }
    if (rotated && lane == 0) rotated_any = 1;
    __syncthreads();
    const int any = rotated_any;
    __syncthreads();
    if (!any) break;
  }
  // column norms = singular values (unsorted)
  for (int col = wave; col < n; col += 4) {
    double a = 0.0;
    for (int i = lane; i < m; i += 64) a += Wb[(int64_t)col * m + i] * Wb[(int64_t)col * m + i];
    a = wave_sum(a);
    if (lane == 0) sigma[t.s_off + col] = sqrt(a);
  }
  if (threadIdx.x == 0) sweeps_used[blockIdx.x] = sweep;
}

// ---- no-allocation variant --------------------------------------------------
// Block b of the packed input A (rows x cols) into the Jacobi work array:
// W = A (rows >= cols) or W = A^T; one workgroup per block.
__global__ __launch_bounds__(256) void k_svd_load(const SvdTask* __restrict__ tasks, const int64_t* __restrict__ a_off,
                                                  const int8_t* __restrict__ transposed, const double* __restrict__ A,
                                                  double* __restrict__ W) {
  const SvdTask t = tasks[blockIdx.x];
  const double* __restrict__ a = A + a_off[blockIdx.x];
  double* __restrict__ w = W + t.w_off;
  const int64_t mn = (int64_t)t.m * t.n;
  if (!transposed[blockIdx.x]) {
    for (int64_t e = threadIdx.x; e < mn; e += 256) w[e] = a[e];
  } else {  // A is n x m (rows x cols): W[j + m i] = A[i + n j]
    for (int64_t e = threadIdx.x; e < mn; e += 256) {
      const int64_t j = e % t.m, i = e / t.m;
      w[e] = a[i + (int64_t)t.n * j];
    }
  }
}

// Singular values sorted descending (ties in column order, as a stable sort),
// U, S and Vt of block b written packed as vmps_block_svd returns them.
constexpr int kSvdFinishMaxN = 8192;
__global__ __launch_bounds__(256) void k_svd_finish(const SvdTask* __restrict__ tasks,
                                                    const int64_t* __restrict__ out_off,
                                                    const int8_t* __restrict__ transposed,
                                                    const double* __restrict__ W, const double* __restrict__ V,
                                                    const double* __restrict__ sigma, double* __restrict__ U,
                                                    double* __restrict__ S, double* __restrict__ Vt) {
  __shared__ int rank[kSvdFinishMaxN];
  const SvdTask t = tasks[blockIdx.x];
  const int m = t.m, n = t.n;
  const double* __restrict__ sg = sigma + t.s_off;
  for (int i = threadIdx.x; i < n; i += 256) {
    const double si = sg[i];
    int r = 0;
    for (int j = 0; j < n; ++j) r += (sg[j] > si) || (sg[j] == si && j < i);
    rank[i] = r;
  }
  __syncthreads();
  const bool tr = transposed[blockIdx.x] != 0;
  const int64_t rows = tr ? n : m, cols = tr ? m : n;
  double* __restrict__ u = U + out_off[3 * blockIdx.x];
  double* __restrict__ sv = S + out_off[3 * blockIdx.x + 1];
  double* __restrict__ vt = Vt + out_off[3 * blockIdx.x + 2];
  const double* __restrict__ w = W + t.w_off;
  const double* __restrict__ v = V + t.v_off;
  for (int i = threadIdx.x; i < n; i += 256) sv[rank[i]] = sg[i];
  // left factor of W: w_i / sigma_i (m x n); right factor V (n x n)
  for (int64_t e = threadIdx.x; e < (int64_t)n * m; e += 256) {
    const int i = (int)(e / m);
    const int64_t x = e % m;
    const double inv = sg[i] > 0.0 ? 1.0 / sg[i] : 0.0;
    const double val = w[x + (int64_t)m * i] * inv;
    if (!tr)
      u[x + rows * rank[i]] = val;        // U = left (rows = m)
    else
      vt[rank[i] + (int64_t)n * x] = val;  // Vt = left^T (cols = m)
  }
  for (int64_t e = threadIdx.x; e < (int64_t)n * n; e += 256) {
    const int i = (int)(e / n);
    const int64_t y = e % n;
    const double val = v[y + (int64_t)n * i];
    if (!tr)
      vt[rank[i] + (int64_t)n * y] = val;  // Vt = V^T (cols = n)
    else
      u[y + rows * rank[i]] = val;         // U = V (rows = n)
  }
  (void)cols;
}

// Scratch of vmps_block_svd_noalloc: one array per device (each GPU holds its
// own copy), reached from the host only through the address
// hipGetSymbolAddress returns for the current device.
constexpr size_t kNoAllocScratchBytes = size_t(64) << 20;
__device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
void* g_scratch_ptr[kMaxDevices] = {};

void* scratch_for_current_device() {
  int dev = 0;
  VMPS_HIP_CHECK(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) throw VmpsError(VMPS_ERR_ARG, "device index beyond the scratch table");
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      throw VmpsError(VMPS_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_ptr[dev] = p;
  }
  return g_scratch_ptr[dev];
}

}  // namespace

}  // namespace vmps

using vmps::VmpsError;

extern "C" vmps_status vmps_block_svd_noalloc(int nblocks, const int64_t* rows, const int64_t* cols,
                                              const double* A, double* U, double* S, double* Vt,
                                              int* max_sweeps) {
  try {
    using vmps::SvdTask;
    if (nblocks < 0 || (nblocks > 0 && (!rows || !cols || !A || !U || !S || !Vt)))
      throw VmpsError(VMPS_ERR_ARG, "bad argument");
    // scratch layout (8-byte words, each region 256-byte aligned): tasks,
    // offsets and flags, A, W, V, sigma, U, S, Vt, sweep counts
    auto align = [](size_t b) { return (b + 255) & ~size_t(255); };
    int64_t asize = 0, wsize = 0, vsize = 0, ssize = 0, usize = 0, vtsize = 0;
    for (int b = 0; b < nblocks; ++b) {
      if (rows[b] < 0 || cols[b] < 0 || rows[b] > (1 << 20) || cols[b] > (1 << 20))
        throw VmpsError(VMPS_ERR_ARG, "bad block dimension");
      const int64_t r = rows[b], c = cols[b], k = std::min(r, c);
      if (std::max(r, c) > 0 && k > vmps::kSvdFinishMaxN)
        throw VmpsError(VMPS_ERR_ARG, "block dimension beyond the no-alloc limit (min(rows, cols) <= 8192)");
      asize += r * c;
      wsize += r * c;
      vsize += k * k;
      ssize += k;
      usize += r * k;
      vtsize += k * c;
    }
    const size_t nb = (size_t)std::max(0, nblocks);
    size_t off = 0;
    const size_t o_tasks = off;
    off = align(off + nb * sizeof(SvdTask));
    const size_t o_aoff = off;
    off = align(off + nb * sizeof(int64_t));
    const size_t o_ooff = off;
    off = align(off + 3 * nb * sizeof(int64_t));
    const size_t o_tr = off;
    off = align(off + nb);
    const size_t o_A = off;
    off = align(off + (size_t)asize * 8);
    const size_t o_W = off;
    off = align(off + (size_t)wsize * 8);
    const size_t o_V = off;
    off = align(off + (size_t)vsize * 8);
    const size_t o_sig = off;
    off = align(off + (size_t)ssize * 8);
    const size_t o_U = off;
    off = align(off + (size_t)usize * 8);
    const size_t o_S = off;
    off = align(off + (size_t)ssize * 8);
    const size_t o_Vt = off;
    off = align(off + (size_t)vtsize * 8);
    const size_t o_sw = off;
    off = align(off + nb * sizeof(int));
    if (off > vmps::kNoAllocScratchBytes)
      throw VmpsError(VMPS_ERR_ARG, "blocks exceed the 64 MiB scratch of vmps_block_svd_noalloc");
    if (nblocks == 0 || wsize == 0) return VMPS_OK;
    unsigned char* base = static_cast<unsigned char*>(vmps::scratch_for_current_device());
    // host descriptor tables (small: a few words per block)
    std::vector<SvdTask> tasks(nb);
    std::vector<int64_t> aoff(nb), ooff(3 * nb);
    std::vector<int8_t> tr(nb);
    int64_t ao = 0, wo = 0, vo = 0, so = 0, uo = 0, vto = 0;
    for (int b = 0; b < nblocks; ++b) {
      const int64_t r = rows[b], c = cols[b];
      tr[b] = r < c;
      const int64_t m = tr[b] ? c : r, n = tr[b] ? r : c;
      tasks[b] = {wo, vo, so, (int32_t)m, (int32_t)n};
      aoff[b] = ao;
      ooff[3 * b] = uo;
      ooff[3 * b + 1] = so;
      ooff[3 * b + 2] = vto;
      ao += r * c;
      wo += m * n;
      vo += n * n;
      so += n;
      uo += r * n;
      vto += n * c;
    }
    hipStream_t s = nullptr;  // the null stream: synchronous with the copies below
    VMPS_HIP_CHECK(hipMemcpy(base + o_tasks, tasks.data(), nb * sizeof(SvdTask), hipMemcpyHostToDevice));
    VMPS_HIP_CHECK(hipMemcpy(base + o_aoff, aoff.data(), nb * sizeof(int64_t), hipMemcpyHostToDevice));
    VMPS_HIP_CHECK(hipMemcpy(base + o_ooff, ooff.data(), 3 * nb * sizeof(int64_t), hipMemcpyHostToDevice));
    VMPS_HIP_CHECK(hipMemcpy(base + o_tr, tr.data(), nb, hipMemcpyHostToDevice));
    VMPS_HIP_CHECK(hipMemcpy(base + o_A, A, (size_t)asize * 8, hipMemcpyHostToDevice));
    auto* dT = reinterpret_cast<const SvdTask*>(base + o_tasks);
    auto* dW = reinterpret_cast<double*>(base + o_W);
    auto* dV = reinterpret_cast<double*>(base + o_V);
    auto* dsig = reinterpret_cast<double*>(base + o_sig);
    auto* dsw = reinterpret_cast<int*>(base + o_sw);
    hipLaunchKernelGGL(vmps::k_svd_load, dim3((unsigned)nblocks), dim3(256), 0, s, dT,
                       reinterpret_cast<const int64_t*>(base + o_aoff), reinterpret_cast<const int8_t*>(base + o_tr),
                       reinterpret_cast<const double*>(base + o_A), dW);
    hipLaunchKernelGGL(vmps::k_jacobi_svd, dim3((unsigned)nblocks), dim3(256), 0, s, dT, dW, dV, dsig, dsw);
    hipLaunchKernelGGL(vmps::k_svd_finish, dim3((unsigned)nblocks), dim3(256), 0, s, dT,
                       reinterpret_cast<const int64_t*>(base + o_ooff), reinterpret_cast<const int8_t*>(base + o_tr),
                       dW, dV, dsig, reinterpret_cast<double*>(base + o_U), reinterpret_cast<double*>(base + o_S),
                       reinterpret_cast<double*>(base + o_Vt));
    VMPS_HIP_CHECK(hipGetLastError());
    VMPS_HIP_CHECK(hipMemcpy(U, base + o_U, (size_t)usize * 8, hipMemcpyDeviceToHost));
    VMPS_HIP_CHECK(hipMemcpy(S, base + o_S, (size_t)ssize * 8, hipMemcpyDeviceToHost));
    VMPS_HIP_CHECK(hipMemcpy(Vt, base + o_Vt, (size_t)vtsize * 8, hipMemcpyDeviceToHost));
    std::vector<int> sweeps(nb);
    VMPS_HIP_CHECK(hipMemcpy(sweeps.data(), base + o_sw, nb * sizeof(int), hipMemcpyDeviceToHost));
    int worst = 0;
    for (int sw : sweeps) worst = std::max(worst, sw);
    if (max_sweeps) *max_sweeps = worst;
    if (worst >= vmps::kMaxSweeps) {
      vmps::set_last_error("block SVD: a block was still rotating after the Jacobi sweep limit");
      return VMPS_ERR_NOCONV;
    }
    return VMPS_OK;
  } catch (const VmpsError& e) {
    vmps::set_last_error(e.what());
    return e.status;
  } catch (const std::exception& e) {
    vmps::set_last_error(e.what());
    return VMPS_ERR_INTERNAL;
  }
}

extern "C" vmps_status vmps_block_svd(int nblocks, const int64_t* rows, const int64_t* cols, const double* A,
                                      double* U, double* S, double* Vt, int* max_sweeps) {
  try {
    if (nblocks < 0 || (nblocks > 0 && (!rows || !cols || !A || !U || !S || !Vt)))
      throw VmpsError(VMPS_ERR_ARG, "bad argument");
    // host layout: every block as W = A (m >= n) or A^T (m < n), W column-major
    std::vector<vmps::SvdTask> tasks;
    std::vector<char> transposed;
    int64_t wsize = 0, vsize = 0, ssize = 0;
    for (int b = 0; b < nblocks; ++b) {
      if (rows[b] < 0 || cols[b] < 0 || rows[b] > (1 << 20) || cols[b] > (1 << 20))
        throw VmpsError(VMPS_ERR_ARG, "bad block dimension");
      const bool tr = rows[b] < cols[b];
      const int64_t m = tr ? cols[b] : rows[b], n = tr ? rows[b] : cols[b];
      tasks.push_back({wsize, vsize, ssize, (int32_t)m, (int32_t)n});
      transposed.push_back(tr);
      wsize += m * n;
      vsize += n * n;
      ssize += n;
    }
    std::vector<double> hW((size_t)std::max<int64_t>(1, wsize));
    int64_t ao = 0;
    for (int b = 0; b < nblocks; ++b) {
      const int64_t r = rows[b], c = cols[b];
      double* w = hW.data() + tasks[b].w_off;
      if (!transposed[b]) {
        std::memcpy(w, A + ao, sizeof(double) * r * c);
      } else {  // W = A^T (c x r), column-major: W[j + c i] = A[i + r j]
        for (int64_t j = 0; j < c; ++j)
          for (int64_t i = 0; i < r; ++i) w[j + c * i] = A[ao + i + r * j];
      }
      ao += r * c;
    }
    if (nblocks == 0 || wsize == 0) {
      // empty blocks: nothing to factor
      return VMPS_OK;
    }
    struct Stream {  // synchronised and returned to the pool on every exit path
      hipStream_t s = nullptr;
      ~Stream() {
        if (s) {
          (void)hipStreamSynchronize(s);
          vmps::release_stream(s);
        }
      }
    } st;
    st.s = vmps::acquire_stream();
    hipStream_t s = st.s;
    vmps::DevBuf<vmps::SvdTask> dT;
    vmps::DevBuf<double> dW, dV, dS;
    vmps::DevBuf<int> dSw;
    dT.upload(tasks, s);
    dW.upload(hW, s);
    dV.alloc((size_t)std::max<int64_t>(1, vsize));
    dS.alloc((size_t)std::max<int64_t>(1, ssize));
    dSw.alloc((size_t)nblocks);
    hipLaunchKernelGGL(vmps::k_jacobi_svd, dim3((unsigned)nblocks), dim3(256), 0, s, dT.p, dW.p, dV.p, dS.p, dSw.p);
    VMPS_HIP_CHECK(hipGetLastError());
    std::vector<double> hV((size_t)std::max<int64_t>(1, vsize)), hS((size_t)std::max<int64_t>(1, ssize));
    std::vector<int> hSw(nblocks);
    VMPS_HIP_CHECK(hipMemcpyAsync(hW.data(), dW.p, sizeof(double) * wsize, hipMemcpyDeviceToHost, s));
    if (vsize) VMPS_HIP_CHECK(hipMemcpyAsync(hV.data(), dV.p, sizeof(double) * vsize, hipMemcpyDeviceToHost, s));
    if (ssize) VMPS_HIP_CHECK(hipMemcpyAsync(hS.data(), dS.p, sizeof(double) * ssize, hipMemcpyDeviceToHost, s));
    VMPS_HIP_CHECK(hipMemcpyAsync(hSw.data(), dSw.p, sizeof(int) * nblocks, hipMemcpyDeviceToHost, s));
    VMPS_HIP_CHECK(hipStreamSynchronize(s));
    // sort each block's singular values descending and write U, S, Vt
    int64_t uo = 0, so = 0, vto = 0;
    int worst = 0;
    for (int b = 0; b < nblocks; ++b) {
      const vmps::SvdTask& t = tasks[b];
      const int64_t m = t.m, n = t.n, r = rows[b], c = cols[b];
      worst = std::max(worst, hSw[b]);
      std::vector<int64_t> ord(n);
      for (int64_t i = 0; i < n; ++i) ord[i] = i;
      const double* sg = hS.data() + t.s_off;
      std::stable_sort(ord.begin(), ord.end(), [&](int64_t x, int64_t y) { return sg[x] > sg[y]; });
      const double* w = hW.data() + t.w_off;
      const double* v = hV.data() + t.v_off;
      // left factor of W: w_i / sigma_i (m x n); right factor: V (n x n)
      // A = W (not transposed): U = left (r x n), Vt = V^T (n x c)
      // A^T = W: A = V S left^T: U = V (r x n, r = n), Vt = left^T (n x c, c = m)
      for (int64_t k = 0; k < n; ++k) {
        const int64_t i = ord[k];
        const double sv = sg[i];
        S[so + k] = sv;
        const double inv = sv > 0.0 ? 1.0 / sv : 0.0;
        if (!transposed[b]) {
          for (int64_t x = 0; x < r; ++x) U[uo + x + r * k] = w[x + m * i] * inv;
          for (int64_t y = 0; y < c; ++y) Vt[vto + k + n * y] = v[y + n * i];
        } else {
          for (int64_t x = 0; x < r; ++x) U[uo + x + r * k] = v[x + n * i];
          for (int64_t y = 0; y < c; ++y) Vt[vto + k + n * y] = w[y + m * i] * inv;
        }
      }
      uo += r * n;
      so += n;
      vto += n * c;
    }
    if (max_sweeps) *max_sweeps = worst;
    if (worst >= vmps::kMaxSweeps) {
      vmps::set_last_error("block SVD: a block was still rotating after the Jacobi sweep limit");
      return VMPS_ERR_NOCONV;
    }
    return VMPS_OK;
  } catch (const VmpsError& e) {
    vmps::set_last_error(e.what());
    return e.status;
  } catch (const std::exception& e) {
    vmps::set_last_error(e.what());
    return VMPS_ERR_INTERNAL;
  }
}
