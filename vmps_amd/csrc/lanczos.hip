// Krylov eigensolver and PivotVector algebra on device-resident vectors.
//
// The reference instantiates an external LanczosSolver<PivotMatrix2,
// PivotVector, double> (called at solvers/DmrgSolver.h:1466-1471 with
// REORTHO=FULL, tol_eigval 1e-8, tol_state 1e-5, dimK 500 from
// DmrgTypedefs.h:370-373); only the vector operations it needs are in the
// reference (pivot/DmrgPivotVector.h:121-424).  This file restates the method:
// Lanczos with full re-orthogonalisation (classical Gram-Schmidt applied twice)
// under the coeff_dot-weighted inner product, the lowest Ritz pair of the
// tridiagonal matrix by Sturm bisection plus inverse iteration, and the
// residual ||H psi - E psi|| evaluated with one extra application.
#include <cmath>
#include <map>
#include <mutex>
#include <cstring>
#include <vector>

#include "engine.hpp"
#include "lanczos.hpp"

namespace vmps {

namespace {

// y = sum_{i<k} c[i] Q_i   (Ritz vector)
__global__ void k_combination(const double* __restrict__ Q, int64_t n, const double* __restrict__ c, int k,
                              double* __restrict__ y) {
  int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (; e < n; e += stride) {
    double s = 0.0;
    for (int i = 0; i < k; ++i) s += c[i] * Q[(int64_t)i * n + e];
    y[e] = s;
  }
}

// y = a * x + b * y ; a, b host scalars
__global__ void k_axpby(double a, const double* __restrict__ x, double b, double* __restrict__ y, int64_t n) {
  int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (; e < n; e += stride) y[e] = a * x[e] + b * y[e];
}

// y = x / sqrt(nrm2[0])
__global__ void k_normalize_into(const double* __restrict__ nrm2, const double* __restrict__ x, double* __restrict__ y,
                                 int64_t n) {
  const double s = nrm2[0] > 0.0 ? 1.0 / sqrt(nrm2[0]) : 0.0;  // a zero vector stays zero
  int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (; e < n; e += stride) y[e] = x[e] * s;
}

// Uniform(-1,1) fill from a counter-based hash (splitmix64 of seed and index).
__global__ void k_fill_uniform(double* __restrict__ v, int64_t n, uint64_t seed) {
  int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (; e < n; e += stride) {
    uint64_t z = seed + 0x9E3779B97F4A7C15ull * (uint64_t)(e + 1);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    v[e] = 2.0 * ((double)(z >> 11) * (1.0 / 9007199254740992.0)) - 1.0;
  }
}

inline unsigned vgrid(int64_t n) { return (unsigned)std::max<int64_t>(1, std::min<int64_t>(2048, (n + 255) / 256)); }

// Sum over the 256 threads of a workgroup in a fixed order (shuffles within
// each wave, then the four wave sums); every thread returns it.
__device__ inline double wg_sum(double x) {
  __shared__ double ws[4];
  for (int m = 32; m > 0; m >>= 1) x += __shfl_xor(x, m);
  if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = x;
  __syncthreads();
  const double r = (ws[0] + ws[1]) + (ws[2] + ws[3]);
  __syncthreads();  // ws is free for the next call
  return r;
}

// ---- kernels of the graph-captured iteration ----------------------------------
// The iteration index k lives in device memory: an iteration's kernels read
// k = *kdev + kofs, where kofs is the iteration's place in a captured batch
// (0 .. kIterBatch - 1), and one k_add_int after the batch advances *kdev.  One
// captured graph thus serves every batch.

// partial[i * nb + b] = sum over elements e of chunk b of ew[e] Q_i[e] w[e]
// (ew: the coeff_dot weight of e's block), for the rows i = 4 y .. 4 y + 3
// that are <= k; workgroup (b, y).  Groups past k exit at once, so the grid is
// sized for the largest k and stays the same in every iteration.
// With `local`, only the rows k-1 and k (one row group): the local
// orthogonalisation of the partially re-orthogonalised iteration.
__global__ __launch_bounds__(256) void k_multidot_rows(const int* __restrict__ kdev, int kofs,
                                                       const double* __restrict__ Q, int64_t n,
                                                       const double* __restrict__ w,
                                                       const double* __restrict__ ew, int64_t chunk,
                                                       double* __restrict__ partial, int nb, int local) {
  const int k = *kdev + kofs;
  const int i0 = (local ? max(0, k - 1) : 0) + 4 * blockIdx.y;
  if (i0 > k) return;  // workgroup-uniform
  __shared__ double ws[4][4];  // [wave][row]
  const int nr = min(4, k + 1 - i0);
  const int b = blockIdx.x;
  const int64_t e0 = (int64_t)b * chunk, e1 = min(n, e0 + chunk);
  // rows past k read row k again (their sums are not stored), so that the
  // loop has no branch and the loads of several elements are in flight together
  const double* __restrict__ q0 = Q + (int64_t)i0 * n;
  const double* __restrict__ q1 = Q + (int64_t)min(i0 + 1, k) * n;
  const double* __restrict__ q2 = Q + (int64_t)min(i0 + 2, k) * n;
  const double* __restrict__ q3 = Q + (int64_t)min(i0 + 3, k) * n;
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll 4
  for (int64_t e = e0 + threadIdx.x; e < e1; e += 256) {
    const double we = w[e] * ew[e];
    acc[0] += q0[e] * we;
    acc[1] += q1[e] * we;
    acc[2] += q2[e] * we;
    acc[3] += q3[e] * we;
  }
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    double x = acc[r];
    for (int m = 32; m > 0; m >>= 1) x += __shfl_xor(x, m);
    if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6][r] = x;
  }
  __syncthreads();
  const int t = threadIdx.x;
  if (t < nr) partial[(int64_t)(i0 + t) * nb + b] = (ws[0][t] + ws[1][t]) + (ws[2][t] + ws[3][t]);
}

// c[i] = sum_b partial[i * nb + b] for i <= k; one workgroup per possible row.
__global__ __launch_bounds__(256) void k_multidot_rows_final(const int* __restrict__ kdev, int kofs,
                                                             const double* __restrict__ partial, int nb,
                                                             double* __restrict__ c) {
  const int i = blockIdx.x;
  if (i > *kdev + kofs) return;  // workgroup-uniform
  __shared__ double red[256];
  double acc = 0.0;
  for (int b = threadIdx.x; b < nb; b += 256) acc += partial[(int64_t)i * nb + b];
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) c[i] = red[0];
}

// sum_{r0<=i<=k} c[i] Q_i[e], added in order of i, with the basis loads issued
// eight rows at a time (eight memory round trips in flight per thread
// instead of the compiler's four)
__device__ inline double combination_at(const double* __restrict__ Q, int64_t n, const double* c, int k,
                                        int64_t e, int r0 = 0) {
  double s = 0.0;
  int i = r0;
  for (; i + 8 <= k + 1; i += 8) {
    double q[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) q[u] = Q[(int64_t)(i + u) * n + e];
#pragma unroll
    for (int u = 0; u < 8; ++u) s += c[i + u] * q[u];
  }
  for (; i <= k; ++i) s += c[i] * Q[(int64_t)i * n + e];
  return s;
}

// w -= sum_{i<=k} c[i] Q_i; with norm_part, also the workgroup's share of the
// weighted ||w||^2 of the result (summed by k_normalize_next)
__global__ __launch_bounds__(256) void k_sub_combination_dk(const double* __restrict__ Q, int64_t n,
                                                            const double* __restrict__ c, const int* __restrict__ kdev,
                                                            int kofs, double* __restrict__ w,
                                                            const double* __restrict__ ew,
                                                            double* __restrict__ norm_part) {
  const int k = *kdev + kofs;
  int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  double nrm = 0.0;
  for (; e < n; e += stride) {
    const double s = combination_at(Q, n, c, k, e);
    const double y = w[e] - s;
    w[e] = y;
    nrm += ew[e] * y * y;
  }
  if (!norm_part) return;  // workgroup-uniform
  const double sum = wg_sum(nrm);
  if (threadIdx.x == 0) norm_part[blockIdx.x] = sum;
}

// k_multidot_rows_final and k_sub_combination_dk in one kernel, for short
// partial lists (nb <= kFuseMaxGroups, k < kFuseMaxRows): every workgroup sums
// the nb partials of each row c[i], i <= k, into LDS (the same sums, in the
// same order, in every workgroup); workgroup 0 also stores them in c for the
// host scalars.  Then w -= sum_i c[i] Q_i as in k_sub_combination_dk.
constexpr int kFuseMaxGroups = 64;
constexpr int kFuseMaxRows = 512;
// With `local`, the rows k-1 and k only; with `gate`, the kernel does
// nothing unless gate[k] is set (the partially re-orthogonalised iteration's
// full pass).
__global__ __launch_bounds__(256) void k_sub_combination_fused(const double* __restrict__ Q, int64_t n,
                                                               const double* __restrict__ partial, int nb,
                                                               const int* __restrict__ kdev, int kofs,
                                                               double* __restrict__ c, double* __restrict__ w,
                                                               const double* __restrict__ ew,
                                                               double* __restrict__ norm_part, int local,
                                                               const int* __restrict__ gate, int poff) {
  __shared__ double cs[kFuseMaxRows];
  const int k = *kdev + kofs;
  if (gate && !gate[k]) return;  // workgroup-uniform
  const int r0 = local ? max(0, k - 1) : 0;
  partial += (int64_t)poff * nb;  // row i's partials at row i + poff
  // four threads per row, each summing every fourth partial, then the four
  // sums combined by shuffles within the quad: a fixed order in every
  // workgroup, and a quarter of the dependent load chain per thread
  // (all of a thread's partials are loaded before the first add, one memory
  // round trip instead of one per partial; absent ones add as 0.0)
  const int quad = threadIdx.x & 3;
  for (int i = r0 + (threadIdx.x >> 2); i <= k; i += 64) {  // quad-uniform
    double p[kFuseMaxGroups / 4];
#pragma unroll
    for (int j = 0; j < kFuseMaxGroups / 4; ++j) {
      const int b = quad + 4 * j;
      p[j] = b < nb ? partial[(int64_t)i * nb + b] : 0.0;
    }
    double acc = 0.0;
#pragma unroll
    for (int j = 0; j < kFuseMaxGroups / 4; ++j) acc += p[j];
    acc += __shfl_xor(acc, 1);
    acc += __shfl_xor(acc, 2);
    if (quad == 0) {
      cs[i] = acc;
      if (blockIdx.x == 0) c[i] = acc;
    }
  }
  __syncthreads();
  int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  double nrm = 0.0;
  for (; e < n; e += stride) {
    const double s = combination_at(Q, n, cs, k, e, r0);
    const double y = w[e] - s;
    w[e] = y;
    nrm += ew[e] * y * y;
  }
  if (!norm_part) return;  // workgroup-uniform
  const double sum = wg_sum(nrm);
  if (threadIdx.x == 0) norm_part[blockIdx.x] = sum;
}

// The first local pass's subtraction and the second local pass's dots in one
// kernel (partially re-orthogonalised iterations, fused path): workgroup b
// sums the partials of rows r0 = max(0, k-1) .. k as k_sub_combination_fused
// does (workgroup 0 stores them in c), subtracts sum_i c[i] Q_i from w over
// chunk b, and forms the chunk's dot partials of the result with the same
// rows, into rows i + 2 of `partial` (the rows the subtraction reads stay
// intact for the other workgroups).  Elements, their order per thread and the
// reductions are those of k_multidot_rows, so the partials are bit-identical
// to a k_sub_combination_fused + k_multidot_rows pair.
__global__ __launch_bounds__(256) void k_sub_local_dots(const double* __restrict__ Q, int64_t n,
                                                        double* __restrict__ partial, int nb,
                                                        const int* __restrict__ kdev, int kofs,
                                                        double* __restrict__ c, double* __restrict__ w,
                                                        const double* __restrict__ ew, int64_t chunk) {
  __shared__ double cs[2];
  __shared__ double ws[4][2];
  const int k = *kdev + kofs;
  const int r0 = max(0, k - 1);
  const int quad = threadIdx.x & 3;
  const int ir = threadIdx.x >> 2;
  if (r0 + ir <= k) {  // quad-uniform: threads 0..7 (rows r0, r0 + 1)
    const int i = r0 + ir;
    double p[kFuseMaxGroups / 4];
#pragma unroll
    for (int j = 0; j < kFuseMaxGroups / 4; ++j) {
      const int b = quad + 4 * j;
      p[j] = b < nb ? partial[(int64_t)i * nb + b] : 0.0;
    }
    double acc = 0.0;
#pragma unroll
    for (int j = 0; j < kFuseMaxGroups / 4; ++j) acc += p[j];
    acc += __shfl_xor(acc, 1);
    acc += __shfl_xor(acc, 2);
    if (quad == 0) {
      cs[ir] = acc;
      if (blockIdx.x == 0) c[i] = acc;
    }
  }
  __syncthreads();
  const int nr = k + 1 - r0;  // 1 or 2
  const double c0 = cs[0], c1 = nr > 1 ? cs[1] : 0.0;
  const double* __restrict__ q0 = Q + (int64_t)r0 * n;
  const double* __restrict__ q1 = Q + (int64_t)k * n;  // row r0 again when k = 0
  const int b = blockIdx.x;
  const int64_t e0 = (int64_t)b * chunk, e1 = min(n, e0 + chunk);
  double a0 = 0.0, a1 = 0.0;
#pragma unroll 4
  for (int64_t e = e0 + threadIdx.x; e < e1; e += 256) {
    const double x0 = q0[e], x1 = q1[e];
    double s = 0.0;
    s += c0 * x0;
    if (nr > 1) s += c1 * x1;
    const double y = w[e] - s;
    w[e] = y;
    const double we = y * ew[e];
    a0 += x0 * we;
    a1 += x1 * we;
  }
  for (int m = 32; m > 0; m >>= 1) {
    a0 += __shfl_xor(a0, m);
    a1 += __shfl_xor(a1, m);
  }
  if ((threadIdx.x & 63) == 0) {
    ws[threadIdx.x >> 6][0] = a0;
    ws[threadIdx.x >> 6][1] = a1;
  }
  __syncthreads();
  const int t = threadIdx.x;
  if (t < nr) partial[(int64_t)(r0 + t + 2) * nb + b] = (ws[0][t] + ws[1][t]) + (ws[2][t] + ws[3][t]);
}

// sum of part[0 .. np-1] (np <= 2048, the largest vgrid) in one workgroup,
// fixed order; every thread returns it.  A thread's eight loads are issued
// together; absent ones add as 0.0.
__device__ inline double block_sum(const double* __restrict__ part, int np) {
  double p[8];
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    const int b = (int)threadIdx.x + 256 * j;
    p[j] = b < np ? part[b] : 0.0;
  }
  double acc = 0.0;
#pragma unroll
  for (int j = 0; j < 8; ++j) acc += p[j];
  return wg_sum(acc);
}

// q_{k+1} = w / ||w|| into the basis and into the fixed HxV input vector v
// (zero for a zero w: the iteration then ends on the breakdown test); every
// workgroup sums the np norm partials itself (block_sum, fixed order).
// Workgroup 0 also stores the numbers the host needs after iteration k
// (<q_k, H q_k> in its Gram-Schmidt parts c1[k] + c2[k] + c3[k], and ||w||^2)
// in out[4k ..].
__global__ __launch_bounds__(256) void k_normalize_next(const int* __restrict__ kdev, int kofs,
                                                        const double* __restrict__ npart, int np,
                                                        const double* __restrict__ w, double* __restrict__ Q,
                                                        int64_t n, double* __restrict__ v,
                                                        const double* __restrict__ c1,
                                                        const double* __restrict__ c2,
                                                        const double* __restrict__ c3, const int* __restrict__ gate,
                                                        double* __restrict__ out) {
  const int k = *kdev + kofs;
  const double nv = block_sum(npart, np);
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    out[4 * k] = c1[k];
    out[4 * k + 1] = c2[k];
    out[4 * k + 2] = nv;
    out[4 * k + 3] = (c3 && gate[k]) ? c3[k] : 0.0;  // the full pass of a partially re-orthogonalised iteration
  }
  const double f = nv > 0.0 ? 1.0 / sqrt(nv) : 0.0;
  double* __restrict__ q = Q + (int64_t)(k + 1) * n;
  int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (; e < n; e += stride) {
    const double y = f * w[e];
    q[e] = y;
    v[e] = y;
  }
}

// ---- partial re-orthogonalisation (FP64 operators) --------------------------
// Simon's recurrence (H. D. Simon, Math. Comp. 42 (1984) 115-142) estimates
// omega_{k+1,j} = <q_{k+1}, q_j> from the Lanczos coefficients alone: with
// H q_j = beta_{j-1} q_{j-1} + alpha_j q_j + beta_j q_{j+1} + f_j,
//   beta_k omega_{k+1,j} = beta_j omega_{k,j+1} + (alpha_j - alpha_k) omega_{k,j}
//                          + beta_{j-1} omega_{k,j-1} - beta_{k-1} omega_{k-1,j} + noise,
// the noise bounding the rounding terms <f_j, q_k> - <q_j, f_k> by
// eps1 (beta_k + beta_j) with the sign that makes |omega| larger.  While
// every |omega_{k+1,j}| stays below sqrt(eps) the basis is semi-orthogonal,
// which keeps the Ritz values accurate to working precision; an iteration
// whose estimate crosses it orthogonalises w against the whole basis (and the
// estimate is reset to eps1).  Rows k-1, k, k+1 of omega live in three slots.
constexpr double kProThreshold = 1.4901161193847656e-08;  // sqrt(2^-52)

__device__ inline double* omega_row(double* om, int row, int ld) { return om + (int64_t)(row % 3) * ld; }

// Every workgroup: beta_k from the norm partials of the local passes, the new
// omega row and the decision (the same in every workgroup); workgroup (0,0)
// stores the row (reset to eps1 when the full pass runs) and gate[k].  With
// the full pass due, workgroup (b, y) then forms the basis dot partials of
// rows 4y .. 4y+3 <= k as k_multidot_rows does.
__global__ __launch_bounds__(256) void k_pro_decide_dots(const int* __restrict__ kdev, int kofs,
                                                         const double* __restrict__ Q, int64_t n,
                                                         const double* __restrict__ w,
                                                         const double* __restrict__ ew, int64_t chunk,
                                                         double* __restrict__ partial, int nb,
                                                         const double* __restrict__ npart, int np,
                                                         const double* __restrict__ c1,
                                                         const double* __restrict__ c2,
                                                         const double* __restrict__ out, double* __restrict__ om,
                                                         int ld, double eps1, int* __restrict__ gate) {
  __shared__ double red[4];
  const int k = *kdev + kofs;
  const double bk = sqrt(block_sum(npart, np));
  const double ak = c1[k] + c2[k];
  const double bkm1 = k > 0 ? sqrt(out[4 * (k - 1) + 2]) : 0.0;
  const double* __restrict__ wk = omega_row(om, k, ld);
  const double* __restrict__ wkm1 = omega_row(om, k + 2, ld);  // row k-1
  double* __restrict__ wnew = omega_row(om, k + 1, ld);
  // omega_{k+1,j}, j < k: thread t takes j = t, t + 256 (k < kFuseMaxRows)
  double mx = 0.0, val[2] = {0.0, 0.0};
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const int j = (int)threadIdx.x + 256 * h;
    if (j >= k || !(bk > 0.0)) continue;
    const double aj = out[4 * j] + out[4 * j + 1] + out[4 * j + 3];
    const double bj = sqrt(out[4 * j + 2]);
    const double bjm1 = j > 0 ? sqrt(out[4 * (j - 1) + 2]) : 0.0;
    double num = bj * wk[j + 1] + (aj - ak) * wk[j] + (j > 0 ? bjm1 * wk[j - 1] : 0.0) - bkm1 * wkm1[j];
    num += (num >= 0.0 ? 1.0 : -1.0) * eps1 * (bk + bj);
    val[h] = num / bk;
    mx = fmax(mx, fabs(val[h]));
  }
  for (int m = 32; m > 0; m >>= 1) mx = fmax(mx, __shfl_xor(mx, m));
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = mx;
  __syncthreads();
  const bool need = fmax(fmax(red[0], red[1]), fmax(red[2], red[3])) > kProThreshold;
  if (blockIdx.x == 0 && blockIdx.y == 0) {
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int j = (int)threadIdx.x + 256 * h;
      if (j < k) wnew[j] = need ? eps1 : val[h];
    }
    if (threadIdx.x == 0) {
      wnew[k] = eps1;  // locally orthogonal
      wnew[k + 1] = 1.0;
      gate[k] = need ? 1 : 0;
    }
  }
  if (!need) return;  // workgroup-uniform
  const int i0 = 4 * blockIdx.y;
  if (i0 > k) return;
  __shared__ double ws[4][4];
  const int nr = min(4, k + 1 - i0);
  const int b = blockIdx.x;
  const int64_t e0 = (int64_t)b * chunk, e1 = min(n, e0 + chunk);
  const double* __restrict__ q0 = Q + (int64_t)i0 * n;
  const double* __restrict__ q1 = Q + (int64_t)min(i0 + 1, k) * n;
  const double* __restrict__ q2 = Q + (int64_t)min(i0 + 2, k) * n;
  const double* __restrict__ q3 = Q + (int64_t)min(i0 + 3, k) * n;
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll 4
  for (int64_t e = e0 + threadIdx.x; e < e1; e += 256) {
    const double we = w[e] * ew[e];
    acc[0] += q0[e] * we;
    acc[1] += q1[e] * we;
    acc[2] += q2[e] * we;
    acc[3] += q3[e] * we;
  }
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    double x = acc[r];
    for (int m = 32; m > 0; m >>= 1) x += __shfl_xor(x, m);
    if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6][r] = x;
  }
  __syncthreads();
  const int t = threadIdx.x;
  if (t < nr) partial[(int64_t)(i0 + t) * nb + b] = (ws[0][t] + ws[1][t]) + (ws[2][t] + ws[3][t]);
}

// omega_{0,0} = 1 before the first iteration of a solve
__global__ void k_pro_init(double* __restrict__ om) {
  if (threadIdx.x == 0 && blockIdx.x == 0) om[0] = 1.0;
}

// *p += d, after every kernel of a batch has read *p
__global__ void k_add_int(int* __restrict__ p, int d) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *p += d;
}

__global__ void k_set_int(int* __restrict__ p, int value) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *p = value;
}

__global__ void k_copy(const double* __restrict__ x, double* __restrict__ y, int64_t n) {
  int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (; e < n; e += stride) y[e] = x[e];
}

}  // namespace

// ---- tridiagonal lowest eigenpair (host) -------------------------------------
// Number of eigenvalues of the symmetric tridiagonal (a, b) below x (Sturm count).
static int sturm_count(const std::vector<double>& a, const std::vector<double>& b, int k, double x) {
  int cnt = 0;
  double d = 1.0;
  for (int i = 0; i < k; ++i) {
    const double bb = i > 0 ? b[i - 1] * b[i - 1] : 0.0;
    d = (a[i] - x) - (i > 0 ? bb / d : 0.0);
    if (d == 0.0) d = -1e-300;
    if (d < 0.0) ++cnt;
  }
  return cnt;
}

void tridiag_lowest(const std::vector<double>& a, const std::vector<double>& b, int k, double& E,
                    std::vector<double>& y) {
  // Gershgorin bounds
  double lo = 1e300, hi = -1e300;
  for (int i = 0; i < k; ++i) {
    double r = (i > 0 ? std::fabs(b[i - 1]) : 0.0) + (i + 1 < k ? std::fabs(b[i]) : 0.0);
    lo = std::min(lo, a[i] - r);
    hi = std::max(hi, a[i] + r);
  }
  for (int it = 0; it < 200; ++it) {  // bisection for the lowest eigenvalue
    const double mid = 0.5 * (lo + hi);
    if (mid <= lo || mid >= hi) break;
    if (sturm_count(a, b, k, mid) >= 1)
      hi = mid;
    else
      lo = mid;
  }
  E = 0.5 * (lo + hi);
  // inverse iteration with a slightly shifted value (tridiagonal LU, partial pivoting free for small k)
  y.assign(k, 1.0 / std::sqrt((double)k));
  const double scale = std::max(1.0, std::fabs(E));
  const double shift = E - 1e-13 * scale;
  std::vector<double> dl(k), dd(k), du(k), r(k);
  for (int pass = 0; pass < 3; ++pass) {
    for (int i = 0; i < k; ++i) {
      dd[i] = a[i] - shift;
      du[i] = i + 1 < k ? b[i] : 0.0;
      dl[i] = i > 0 ? b[i - 1] : 0.0;
      r[i] = y[i];
    }
    // Gaussian elimination with row swaps on the tridiagonal band (stored as dense band of width 3)
    std::vector<double> u2(k, 0.0);
    for (int i = 0; i + 1 < k; ++i) {
      if (std::fabs(dl[i + 1]) > std::fabs(dd[i])) {  // swap rows i and i+1
        std::swap(dd[i], dl[i + 1]);
        std::swap(du[i], dd[i + 1]);
        if (i + 2 < k) std::swap(u2[i], du[i + 1]);
        std::swap(r[i], r[i + 1]);
      }
      if (dd[i] == 0.0) dd[i] = 1e-300;
      const double f = dl[i + 1] / dd[i];
      dd[i + 1] -= f * du[i];
      if (i + 2 < k) du[i + 1] -= f * u2[i];
      r[i + 1] -= f * r[i];
    }
    if (dd[k - 1] == 0.0) dd[k - 1] = 1e-300;
    r[k - 1] /= dd[k - 1];
    if (k >= 2) r[k - 2] = (r[k - 2] - du[k - 2] * r[k - 1]) / dd[k - 2];
    for (int i = k - 3; i >= 0; --i) r[i] = (r[i] - du[i] * r[i + 1] - u2[i] * r[i + 2]) / dd[i];
    double nrm = 0.0;
    for (int i = 0; i < k; ++i) nrm += r[i] * r[i];
    nrm = std::sqrt(nrm);
    for (int i = 0; i < k; ++i) y[i] = r[i] / nrm;
  }
  if (y[0] < 0)
    for (auto& x : y) x = -x;
}

// ---- device vector helpers ---------------------------------------------------
void dev_axpby(double a, const double* x, double b, double* y, int64_t n, hipStream_t s) {
  hipLaunchKernelGGL(k_axpby, dim3(vgrid(n)), dim3(256), 0, s, a, x, b, y, n);
  VMPS_HIP_CHECK(hipGetLastError());
}

void dev_fill_uniform(double* v, int64_t n, uint64_t seed, hipStream_t s) {
  hipLaunchKernelGGL(k_fill_uniform, dim3(vgrid(n)), dim3(256), 0, s, v, n, seed);
  VMPS_HIP_CHECK(hipGetLastError());
}

// ---- Lanczos ---------------------------------------------------------------
// One iteration (HxV of q_k, two classical Gram-Schmidt passes against
// q_0 .. q_k, ||w||^2, q_{k+1} = w / ||w||) is a fixed sequence of kernels whose
// only varying input, k, is read from device memory.  With `graph` the sequence
// is captured once per plan configuration as a HIP graph and replayed each
// iteration; without it the same kernels are launched one by one.  Either way
// the host waits once per iteration, for three numbers.
// Below 64k elements a pass over the whole basis costs about what the local
// passes and the decision cost (every kernel is at its launch floor), so
// small vectors keep full re-orthogonalisation.
constexpr int64_t kProMinElements = 65536;

bool Krylov::partial_reortho(const DevPlan& D) const {
  return !full_reortho && nvec >= kProMinElements && ngrp <= kFuseMaxGroups && cap < kFuseMaxRows &&
         D.prec == VMPS_PREC_FP64;
}

void Krylov::enqueue_iteration(DevPlan& D, hipStream_t s, int kofs) {
  const int64_t n = D.H.vsize;
  const unsigned gy = (unsigned)((cap + 3) / 4);
  const unsigned gs = vgrid(n);
  const bool fused = ngrp <= kFuseMaxGroups && cap < kFuseMaxRows;
  double* c1 = c.p;
  double* c2 = c.p + cap;
  double* c3 = c.p + 2 * cap;
  auto pass = [&](double* c_out, double* norm_part, int local) {
    hipLaunchKernelGGL(k_multidot_rows, dim3(ngrp, local ? 1u : gy), dim3(256), 0, s, kdev.p, kofs, Q.p, n, w.p,
                       ew.p, chunk, gpart.p, ngrp, local);
    if (fused) {
      hipLaunchKernelGGL(k_sub_combination_fused, dim3(gs), dim3(256), 0, s, Q.p, n, gpart.p, ngrp, kdev.p, kofs,
                         c_out, w.p, ew.p, norm_part, local, (const int*)nullptr, 0);
    } else {
      hipLaunchKernelGGL(k_multidot_rows_final, dim3(cap), dim3(256), 0, s, kdev.p, kofs, gpart.p, ngrp, c_out);
      hipLaunchKernelGGL(k_sub_combination_dk, dim3(gs), dim3(256), 0, s, Q.p, n, c_out, kdev.p, kofs, w.p, ew.p,
                         norm_part);
    }
  };
  D.hxv(v.p, w.p, s);
  if (partial_reortho(D)) {
    // FP64 operator: classical Gram-Schmidt twice against q_{k-1} and q_k
    // (the first pass's subtraction and the second pass's dots in one
    // kernel), then the whole basis only where the omega estimate asks for it
    hipLaunchKernelGGL(k_multidot_rows, dim3(ngrp, 1u), dim3(256), 0, s, kdev.p, kofs, Q.p, n, w.p, ew.p, chunk,
                       gpart.p, ngrp, 1);
    hipLaunchKernelGGL(k_sub_local_dots, dim3(ngrp), dim3(256), 0, s, Q.p, n, gpart.p, ngrp, kdev.p, kofs, c1, w.p,
                       ew.p, chunk);
    hipLaunchKernelGGL(k_sub_combination_fused, dim3(gs), dim3(256), 0, s, Q.p, n, gpart.p, ngrp, kdev.p, kofs, c2,
                       w.p, ew.p, npart.p, 1, (const int*)nullptr, 2);
    const double eps1 = 1.1102230246251565e-16 * std::sqrt((double)n);
    hipLaunchKernelGGL(k_pro_decide_dots, dim3(ngrp, gy), dim3(256), 0, s, kdev.p, kofs, Q.p, n, w.p, ew.p, chunk,
                       gpart.p, ngrp, npart.p, (int)gs, c1, c2, out3.p, om.p, cap + 2, eps1, gate.p);
    hipLaunchKernelGGL(k_sub_combination_fused, dim3(gs), dim3(256), 0, s, Q.p, n, gpart.p, ngrp, kdev.p, kofs, c3,
                       w.p, ew.p, npart.p, 0, (const int*)gate.p, 0);
    hipLaunchKernelGGL(k_normalize_next, dim3(gs), dim3(256), 0, s, kdev.p, kofs, npart.p, (int)gs, w.p, Q.p, n,
                       v.p, c1, c2, c3, (const int*)gate.p, out3.p);
  } else {
    pass(c1, nullptr, 0);
    pass(c2, npart.p, 0);
    hipLaunchKernelGGL(k_normalize_next, dim3(gs), dim3(256), 0, s, kdev.p, kofs, npart.p, (int)gs, w.p, Q.p, n,
                       v.p, c1, c2, (const double*)nullptr, (const int*)nullptr, out3.p);
  }
  VMPS_HIP_CHECK(hipGetLastError());
}

void Krylov::release_graph() {
  for (int i = 0; i < 3; ++i) {
    if (exec[i]) (void)hipGraphExecDestroy(exec[i]);
    if (exec2[i]) (void)hipGraphExecDestroy(exec2[i]);
    if (graph[i]) (void)hipGraphDestroy(graph[i]);
    exec[i] = nullptr;
    exec2[i] = nullptr;
    graph[i] = nullptr;
    graph_gen[i] = -1;
  }
}

hipGraphExec_t Krylov::ensure_graph(DevPlan& D, int copy) {
  const int p = (int)D.prec;
  if (p < 0 || p > 2) throw VmpsError(VMPS_ERR_ARG, "unsupported precision");
  if (exec[p] && graph_gen[p] == D.gen) {
    if (copy == 0) return exec[p];
    if (!exec2[p]) VMPS_HIP_CHECK(hipGraphInstantiate(&exec2[p], graph[p], nullptr, nullptr, 0));
    return exec2[p];
  }
  if (exec[p]) (void)hipGraphExecDestroy(exec[p]);
  if (exec2[p]) (void)hipGraphExecDestroy(exec2[p]);
  if (graph[p]) (void)hipGraphDestroy(graph[p]);
  exec[p] = nullptr;
  exec2[p] = nullptr;
  graph[p] = nullptr;
  hipStream_t s = D.stream;
  VMPS_HIP_CHECK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
  try {
    for (int i = 0; i < kIterBatch; ++i) enqueue_iteration(D, s, i);
    hipLaunchKernelGGL(k_add_int, dim3(1), dim3(64), 0, s, kdev.p, kIterBatch);
  } catch (...) {
    hipGraph_t g = nullptr;
    (void)hipStreamEndCapture(s, &g);
    if (g) (void)hipGraphDestroy(g);
    throw;
  }
  VMPS_HIP_CHECK(hipStreamEndCapture(s, &graph[p]));
  VMPS_HIP_CHECK(hipGraphInstantiate(&exec[p], graph[p], nullptr, nullptr, 0));
  graph_gen[p] = D.gen;
  if (copy == 0) return exec[p];
  VMPS_HIP_CHECK(hipGraphInstantiate(&exec2[p], graph[p], nullptr, nullptr, 0));
  return exec2[p];
}

LanczosResult lanczos_solve(DevPlan& D, double* d_v, const LanczosOpts& o, vmps_precision prec) {
  hipStream_t s = D.stream;
  const int64_t n = D.H.vsize;
  LanczosResult res;
  if (n == 0) return res;
  const int kmax = std::max(2, (int)std::min<int64_t>(o.max_iter, n));
  Krylov& K = D.krylov;
  const vmps_precision saved = D.prec;
  D.prec = prec;
  D.ensure_precision_buffers();
  K.ensure(D, kmax);
  // A captured batch runs kIterBatch iterations, so a solve runs (kIterBatch
  // - 1) / 2 HxV past convergence on average; where one HxV costs more than
  // the launches and the per-iteration readback the graph saves (~0.1 ms per
  // iteration against ~1.5 ms for an HxV of 5e10 FLOP), the iterations are
  // launched one at a time instead.
  constexpr double kGraphMaxFlops = 5e10;
  hipGraphExec_t iteration =
      o.graph && D.H.stats.flops_alg < kGraphMaxFlops ? K.ensure_graph(D, 0) : nullptr;

  // q0 = v / ||v|| into the basis and into the HxV input; k = 0.  A zero
  // (or non-finite) start vector has no direction to start from.
  D.dot_async(d_v, d_v, K.scal.p, s);
  VMPS_HIP_CHECK(hipMemcpyAsync(K.host3, K.scal.p, sizeof(double), hipMemcpyDeviceToHost, s));
  VMPS_HIP_CHECK(hipStreamSynchronize(s));
  if (!(K.host3[0] > 0.0) || !std::isfinite(K.host3[0])) {
    D.prec = saved;
    throw VmpsError(K.host3[0] == 0.0 ? VMPS_ERR_ARG : VMPS_ERR_NONFINITE,
                    K.host3[0] == 0.0 ? "the start vector is zero" : "the start vector is not finite");
  }
  hipLaunchKernelGGL(k_normalize_into, dim3(vgrid(n)), dim3(256), 0, s, K.scal.p, d_v, K.Q.p, n);
  hipLaunchKernelGGL(k_copy, dim3(vgrid(n)), dim3(256), 0, s, K.Q.p, K.v.p, n);
  hipLaunchKernelGGL(k_set_int, dim3(1), dim3(64), 0, s, K.kdev.p, 0);
  const bool pro = K.partial_reortho(D);
  if (pro) hipLaunchKernelGGL(k_pro_init, dim3(1), dim3(64), 0, s, K.om.p);
  VMPS_HIP_CHECK(hipGetLastError());

  std::vector<double> alpha, beta, y;
  double E = 0.0, Eprev = 0.0, rest_last = 0.0;
  int k = 0;
  bool converged = false;
  // The graph runs kIterBatch iterations per launch and the host reads their
  // scalars with one synchronisation; it then tests them one by one, exactly
  // as after single iterations.  Iterations a batch runs past convergence (or
  // past kmax) only extend the basis beyond the part the Ritz vector uses.
  // On cheap operators, while the last batch read left the residual estimate
  // or the eigenvalue change at least 100x above its tolerance (so that
  // convergence within the next batch is unlikely), one more batch is
  // queued before the host reads the oldest, and the device does not wait for
  // the readback and the tridiagonal solve.  A batch queued past convergence
  // is the price (kIterBatch HxV, at most ~0.2 ms under kAheadMaxFlops); short
  // warm-started solves, which converge within a batch or two, never pay it.
  // Consecutive batches alternate between two executable instances of the
  // captured graph, so an instance is launched again only after its previous
  // launch has completed (the host waited for it).
  constexpr int B = Krylov::kIterBatch;
  constexpr double kAheadMaxFlops = 1e9;
  const bool ahead = iteration && D.H.stats.flops_alg <= kAheadMaxFlops;
  hipGraphExec_t instance[2] = {iteration, ahead ? K.ensure_graph(D, 1) : iteration};
  bool far = false;  // the last batch read left both tests far from their tolerances
  int kq = 0;        // first iteration of the next batch to queue
  int inflight = 0;  // batches queued whose scalars the host has not read
  int head = 0;      // readback slot of the oldest of them
  auto queue_batch = [&]() {
    const int slot = (head + inflight) & 1;
    VMPS_HIP_CHECK(hipGraphLaunch(instance[(kq / B) & 1], s));
    VMPS_HIP_CHECK(hipMemcpyAsync(K.host3 + 4 * B * slot, K.out3.p + 4 * (size_t)kq, 4 * B * sizeof(double),
                                  hipMemcpyDeviceToHost, s));
    VMPS_HIP_CHECK(hipEventRecord(K.done[slot], s));
    kq += B;
    ++inflight;
  };
  while (k < kmax && !converged) {
    int nb = 1;
    const double* hs = K.host3;
    if (iteration) {
      if (inflight == 0) queue_batch();
      if (ahead && far && inflight == 1 && kq < kmax) queue_batch();
      VMPS_HIP_CHECK(hipEventSynchronize(K.done[head]));
      hs = K.host3 + 4 * B * head;
      head ^= 1;
      --inflight;
      nb = B;
    } else {
      K.enqueue_iteration(D, s, 0);
      hipLaunchKernelGGL(k_add_int, dim3(1), dim3(64), 0, s, K.kdev.p, 1);
      VMPS_HIP_CHECK(hipMemcpyAsync(K.host3, K.out3.p + 4 * (size_t)k, 4 * sizeof(double), hipMemcpyDeviceToHost, s));
      VMPS_HIP_CHECK(hipStreamSynchronize(s));
    }
    res.hxv_calls += nb;
    for (int j = 0; j < nb && k < kmax; ++j) {
      const double a_k = hs[4 * j] + hs[4 * j + 1] + hs[4 * j + 3];
      const double b_k = std::sqrt(std::max(0.0, hs[4 * j + 2]));
      alpha.push_back(a_k);
      beta.push_back(b_k);
      Eprev = E;
      tridiag_lowest(alpha, beta, k + 1, E, y);
      // residual estimate of the Ritz pair: beta_k |y_k|
      const double rest = b_k * std::fabs(y[k]);
      rest_last = rest;
      res.iterations = k + 1;
      ++k;
      // convergence after min_iter iterations; a breakdown (the Krylov space is
      // invariant, the Ritz pair exact) ends the iteration at any point, since
      // the next basis vector would be the normalised zero vector
      const bool breakdown = b_k < 1e-14 * std::max(1.0, std::fabs(a_k));
      if (breakdown || (k >= o.min_iter && k > 1 && std::fabs(E - Eprev) < o.tol_eigval && rest < o.tol_state)) {
        converged = true;
        break;
      }
      far = k > 1 && (std::fabs(E - Eprev) > 100.0 * o.tol_eigval || rest > 100.0 * o.tol_state);
    }
  }
  // a batch queued past convergence runs before the work below (stream order)
  // and only extends the basis; its HxV applications are counted as done
  res.hxv_calls += B * inflight;
  k = res.iterations;
  // Ritz vector psi = Q y (k basis vectors), normalised, then the true residual
  VMPS_HIP_CHECK(hipMemcpyAsync(K.c.p, y.data(), sizeof(double) * k, hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(k_combination, dim3(vgrid(n)), dim3(256), 0, s, K.Q.p, n, K.c.p, k, K.w.p);
  D.dot_async(K.w.p, K.w.p, K.scal.p, s);
  hipLaunchKernelGGL(k_normalize_into, dim3(vgrid(n)), dim3(256), 0, s, K.scal.p, K.w.p, d_v, n);
  VMPS_HIP_CHECK(hipGetLastError());
  res.ritz = E;
  res.converged = converged;
  double rq = E;
  if (o.final_rq || pro) {
    // Rayleigh quotient and residual with the requested precision's operator
    D.hxv(d_v, K.w.p, s);
    ++res.hxv_calls;
    D.dot_async(d_v, K.w.p, K.scal.p, s);
    VMPS_HIP_CHECK(hipMemcpyAsync(K.host3, K.scal.p, sizeof(double), hipMemcpyDeviceToHost, s));
    VMPS_HIP_CHECK(hipStreamSynchronize(s));
    rq = K.host3[0];
    dev_axpby(-rq, d_v, 1.0, K.w.p, n, s);
    D.dot_async(K.w.p, K.w.p, K.scal.p, s);
    VMPS_HIP_CHECK(hipMemcpyAsync(K.host3, K.scal.p, sizeof(double), hipMemcpyDeviceToHost, s));
    VMPS_HIP_CHECK(hipStreamSynchronize(s));
    res.energy = rq;
    res.residual = std::sqrt(std::max(0.0, K.host3[0]));
  } else {
    res.energy = E;
    res.residual = rest_last;
  }
  D.prec = saved;
  if (pro) {
    // With a semi-orthogonal basis the Ritz value equals the Rayleigh
    // quotient of the Ritz vector to O(eps ||T||).  A larger gap means the
    // basis lost orthogonality that the omega estimate did not see (it
    // assumes a symmetric operator): solve again from the same start vector
    // with the whole basis re-orthogonalised at every iteration.
    double tnorm = 0.0;
    for (int j = 0; j < k; ++j)
      tnorm = std::max(tnorm, std::fabs(alpha[j]) + beta[j] + (j > 0 ? beta[j - 1] : 0.0));
    if (!(std::fabs(E - rq) <= 1e-8 * std::max(tnorm, 1e-300))) {
      hipLaunchKernelGGL(k_copy, dim3(vgrid(n)), dim3(256), 0, s, K.Q.p, d_v, n);  // q_0
      VMPS_HIP_CHECK(hipGetLastError());
      K.full_reortho = true;
      K.release_graph();
      LanczosResult again;
      try {
        again = lanczos_solve(D, d_v, o, prec);
      } catch (...) {
        K.full_reortho = false;
        K.release_graph();
        throw;
      }
      K.full_reortho = false;
      K.release_graph();
      again.hxv_calls += res.hxv_calls;
      return again;
    }
  }
  return res;
}

// ---- mixed-precision refinement ---------------------------------------------
// Iterative refinement of an eigenpair in the manner of Jacobi-Davidson: the
// residual r = H psi - E psi is formed with the FP64 operator H; the
// correction t, orthogonal to psi, solves the projected equation
//   (1 - psi psi^T)(H' - E)(1 - psi psi^T) t = -r
// by conjugate gradients with the mixed-precision operator H' (the projected
// operator is positive definite on the complement of psi while E is below
// the second eigenvalue); then the FP64 operator's Rayleigh-Ritz pair on
// span{psi, t} replaces psi.  An outer step costs one FP64 HxV (H psi is
// carried along as w), so the FP64 work no longer grows with the Krylov
// dimension: the error shrinks per step by about max(CG tolerance,
// eps(H') ||H|| / gap).  Inner products carry the coeff_dot weights.
RefineResult refine_eigenpair(DevPlan& D, double* d_v, const RefineOpts& o, vmps_precision inner) {
  hipStream_t s = D.stream;
  const int64_t n = D.H.vsize;
  RefineResult res;
  if (n == 0) return res;
  Krylov& K = D.krylov;
  K.ensure(D, std::max(2, K.cap - Krylov::kIterBatch));  // at least six vectors of workspace in Q
  double* psi = d_v;
  double* w = K.Q.p;   // H psi (FP64)
  double* t = w + n;   // correction
  double* g = t + n;   // CG residual; -r before the CG
  double* pv = g + n;  // CG direction
  double* q = pv + n;  // projected operator on pv; H t afterwards
  const vmps_precision saved = D.prec;
  auto apply = [&](vmps_precision p, const double* x, double* y) {
    D.prec = p;
    D.ensure_precision_buffers();
    D.hxv(x, y, s);
  };
  auto dot = [&](const double* x, const double* y) { return D.dot(x, y, s); };
  auto axpby = [&](double a, const double* x, double b, double* y) { dev_axpby(a, x, b, y, n, s); };
  auto copy = [&](const double* x, double* y) {
    hipLaunchKernelGGL(k_copy, dim3(vgrid(n)), dim3(256), 0, s, x, y, n);
    VMPS_HIP_CHECK(hipGetLastError());
  };
  try {
    apply(VMPS_PREC_FP64, psi, w);
    ++res.hxv_fp64;
    double E = dot(psi, w);
    copy(w, g);
    axpby(E, psi, -1.0, g);  // g = E psi - H psi = -r
    double rho = std::sqrt(std::max(0.0, dot(g, g)));
    res.energy0 = E;
    res.residual0 = rho;
    for (;;) {
      if (rho <= o.tol_state && rho * rho <= o.tol_eigval) {
        res.converged = true;
        break;
      }
      if (res.outer >= o.max_outer || !std::isfinite(rho)) break;
      ++res.outer;
      // CG on the projected correction equation, from t = 0 (residual g = -r);
      // it stops once the correction reaches half the residual the step needs,
      // down to 1e-4 of the residual (split-BF16x3 Lanczos vectors, residual
      // ~3e-4, then reach a 1e-7 residual in one step instead of two)
      const double cg_tol = std::min(0.1, std::max(o.cg_tol_min, 0.5 * o.tol_state / rho));
      VMPS_HIP_CHECK(hipMemsetAsync(t, 0, sizeof(double) * n, s));
      copy(g, pv);
      double gg = rho * rho;
      for (int it = 0; it < o.cg_max; ++it) {
        apply(inner, pv, q);
        ++res.cg_iterations;
        axpby(-E, pv, 1.0, q);
        const double c = dot(psi, q);
        axpby(-c, psi, 1.0, q);
        const double pq = dot(pv, q);
        if (!(pq > 0.0)) break;  // not positive definite here: keep the correction so far
        const double alpha = gg / pq;
        axpby(alpha, pv, 1.0, t);
        axpby(-alpha, q, 1.0, g);
        const double gg_new = dot(g, g);
        if (std::sqrt(gg_new) <= cg_tol * rho) break;
        axpby(1.0, g, gg_new / gg, pv);  // pv = g + beta pv
        gg = gg_new;
      }
      // t -> unit vector orthogonal to psi
      const double ct = dot(psi, t);
      axpby(-ct, psi, 1.0, t);
      const double nt = std::sqrt(std::max(0.0, dot(t, t)));
      if (!(nt > 1e-300) || !std::isfinite(nt)) break;
      axpby(0.0, t, 1.0 / nt, t);
      // Rayleigh-Ritz of the FP64 operator on span{psi, t}
      apply(VMPS_PREC_FP64, t, q);
      ++res.hxv_fp64;
      const double h11 = E, h12 = dot(psi, q), h22 = dot(t, q);
      const double mean = 0.5 * (h11 + h22), half = 0.5 * (h11 - h22);
      const double lam = mean - std::hypot(half, h12);
      double c1 = h12, c2 = lam - h11;  // (H - lam) c = 0, two equivalent rows: take the longer
      if (std::hypot(lam - h22, h12) > std::hypot(c1, c2)) {
        c1 = lam - h22;
        c2 = h12;
      }
      const double cn = std::hypot(c1, c2);
      if (!(cn > 0.0)) break;
      c1 /= cn;
      c2 /= cn;
      if (c1 < 0.0) {
        c1 = -c1;
        c2 = -c2;
      }
      axpby(c2, t, c1, psi);
      axpby(c2, q, c1, w);
      const double nrm = std::sqrt(dot(psi, psi));
      axpby(0.0, psi, 1.0 / nrm, psi);
      axpby(0.0, w, 1.0 / nrm, w);
      E = dot(psi, w);
      copy(w, g);
      axpby(E, psi, -1.0, g);
      rho = std::sqrt(std::max(0.0, dot(g, g)));
    }
    res.energy = E;
    res.residual = rho;
  } catch (...) {
    D.prec = saved;
    throw;
  }
  D.prec = saved;
  return res;
}

// The Krylov basis is the largest buffer of a solve (kmax + 5 vectors: 37 GB
// for 300 iterations on the C60 M = 4096 pivots), and a DMRG sweep builds a
// new plan, hence a new basis, at every step; allocating it took ~25 ms.  One
// released basis per device is kept (until the process ends) for the next
// plan that fits in it.  A basis is released only after its plan's stream has
// finished (the plan synchronises its stream before its members go).
namespace {
std::mutex g_basis_mutex;
std::map<int, std::pair<double*, size_t>> g_basis;  // device -> (pointer, doubles)

void basis_acquire(DevBuf<double>& Q, size_t count, int device) {
  Q.release();
  {
    std::lock_guard<std::mutex> lock(g_basis_mutex);
    auto it = g_basis.find(device);
    if (it != g_basis.end() && it->second.second >= count) {
      Q.p = it->second.first;
      Q.n = it->second.second;
      g_basis.erase(it);
      return;
    }
  }
  Q.alloc(count);
}

void basis_release(DevBuf<double>& Q, int device) {
  if (!Q.p) return;
  double* give = Q.p;
  size_t n = Q.n;
  Q.p = nullptr;
  Q.n = 0;
  double* drop = nullptr;
  {
    std::lock_guard<std::mutex> lock(g_basis_mutex);
    auto& slot = g_basis[device];
    if (slot.first && slot.second >= n) {
      drop = give;  // the kept one is at least as large
    } else {
      drop = slot.first;
      slot = {give, n};
    }
  }
  if (drop) (void)hipFree(drop);
}
}  // namespace

void Krylov::ensure(DevPlan& D, int kmax) {
  const int64_t n = D.H.vsize;
  const int need = kmax + kIterBatch;
  if (cap >= need && nvec == n) return;
  release_graph();
  VMPS_HIP_CHECK(hipStreamSynchronize(D.stream));  // a basis given back has no work pending
  basis_release(Q, device);
  device = D.device;
  basis_acquire(Q, (size_t)(need + 1) * n, device);
  w.alloc(n);
  v.alloc(n);
  c.alloc(3 * (size_t)need + 3);
  om.alloc(3 * ((size_t)need + 2));
  gate.alloc((size_t)need + 1);
  scal.alloc(8);
  out3.alloc(4 * (size_t)need);
  kdev.alloc(1);
  if (!host3) VMPS_HIP_CHECK(hipHostMalloc((void**)&host3, 8 * kIterBatch * sizeof(double), hipHostMallocDefault));
  for (auto& e : done)
    if (!e) VMPS_HIP_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  // per-element coeff_dot weights and element chunks of the row dot products
  std::vector<double> hw((size_t)n, 1.0);
  for (size_t sg = 0; sg < D.H.seg_off.size(); ++sg)
    for (int32_t e = 0; e < D.H.seg_len[sg]; ++e) hw[(size_t)(D.H.seg_off[sg] + e)] = D.H.seg_w[sg];
  ew.upload(hw, D.stream);
  // chunks of 1024 elements (4 per thread) for long vectors; vectors of up to
  // 64 x 8192 elements are cut into at most kFuseMaxGroups chunks instead, so
  // an iteration takes the fused path (two kernels fewer, the launch-bound
  // regime of small pivots)
  ngrp = (int)std::max<int64_t>(1, std::min<int64_t>(1024, (n + 1023) / 1024));
  if (n <= (int64_t)kFuseMaxGroups * 8192) ngrp = std::min<int>(ngrp, kFuseMaxGroups);
  chunk = (n + ngrp - 1) / ngrp;
  npart.alloc(vgrid(n));
  gpart.alloc((size_t)std::max(1, ngrp) * (need + 3));  // rows 0 .. k and k + 1, k + 2 (k_sub_local_dots)
  VMPS_HIP_CHECK(hipStreamSynchronize(D.stream));
  cap = need;
  nvec = n;
}

Krylov::~Krylov() {
  release_graph();
  basis_release(Q, device);
  if (host3) (void)hipHostFree(host3);
  host3 = nullptr;
  for (auto& e : done) {
    if (e) (void)hipEventDestroy(e);
    e = nullptr;
  }
}

}  // namespace vmps
