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
#include <cstring>
#include <vector>

#include "engine.hpp"
#include "lanczos.hpp"

namespace vmps {

namespace {

// c[i] = <Q_i, w>_weighted for i < k; one workgroup per (segment group, i).
// Partial sums go to partial[i * nb + b]; reduced by k_multidot_final.
__global__ __launch_bounds__(256) void k_multidot_partial(const int64_t* __restrict__ seg_off,
                                                          const int32_t* __restrict__ seg_len,
                                                          const double* __restrict__ seg_w, int nseg, int spb,
                                                          const double* __restrict__ Q, int64_t n,
                                                          const double* __restrict__ w, double* __restrict__ partial,
                                                          int nb) {
  __shared__ double red[256];
  const int i = blockIdx.y;
  const double* __restrict__ q = Q + (int64_t)i * n;
  double acc = 0.0;
  const int s0 = blockIdx.x * spb;
  const int s1 = min(nseg, s0 + spb);
  for (int s = s0; s < s1; ++s) {
    const int64_t off = seg_off[s];
    const int len = seg_len[s];
    double part = 0.0;
    for (int e = threadIdx.x; e < len; e += 256) part += q[off + e] * w[off + e];
    acc += seg_w[s] * part;
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[(int64_t)i * nb + blockIdx.x] = red[0];
}

__global__ __launch_bounds__(256) void k_multidot_final(const double* __restrict__ partial, int nb,
                                                        double* __restrict__ c) {
  __shared__ double red[256];
  const int i = blockIdx.x;
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

// w -= sum_{i<k} c[i] Q_i
__global__ void k_sub_combination(const double* __restrict__ Q, int64_t n, const double* __restrict__ c, int k,
                                  double* __restrict__ w) {
  int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (; e < n; e += stride) {
    double s = 0.0;
    for (int i = 0; i < k; ++i) s += c[i] * Q[(int64_t)i * n + e];
    w[e] -= s;
  }
}

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
LanczosResult lanczos_solve(DevPlan& D, double* d_v, const LanczosOpts& o, vmps_precision prec) {
  hipStream_t s = D.stream;
  const int64_t n = D.H.vsize;
  LanczosResult res;
  if (n == 0) return res;
  const int kmax = std::max(2, (int)std::min<int64_t>(o.max_iter, n));
  Krylov& K = D.krylov;
  K.ensure(D, kmax);
  const vmps_precision saved = D.prec;
  D.prec = prec;
  D.ensure_precision_buffers();

  const int nseg = (int)D.H.seg_off.size();
  const int spb = 16;
  const int nb = (nseg + spb - 1) / spb;

  auto multidot = [&](int k, const double* w, double* c_dev) {
    hipLaunchKernelGGL(k_multidot_partial, dim3(nb, k), dim3(256), 0, s, D.seg_off.p, D.seg_len.p, D.seg_w.p, nseg,
                       spb, K.Q.p, n, w, K.part.p, nb);
    hipLaunchKernelGGL(k_multidot_final, dim3(k), dim3(256), 0, s, K.part.p, nb, c_dev);
    VMPS_HIP_CHECK(hipGetLastError());
  };

  // q0 = v / ||v||
  D.dot_async(d_v, d_v, K.scal.p, s);
  hipLaunchKernelGGL(k_normalize_into, dim3(vgrid(n)), dim3(256), 0, s, K.scal.p, d_v, K.Q.p, n);
  VMPS_HIP_CHECK(hipGetLastError());

  std::vector<double> alpha, beta, y, c_host(kmax + 1);
  double E = 0.0, Eprev = 0.0;
  int k = 0;
  bool converged = false;
  for (k = 0; k < kmax; ++k) {
    double* qk = K.Q.p + (int64_t)k * n;
    D.hxv(qk, K.w.p, s);
    ++res.hxv_calls;
    // two passes of classical Gram-Schmidt against q_0 .. q_k
    multidot(k + 1, K.w.p, K.c.p);
    hipLaunchKernelGGL(k_sub_combination, dim3(vgrid(n)), dim3(256), 0, s, K.Q.p, n, K.c.p, k + 1, K.w.p);
    multidot(k + 1, K.w.p, K.c.p + kmax);
    hipLaunchKernelGGL(k_sub_combination, dim3(vgrid(n)), dim3(256), 0, s, K.Q.p, n, K.c.p + kmax, k + 1, K.w.p);
    D.dot_async(K.w.p, K.w.p, K.scal.p, s);
    VMPS_HIP_CHECK(hipGetLastError());
    double hc[2];
    VMPS_HIP_CHECK(hipMemcpyAsync(c_host.data(), K.c.p, sizeof(double) * (k + 1), hipMemcpyDeviceToHost, s));
    VMPS_HIP_CHECK(hipMemcpyAsync(hc, K.c.p + kmax + k, sizeof(double), hipMemcpyDeviceToHost, s));
    VMPS_HIP_CHECK(hipMemcpyAsync(hc + 1, K.scal.p, sizeof(double), hipMemcpyDeviceToHost, s));
    VMPS_HIP_CHECK(hipStreamSynchronize(s));
    const double a_k = c_host[k] + hc[0];
    const double b_k = std::sqrt(std::max(0.0, hc[1]));
    alpha.push_back(a_k);
    beta.push_back(b_k);
    Eprev = E;
    tridiag_lowest(alpha, beta, k + 1, E, y);
    // residual estimate of the Ritz pair: beta_k |y_k|
    const double rest = b_k * std::fabs(y[k]);
    res.iterations = k + 1;
    // convergence after min_iter iterations; a breakdown (the Krylov space is
    // invariant, the Ritz pair exact) ends the iteration at any point, since
    // the next basis vector would be the normalised zero vector
    const bool breakdown = b_k < 1e-14 * std::max(1.0, std::fabs(a_k));
    if (breakdown ||
        (k + 1 >= o.min_iter && k > 0 && std::fabs(E - Eprev) < o.tol_eigval && rest < o.tol_state)) {
      converged = true;
      ++k;
      break;
    }
    if (k + 1 < kmax) {
      hipLaunchKernelGGL(k_normalize_into, dim3(vgrid(n)), dim3(256), 0, s, K.scal.p, K.w.p, K.Q.p + (int64_t)(k + 1) * n,
                         n);
      VMPS_HIP_CHECK(hipGetLastError());
    }
  }
  if (!converged) k = res.iterations;
  // Ritz vector psi = Q y (k basis vectors), normalised, then the true residual
  VMPS_HIP_CHECK(hipMemcpyAsync(K.c.p, y.data(), sizeof(double) * k, hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(k_combination, dim3(vgrid(n)), dim3(256), 0, s, K.Q.p, n, K.c.p, k, K.w.p);
  D.dot_async(K.w.p, K.w.p, K.scal.p, s);
  hipLaunchKernelGGL(k_normalize_into, dim3(vgrid(n)), dim3(256), 0, s, K.scal.p, K.w.p, d_v, n);
  VMPS_HIP_CHECK(hipGetLastError());
  // Rayleigh quotient and residual with the requested precision's operator
  D.hxv(d_v, K.w.p, s);
  ++res.hxv_calls;
  D.dot_async(d_v, K.w.p, K.scal.p, s);
  double rq = 0.0;
  VMPS_HIP_CHECK(hipMemcpyAsync(&rq, K.scal.p, sizeof(double), hipMemcpyDeviceToHost, s));
  VMPS_HIP_CHECK(hipStreamSynchronize(s));
  dev_axpby(-rq, d_v, 1.0, K.w.p, n, s);
  D.dot_async(K.w.p, K.w.p, K.scal.p, s);
  double r2 = 0.0;
  VMPS_HIP_CHECK(hipMemcpyAsync(&r2, K.scal.p, sizeof(double), hipMemcpyDeviceToHost, s));
  VMPS_HIP_CHECK(hipStreamSynchronize(s));
  res.energy = rq;
  res.ritz = E;
  res.residual = std::sqrt(std::max(0.0, r2));
  res.converged = converged;
  D.prec = saved;
  return res;
}

void Krylov::ensure(DevPlan& D, int kmax) {
  const int64_t n = D.H.vsize;
  if (cap >= kmax && nvec == n) return;
  Q.alloc((size_t)(kmax + 1) * n);
  w.alloc(n);
  c.alloc(2 * (size_t)kmax + 2);
  scal.alloc(8);
  const int nseg = (int)D.H.seg_off.size();
  const int nb = (nseg + 15) / 16;
  part.alloc((size_t)std::max(1, nb) * (kmax + 1));
  cap = kmax;
  nvec = n;
}

}  // namespace vmps
