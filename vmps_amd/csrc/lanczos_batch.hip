// One Lanczos solve per pivot of a batch plan, all pivots iterating together.
//
// The reference solves the independent pivots of a VUMPS unit cell one per
// thread (VUMPS/VumpsSolver.h:1168-1211: an OpenMP loop of LanczosSolver
// calls).  A batch plan (vmps_plan_create_batch) applies all of them in one
// launch per stage, with the pivots' vectors concatenated.  Here the Lanczos
// recurrence runs on that concatenated vector with every reduction segmented
// by pivot: each pivot keeps its own Gram-Schmidt coefficients, alpha_k,
// beta_k, tridiagonal matrix and convergence test, exactly as if it were
// solved alone (classical Gram-Schmidt twice against its own basis, weighted
// by coeff_dot).  The vector is cut into chunks that never straddle a pivot;
// dot products are chunk partial sums reduced per pivot in a fixed order.
// One iteration of the whole batch is a fixed kernel sequence reading the
// iteration index from device memory, captured as one HIP graph (four
// iterations per launch) like the single-pivot solver.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "engine.hpp"
#include "lanczos.hpp"

namespace vmps {

namespace {

inline unsigned bgrid(int64_t items) { return (unsigned)std::max<int64_t>(1, std::min<int64_t>(4096, items)); }

__device__ inline double wg_sum_b(double x) {
  __shared__ double ws[4];
  for (int m = 32; m > 0; m >>= 1) x += __shfl_xor(x, m);
  if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = x;
  __syncthreads();
  const double r = (ws[0] + ws[1]) + (ws[2] + ws[3]);
  __syncthreads();
  return r;
}

// partial[i * nc + b] = sum over chunk b of ew Q_i w, rows i = 4y .. 4y+3 <= k
__global__ __launch_bounds__(256) void k_bdot_rows(const int* __restrict__ kdev, int kofs,
                                                   const double* __restrict__ Q, int64_t n,
                                                   const double* __restrict__ w, const double* __restrict__ ew,
                                                   const BChunk* __restrict__ ch, int nc,
                                                   double* __restrict__ partial) {
  const int k = *kdev + kofs;
  const int i0 = 4 * blockIdx.y;
  if (i0 > k) return;  // workgroup-uniform
  __shared__ double ws[4][4];
  const int nr = min(4, k + 1 - i0);
  const BChunk c = ch[blockIdx.x];
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  for (int64_t e = c.start + threadIdx.x; e < c.start + c.len; e += 256) {
    const double we = w[e] * ew[e];
#pragma unroll
    for (int r = 0; r < 4; ++r)
      if (r < nr) acc[r] += Q[(int64_t)(i0 + r) * n + e] * we;
  }
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    double x = acc[r];
    for (int m = 32; m > 0; m >>= 1) x += __shfl_xor(x, m);
    if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6][r] = x;
  }
  __syncthreads();
  const int t = threadIdx.x;
  if (t < nr) partial[(int64_t)(i0 + t) * nc + blockIdx.x] = (ws[0][t] + ws[1][t]) + (ws[2][t] + ws[3][t]);
}

// c[i * P + p] = sum of partial[i * nc + b] over the chunks b of pivot p;
// workgroup (p, i), rows past k exit
__global__ __launch_bounds__(256) void k_bdot_final(const int* __restrict__ kdev, int kofs,
                                                    const double* __restrict__ partial, int nc,
                                                    const int32_t* __restrict__ cb, int P, double* __restrict__ c) {
  const int p = blockIdx.x, i = blockIdx.y;
  if (i > *kdev + kofs) return;  // workgroup-uniform
  double acc = 0.0;
  for (int b = cb[p] + threadIdx.x; b < cb[p + 1]; b += 256) acc += partial[(int64_t)i * nc + b];
  acc = wg_sum_b(acc);
  if (threadIdx.x == 0) c[(int64_t)i * P + p] = acc;
}

// w -= sum_{i<=k} c[i * P + p] Q_i over every chunk (p its pivot); with
// npart, the chunk's weighted ||w||^2 after the subtraction
constexpr int kBatchMaxRows = 1024;
__global__ __launch_bounds__(256) void k_bsub(const double* __restrict__ Q, int64_t n, const double* __restrict__ c,
                                              int P, const int* __restrict__ kdev, int kofs,
                                              const BChunk* __restrict__ ch, int nc, double* __restrict__ w,
                                              const double* __restrict__ ew, double* __restrict__ npart) {
  __shared__ double cs[kBatchMaxRows];
  const int k = *kdev + kofs;
  for (int b = blockIdx.x; b < nc; b += gridDim.x) {  // workgroup-uniform loop
    const BChunk cbk = ch[b];
    __syncthreads();  // cs of the previous chunk is no longer read
    for (int i = threadIdx.x; i <= k; i += 256) cs[i] = c[(int64_t)i * P + cbk.piv];
    __syncthreads();
    double nrm = 0.0;
    for (int64_t e = cbk.start + threadIdx.x; e < cbk.start + cbk.len; e += 256) {
      double s = 0.0;
      for (int i = 0; i <= k; ++i) s += cs[i] * Q[(int64_t)i * n + e];
      const double y = w[e] - s;
      w[e] = y;
      nrm += ew[e] * y * y;
    }
    if (npart) {  // kernel-uniform
      const double t = wg_sum_b(nrm);
      if (threadIdx.x == 0) npart[b] = t;
    }
  }
}

// nv[p] = sum of npart over the chunks of p; the iteration's three numbers
// per pivot into out[4 (k P + p) ..]
__global__ __launch_bounds__(256) void k_bnorm_final(const int* __restrict__ kdev, int kofs,
                                                     const double* __restrict__ npart, const int32_t* __restrict__ cb,
                                                     int P, const double* __restrict__ c1,
                                                     const double* __restrict__ c2, double* __restrict__ nv,
                                                     double* __restrict__ out) {
  const int p = blockIdx.x;
  const int k = *kdev + kofs;
  double acc = 0.0;
  for (int b = cb[p] + threadIdx.x; b < cb[p + 1]; b += 256) acc += npart[b];
  acc = wg_sum_b(acc);
  if (threadIdx.x == 0) {
    nv[p] = acc;
    const int64_t o = 4 * ((int64_t)k * P + p);
    out[o] = c1[(int64_t)k * P + p];
    out[o + 1] = c2[(int64_t)k * P + p];
    out[o + 2] = acc;
  }
}

// q_{k+1} = w / sqrt(nv[p]) into row k+1 of the basis and into the HxV input
// (zero for a zero w: that pivot's recurrence has broken down)
__global__ __launch_bounds__(256) void k_bnormalize(const int* __restrict__ kdev, int kofs,
                                                    const double* __restrict__ nv, const BChunk* __restrict__ ch,
                                                    int nc, const double* __restrict__ w, double* __restrict__ Q,
                                                    int64_t n, double* __restrict__ v) {
  const int k = *kdev + kofs;
  double* __restrict__ q = Q + (int64_t)(k + 1) * n;
  for (int b = blockIdx.x; b < nc; b += gridDim.x) {
    const BChunk cbk = ch[b];
    const double x = nv[cbk.piv];
    const double f = x > 0.0 ? 1.0 / sqrt(x) : 0.0;
    for (int64_t e = cbk.start + threadIdx.x; e < cbk.start + cbk.len; e += 256) {
      const double y = f * w[e];
      q[e] = y;
      v[e] = y;
    }
  }
}

// segmented weighted dot: part[b] = sum over chunk b of ew x y
__global__ __launch_bounds__(256) void k_bseg_dot(const double* __restrict__ x, const double* __restrict__ y,
                                                  const double* __restrict__ ew, const BChunk* __restrict__ ch,
                                                  int nc, double* __restrict__ part) {
  for (int b = blockIdx.x; b < nc; b += gridDim.x) {
    const BChunk cbk = ch[b];
    double acc = 0.0;
    for (int64_t e = cbk.start + threadIdx.x; e < cbk.start + cbk.len; e += 256) acc += ew[e] * x[e] * y[e];
    acc = wg_sum_b(acc);
    if (threadIdx.x == 0) part[b] = acc;
  }
}

__global__ __launch_bounds__(256) void k_bseg_final(const double* __restrict__ part, const int32_t* __restrict__ cb,
                                                    int P, double* __restrict__ res) {
  const int p = blockIdx.x;
  double acc = 0.0;
  for (int b = cb[p] + threadIdx.x; b < cb[p + 1]; b += 256) acc += part[b];
  acc = wg_sum_b(acc);
  if (threadIdx.x == 0) res[p] = acc;
}

// y = alpha[p] x + beta[p] y per chunk (p its pivot)
__global__ __launch_bounds__(256) void k_bpiv_axpby(const double* __restrict__ alpha, const double* __restrict__ x,
                                                    const double* __restrict__ beta, double* __restrict__ y,
                                                    const BChunk* __restrict__ ch, int nc) {
  for (int b = blockIdx.x; b < nc; b += gridDim.x) {
    const BChunk cbk = ch[b];
    const double a = alpha[cbk.piv], bb = beta[cbk.piv];
    for (int64_t e = cbk.start + threadIdx.x; e < cbk.start + cbk.len; e += 256) y[e] = a * x[e] + bb * y[e];
  }
}

// y = sum_{i < kuse} Y[p * ldy + i] Q_i per chunk (Ritz vectors; Y zero past each pivot's own count)
__global__ __launch_bounds__(256) void k_bcombination(const double* __restrict__ Q, int64_t n,
                                                      const double* __restrict__ Y, int ldy, int kuse,
                                                      const BChunk* __restrict__ ch, int nc, double* __restrict__ y) {
  for (int b = blockIdx.x; b < nc; b += gridDim.x) {
    const BChunk cbk = ch[b];
    const double* __restrict__ yc = Y + (int64_t)cbk.piv * ldy;
    for (int64_t e = cbk.start + threadIdx.x; e < cbk.start + cbk.len; e += 256) {
      double s = 0.0;
      for (int i = 0; i < kuse; ++i) s += yc[i] * Q[(int64_t)i * n + e];
      y[e] = s;
    }
  }
}

__global__ void k_badd_int(int* __restrict__ p, int d) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *p += d;
}

__global__ void k_bset_int(int* __restrict__ p, int value) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *p = value;
}

}  // namespace

// ---- workspace ---------------------------------------------------------------
void BatchKrylov::setup(DevPlan& D, const std::vector<int64_t>& offsets) {
  P = (int)offsets.size() - 1;
  n = D.H.vsize;
  off = offsets;
  std::vector<BChunk> hc;
  std::vector<int32_t> hb(1, 0);
  for (int p = 0; p < P; ++p) {
    for (int64_t s = off[p]; s < off[p + 1]; s += kChunk)
      hc.push_back({s, (int32_t)std::min<int64_t>(kChunk, off[p + 1] - s), p});
    hb.push_back((int32_t)hc.size());
  }
  nc = (int)hc.size();
  ch.upload(hc, D.stream);
  cb.upload(hb, D.stream);
  std::vector<double> hw((size_t)n, 1.0);
  for (size_t sg = 0; sg < D.H.seg_off.size(); ++sg)
    for (int32_t e = 0; e < D.H.seg_len[sg]; ++e) hw[(size_t)(D.H.seg_off[sg] + e)] = D.H.seg_w[sg];
  ew.upload(hw, D.stream);
  nv.alloc(std::max(1, P));
  sres.alloc(std::max(1, P));
  coefa.alloc(std::max(1, P));
  coefb.alloc(std::max(1, P));
  spart.alloc(std::max(1, nc));
  npart.alloc(std::max(1, nc));
  kdev.alloc(1);
  VMPS_HIP_CHECK(hipStreamSynchronize(D.stream));
  cap = 0;
  release_graph();
}

void BatchKrylov::ensure(DevPlan& D, int kmax) {
  const int need = kmax + kIterBatch;
  if (need + 1 > kBatchMaxRows) throw VmpsError(VMPS_ERR_ARG, "max_iter too large for the batch solver (<= 1000)");
  if (cap >= need) return;
  release_graph();
  Q.alloc((size_t)(need + 1) * n);
  w.alloc(n);
  v.alloc(n);
  c.alloc(2 * (size_t)(need + 1) * P);
  gpart.alloc((size_t)(need + 1) * nc);
  out3.alloc(4 * (size_t)need * P);
  Y.alloc((size_t)P * (need + 1));
  if (host) (void)hipHostFree(host);
  host = nullptr;
  VMPS_HIP_CHECK(hipHostMalloc((void**)&host, 4 * (size_t)kIterBatch * P * sizeof(double), hipHostMallocDefault));
  VMPS_HIP_CHECK(hipStreamSynchronize(D.stream));
  cap = need;
}

void BatchKrylov::enqueue_iteration(DevPlan& D, hipStream_t s, int kofs) {
  const unsigned gy = (unsigned)((cap + 3) / 4);
  double* c1 = c.p;
  double* c2 = c.p + (size_t)(cap + 1) * P;
  auto pass = [&](double* c_out, double* norm_part) {
    hipLaunchKernelGGL(k_bdot_rows, dim3(nc, gy), dim3(256), 0, s, kdev.p, kofs, Q.p, n, w.p, ew.p, ch.p, nc,
                       gpart.p);
    hipLaunchKernelGGL(k_bdot_final, dim3(P, cap), dim3(256), 0, s, kdev.p, kofs, gpart.p, nc, cb.p, P, c_out);
    hipLaunchKernelGGL(k_bsub, dim3(bgrid(nc)), dim3(256), 0, s, Q.p, n, c_out, P, kdev.p, kofs, ch.p, nc, w.p,
                       ew.p, norm_part);
  };
  D.hxv(v.p, w.p, s);
  pass(c1, nullptr);
  pass(c2, npart.p);
  hipLaunchKernelGGL(k_bnorm_final, dim3(P), dim3(256), 0, s, kdev.p, kofs, npart.p, cb.p, P, c1, c2, nv.p, out3.p);
  hipLaunchKernelGGL(k_bnormalize, dim3(bgrid(nc)), dim3(256), 0, s, kdev.p, kofs, nv.p, ch.p, nc, w.p, Q.p, n,
                     v.p);
  VMPS_HIP_CHECK(hipGetLastError());
}

void BatchKrylov::release_graph() {
  for (int i = 0; i < 3; ++i) {
    if (exec[i]) (void)hipGraphExecDestroy(exec[i]);
    if (graph[i]) (void)hipGraphDestroy(graph[i]);
    exec[i] = nullptr;
    graph[i] = nullptr;
    graph_gen[i] = -1;
  }
}

hipGraphExec_t BatchKrylov::ensure_graph(DevPlan& D) {
  const int p = (int)D.prec;
  if (exec[p] && graph_gen[p] == D.gen) return exec[p];
  if (exec[p]) (void)hipGraphExecDestroy(exec[p]);
  if (graph[p]) (void)hipGraphDestroy(graph[p]);
  exec[p] = nullptr;
  graph[p] = nullptr;
  hipStream_t s = D.stream;
  VMPS_HIP_CHECK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
  try {
    for (int i = 0; i < kIterBatch; ++i) enqueue_iteration(D, s, i);
    hipLaunchKernelGGL(k_badd_int, dim3(1), dim3(64), 0, s, kdev.p, kIterBatch);
  } catch (...) {
    hipGraph_t g = nullptr;
    (void)hipStreamEndCapture(s, &g);
    if (g) (void)hipGraphDestroy(g);
    throw;
  }
  VMPS_HIP_CHECK(hipStreamEndCapture(s, &graph[p]));
  VMPS_HIP_CHECK(hipGraphInstantiate(&exec[p], graph[p], nullptr, nullptr, 0));
  graph_gen[p] = D.gen;
  return exec[p];
}

BatchKrylov::~BatchKrylov() {
  release_graph();
  if (host) (void)hipHostFree(host);
}

// per-pivot weighted dots <x, y> of the batch vectors (synchronous)
std::vector<double> BatchKrylov::seg_dot(DevPlan& D, const double* x, const double* y) {
  hipStream_t s = D.stream;
  hipLaunchKernelGGL(k_bseg_dot, dim3(bgrid(nc)), dim3(256), 0, s, x, y, ew.p, ch.p, nc, spart.p);
  hipLaunchKernelGGL(k_bseg_final, dim3(P), dim3(256), 0, s, spart.p, cb.p, P, sres.p);
  VMPS_HIP_CHECK(hipGetLastError());
  std::vector<double> r((size_t)P);
  VMPS_HIP_CHECK(hipMemcpyAsync(r.data(), sres.p, sizeof(double) * P, hipMemcpyDeviceToHost, s));
  VMPS_HIP_CHECK(hipStreamSynchronize(s));
  return r;
}

// y = a[p] x + b[p] y
void BatchKrylov::piv_axpby(DevPlan& D, const std::vector<double>& a, const double* x, const std::vector<double>& b,
                            double* y) {
  hipStream_t s = D.stream;
  VMPS_HIP_CHECK(hipMemcpyAsync(coefa.p, a.data(), sizeof(double) * P, hipMemcpyHostToDevice, s));
  VMPS_HIP_CHECK(hipMemcpyAsync(coefb.p, b.data(), sizeof(double) * P, hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(k_bpiv_axpby, dim3(bgrid(nc)), dim3(256), 0, s, coefa.p, x, coefb.p, y, ch.p, nc);
  VMPS_HIP_CHECK(hipGetLastError());
  VMPS_HIP_CHECK(hipStreamSynchronize(s));  // the host vectors a, b are read before they go out of scope
}

// ---- the solve -----------------------------------------------------------------
std::vector<LanczosResult> lanczos_solve_batch(DevPlan& D, BatchKrylov& K, double* d_v, const LanczosOpts& o,
                                               vmps_precision prec) {
  hipStream_t s = D.stream;
  const int P = K.P;
  std::vector<LanczosResult> res((size_t)P);
  if (K.n == 0 || P == 0) return res;
  const int kmax = std::max(2, (int)std::min<int64_t>(o.max_iter, std::max<int64_t>(1, K.n)));
  const vmps_precision saved = D.prec;
  D.prec = prec;
  try {
    D.ensure_precision_buffers();
    K.ensure(D, kmax);
    hipGraphExec_t iteration = o.graph ? K.ensure_graph(D) : nullptr;
    // q0 = v / ||v|| per pivot, into the basis and the HxV input
    std::vector<double> nrm = K.seg_dot(D, d_v, d_v), a(P), z(P, 0.0);
    for (int p = 0; p < P; ++p) {
      if (K.off[p + 1] > K.off[p] && !(nrm[p] > 0.0))
        throw VmpsError(nrm[p] == 0.0 ? VMPS_ERR_ARG : VMPS_ERR_NONFINITE,
                        "the start vector of a pivot of the batch is zero or not finite");
      a[p] = nrm[p] > 0.0 ? 1.0 / std::sqrt(nrm[p]) : 0.0;
    }
    K.piv_axpby(D, a, d_v, z, K.Q.p);
    K.piv_axpby(D, a, d_v, z, K.v.p);
    hipLaunchKernelGGL(k_bset_int, dim3(1), dim3(64), 0, s, K.kdev.p, 0);
    VMPS_HIP_CHECK(hipGetLastError());

    std::vector<std::vector<double>> alpha(P), beta(P), y(P);
    std::vector<double> E(P, 0.0), Eprev(P, 0.0);
    std::vector<char> done(P, 0);
    for (int p = 0; p < P; ++p)
      if (K.off[p + 1] == K.off[p]) done[p] = 1;  // an empty pivot has nothing to solve
    int k = 0, hxv = 0;
    auto all_done = [&]() {
      for (int p = 0; p < P; ++p)
        if (!done[p]) return false;
      return true;
    };
    while (k < kmax && !all_done()) {
      int nb = 1;
      if (iteration) {
        VMPS_HIP_CHECK(hipGraphLaunch(iteration, s));
        nb = BatchKrylov::kIterBatch;
      } else {
        K.enqueue_iteration(D, s, 0);
        hipLaunchKernelGGL(k_badd_int, dim3(1), dim3(64), 0, s, K.kdev.p, 1);
      }
      hxv += nb;
      VMPS_HIP_CHECK(hipMemcpyAsync(K.host, K.out3.p + 4 * (size_t)k * P, 4 * (size_t)nb * P * sizeof(double),
                                    hipMemcpyDeviceToHost, s));
      VMPS_HIP_CHECK(hipStreamSynchronize(s));
      for (int j = 0; j < nb && k < kmax; ++j, ++k) {
        for (int p = 0; p < P; ++p) {
          if (done[p]) continue;
          const double* h3 = K.host + 4 * ((size_t)j * P + p);
          const double a_k = h3[0] + h3[1];
          const double b_k = std::sqrt(std::max(0.0, h3[2]));
          alpha[p].push_back(a_k);
          beta[p].push_back(b_k);
          Eprev[p] = E[p];
          tridiag_lowest(alpha[p], beta[p], k + 1, E[p], y[p]);
          const double rest = b_k * std::fabs(y[p][k]);
          res[p].iterations = k + 1;
          const bool breakdown = b_k < 1e-14 * std::max(1.0, std::fabs(a_k));
          if (breakdown || (k + 1 >= o.min_iter && k > 0 && std::fabs(E[p] - Eprev[p]) < o.tol_eigval &&
                            rest < o.tol_state)) {
            done[p] = 1;
            res[p].converged = true;
          }
        }
      }
    }
    // Ritz vectors: pivot p combines its own first iterations[p] basis rows
    const int ldy = K.cap + 1;
    std::vector<double> Yh((size_t)P * ldy, 0.0);
    int kuse = 0;
    for (int p = 0; p < P; ++p) {
      for (size_t i = 0; i < y[p].size(); ++i) Yh[(size_t)p * ldy + i] = y[p][i];
      kuse = std::max(kuse, (int)y[p].size());
      res[p].ritz = E[p];
    }
    VMPS_HIP_CHECK(hipMemcpyAsync(K.Y.p, Yh.data(), sizeof(double) * Yh.size(), hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_bcombination, dim3(bgrid(K.nc)), dim3(256), 0, s, K.Q.p, K.n, K.Y.p, ldy, kuse, K.ch.p,
                       K.nc, K.w.p);
    VMPS_HIP_CHECK(hipGetLastError());
    nrm = K.seg_dot(D, K.w.p, K.w.p);  // synchronises: Yh may go
    for (int p = 0; p < P; ++p) a[p] = nrm[p] > 0.0 ? 1.0 / std::sqrt(nrm[p]) : 0.0;
    K.piv_axpby(D, a, K.w.p, z, d_v);
    // Rayleigh quotients and residuals with the solve's operator
    D.hxv(d_v, K.w.p, s);
    ++hxv;
    const std::vector<double> rq = K.seg_dot(D, d_v, K.w.p);
    std::vector<double> neg(P), one(P, 1.0);
    for (int p = 0; p < P; ++p) neg[p] = -rq[p];
    K.piv_axpby(D, neg, d_v, one, K.w.p);
    const std::vector<double> r2 = K.seg_dot(D, K.w.p, K.w.p);
    for (int p = 0; p < P; ++p) {
      res[p].energy = rq[p];
      res[p].residual = std::sqrt(std::max(0.0, r2[p]));
      res[p].hxv_calls = hxv;
    }
  } catch (...) {
    D.prec = saved;
    throw;
  }
  D.prec = saved;
  return res;
}

}  // namespace vmps
