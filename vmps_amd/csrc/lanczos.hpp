// Krylov eigensolver on a device plan (see lanczos.hip).
#pragma once
#include <vector>

#include "engine.hpp"

namespace vmps {

struct LanczosOpts {
  int max_iter = 100;
  int min_iter = 2;
  double tol_eigval = 1e-8;  // DmrgTypedefs.h:371 default
  double tol_state = 1e-5;   // DmrgTypedefs.h:372 default
  bool graph = true;         // replay one captured HIP graph per iteration
  // one more HxV for the returned vector's Rayleigh quotient and residual;
  // without it (a refinement that forms both with the FP64 operator
  // follows) energy / residual are the Ritz value and the Lanczos estimate
  bool final_rq = true;
};

struct LanczosResult {
  double energy = 0.0;    // Rayleigh quotient <psi|H|psi> of the returned vector
  double ritz = 0.0;      // lowest eigenvalue of the tridiagonal matrix
  double residual = 0.0;  // ||H psi - E psi|| (weighted norm)
  int iterations = 0;
  int hxv_calls = 0;
  bool converged = false;
};

LanczosResult lanczos_solve(DevPlan& D, double* d_v, const LanczosOpts& o, vmps_precision prec);

struct RefineOpts {
  double tol_eigval = 1e-12;  // psi accepted when rho <= tol_state and rho^2 <= tol_eigval
  double tol_state = 1e-7;    // (rho = ||H psi - E psi||, FP64 operator)
  int max_outer = 4;
  int cg_max = 200;           // CG iterations per correction
  double cg_tol_min = 1e-4;   // finest CG tolerance relative to the residual
};

struct RefineResult {
  double energy = 0.0;    // <psi|H|psi>, FP64 operator
  double residual = 0.0;  // ||H psi - E psi||, FP64 operator
  double energy0 = 0.0;   // the same two for the start vector
  double residual0 = 0.0;
  int outer = 0;          // correction steps taken
  int cg_iterations = 0;  // mixed-precision HxV of the corrections
  int hxv_fp64 = 0;
  bool converged = false;
};

// A piece of a batch vector inside one pivot's segment (batch Lanczos).
struct BChunk {
  int64_t start;
  int32_t len;
  int32_t piv;
};

// Workspace of the batch Lanczos (lanczos_batch.hip): one Krylov basis over
// the concatenated vector of a batch plan, every reduction segmented by pivot.
struct BatchKrylov {
  static constexpr int kIterBatch = 4;
  static constexpr int64_t kChunk = 4096;
  int P = 0;
  int64_t n = 0;
  std::vector<int64_t> off;   // pivot p owns [off[p], off[p+1])
  DevBuf<BChunk> ch;          // chunks, pivot by pivot
  DevBuf<int32_t> cb;         // pivot p's chunks are [cb[p], cb[p+1])
  int nc = 0;
  DevBuf<double> Q, w, v, c, gpart, npart, nv, out3, ew, Y, spart, sres, coefa, coefb;
  DevBuf<int> kdev;
  double* host = nullptr;     // pinned readback of one graph launch's scalars
  int cap = 0;
  hipGraph_t graph[3] = {nullptr, nullptr, nullptr};
  hipGraphExec_t exec[3] = {nullptr, nullptr, nullptr};
  int graph_gen[3] = {-1, -1, -1};
  void setup(DevPlan& D, const std::vector<int64_t>& offsets);
  void ensure(DevPlan& D, int kmax);
  void enqueue_iteration(DevPlan& D, hipStream_t s, int kofs);
  hipGraphExec_t ensure_graph(DevPlan& D);
  void release_graph();
  std::vector<double> seg_dot(DevPlan& D, const double* x, const double* y);
  void piv_axpby(DevPlan& D, const std::vector<double>& a, const double* x, const std::vector<double>& b, double* y);
  BatchKrylov() = default;
  BatchKrylov(const BatchKrylov&) = delete;
  BatchKrylov& operator=(const BatchKrylov&) = delete;
  ~BatchKrylov();
};

// One lowest eigenpair per pivot of a batch plan; d_v holds the start vectors
// on entry and the eigenvectors on exit.
std::vector<LanczosResult> lanczos_solve_batch(DevPlan& D, BatchKrylov& K, double* d_v, const LanczosOpts& o,
                                               vmps_precision prec);

// FP64 refinement of an approximate lowest eigenpair (d_v: unit vector in
// the weighted inner product) with corrections solved in precision `inner`.
RefineResult refine_eigenpair(DevPlan& D, double* d_v, const RefineOpts& o, vmps_precision inner);
void tridiag_lowest(const std::vector<double>& a, const std::vector<double>& b, int k, double& E,
                    std::vector<double>& y);
void dev_axpby(double a, const double* x, double b, double* y, int64_t n, hipStream_t s);
void dev_fill_uniform(double* v, int64_t n, uint64_t seed, hipStream_t s);

}  // namespace vmps
