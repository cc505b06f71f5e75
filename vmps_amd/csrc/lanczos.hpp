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
void tridiag_lowest(const std::vector<double>& a, const std::vector<double>& b, int k, double& E,
                    std::vector<double>& y);
void dev_axpby(double a, const double* x, double b, double* y, int64_t n, hipStream_t s);
void dev_fill_uniform(double* v, int64_t n, uint64_t seed, hipStream_t s);

}  // namespace vmps
