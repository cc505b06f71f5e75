// Density-matrix split of the two-site wavefunction on the device.
//
// The reference splits every middle sector of the two-site wavefunction with
// a dense SVD (Biped::truncateSVD with Eigen's BDCSVD, tensors/Biped.h:741-897,
// called from split_AA2, tensors/DmrgContractions.h:2460-2630).  Only the
// factor that becomes the new orthonormal site tensor must be orthonormal,
// and the truncation needs the squared singular values; both come from the
// eigendecomposition of the reduced density matrix of that side (White's
// original DMRG truncation):
//   left  (sweep to the right):  rho = Psi Psi^T = W diag(lambda) W^T,
//                                 centre C = W^T Psi  (= S V^T in SVD terms)
//   right (sweep to the left):   rho = Psi^T Psi = W diag(lambda) W^T,
//                                 centre C = Psi W   (= U S)
// with lambda = sigma^2 exactly (eigenvalues to absolute accuracy
// eps ||Psi||^2, which is what the truncation weight sum_discarded deg sigma^2
// needs) and W orthonormal to working accuracy.  No division by a singular
// value occurs.  The dense linear algebra is library work (rocBLAS GEMM and
// the rocSOLVER divide-and-conquer eigensolver, dsyevd); this file only
// stages the blocks.
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "engine.hpp"
#include "errors.hpp"
#include "vmps_hxv.h"

namespace vmps {
void set_last_error(const std::string& msg);

namespace {

void rb_check(rocblas_status st, const char* what) {
  if (st != rocblas_status_success)
    throw VmpsError(VMPS_ERR_HIP, std::string(what) + ": " + rocblas_status_to_string(st));
}

// one rocBLAS handle per device, made on first use (rocBLAS handles keep their
// workspace; the stream is set per call, calls on one handle are serialised)
struct HandleSlot {
  rocblas_handle h = nullptr;
  std::mutex m;
};
std::mutex g_handles_mutex;
std::map<int, HandleSlot*> g_handles;

HandleSlot& handle_for_current_device() {
  int dev = 0;
  VMPS_HIP_CHECK(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lock(g_handles_mutex);
  HandleSlot*& s = g_handles[dev];
  if (!s) {
    auto* slot = new HandleSlot();  // kept until the process ends, like the device's stream pool
    rb_check(rocblas_create_handle(&slot->h), "rocblas_create_handle");
    s = slot;
  }
  return *s;
}

}  // namespace
}  // namespace vmps

using vmps::VmpsError;

extern "C" vmps_status vmps_block_eigh_split(int nblocks, const int64_t* rows, const int64_t* cols, const double* A,
                                             int side, double* lambda, double* W, double* C) {
  try {
    if (nblocks < 0 || (nblocks > 0 && (!rows || !cols || !A || !lambda || !W || !C)))
      throw VmpsError(VMPS_ERR_ARG, "bad argument");
    if (side != 0 && side != 1) throw VmpsError(VMPS_ERR_ARG, "side must be 0 (left) or 1 (right)");
    int64_t maxrc = 0, maxk = 0;
    for (int b = 0; b < nblocks; ++b) {
      if (rows[b] < 0 || cols[b] < 0 || rows[b] > (1 << 20) || cols[b] > (1 << 20))
        throw VmpsError(VMPS_ERR_ARG, "bad block dimension");
      const int64_t k = side == 0 ? rows[b] : cols[b];
      if (k > (1 << 16)) throw VmpsError(VMPS_ERR_ARG, "block too large for the eigensolver (side > 65536)");
      maxrc = std::max(maxrc, rows[b] * cols[b]);
      maxk = std::max(maxk, k);
    }
    if (nblocks == 0 || maxrc == 0) {
      // nothing to factor: empty blocks have empty outputs, but a k x k
      // eigenproblem of a zero block still has k zero eigenvalues
      int64_t lo = 0, wo = 0;
      for (int b = 0; b < nblocks; ++b) {
        const int64_t k = side == 0 ? rows[b] : cols[b];
        for (int64_t i = 0; i < k; ++i) lambda[lo + i] = 0.0;
        for (int64_t i = 0; i < k * k; ++i) W[wo + i] = (i % (k + 1) == 0) ? 1.0 : 0.0;
        lo += k;
        wo += k * k;
      }
      return VMPS_OK;
    }
    vmps::HandleSlot& slot = vmps::handle_for_current_device();
    std::lock_guard<std::mutex> hold(slot.m);
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
    vmps::rb_check(rocblas_set_stream(slot.h, s), "rocblas_set_stream");
    vmps::DevBuf<double> dA, dG, dC, dL, dE;
    vmps::DevBuf<int> dInfo;
    dA.alloc((size_t)maxrc);
    dC.alloc((size_t)maxrc);
    dG.alloc((size_t)std::max<int64_t>(1, maxk * maxk));
    dL.alloc((size_t)std::max<int64_t>(1, maxk));
    dE.alloc((size_t)std::max<int64_t>(1, maxk));
    dInfo.alloc(1);
    const double one = 1.0, zero = 0.0;
    int64_t ao = 0, lo = 0, wo = 0;
    for (int b = 0; b < nblocks; ++b) {
      const int64_t r = rows[b], c = cols[b], k = side == 0 ? r : c;
      if (k == 0) {
        ao += r * c;
        continue;
      }
      if (r * c == 0) {  // a k x k zero density matrix
        for (int64_t i = 0; i < k; ++i) lambda[lo + i] = 0.0;
        for (int64_t i = 0; i < k * k; ++i) W[wo + i] = (i % (k + 1) == 0) ? 1.0 : 0.0;
        lo += k;
        wo += k * k;
        continue;
      }
      VMPS_HIP_CHECK(hipMemcpyAsync(dA.p, A + ao, sizeof(double) * r * c, hipMemcpyHostToDevice, s));
      // rho = A A^T (r x r) or A^T A (c x c), the full matrix
      if (side == 0)
        vmps::rb_check(rocblas_dgemm(slot.h, rocblas_operation_none, rocblas_operation_transpose, (rocblas_int)r,
                                     (rocblas_int)r, (rocblas_int)c, &one, dA.p, (rocblas_int)r, dA.p,
                                     (rocblas_int)r, &zero, dG.p, (rocblas_int)r),
                       "rocblas_dgemm");
      else
        vmps::rb_check(rocblas_dgemm(slot.h, rocblas_operation_transpose, rocblas_operation_none, (rocblas_int)c,
                                     (rocblas_int)c, (rocblas_int)r, &one, dA.p, (rocblas_int)r, dA.p,
                                     (rocblas_int)r, &zero, dG.p, (rocblas_int)c),
                       "rocblas_dgemm");
      vmps::rb_check(rocsolver_dsyevd(slot.h, rocblas_evect_original, rocblas_fill_upper, (rocblas_int)k, dG.p,
                                      (rocblas_int)k, dL.p, dE.p, dInfo.p),
                     "rocsolver_dsyevd");
      // centre: W^T A (k x c) or A W (r x k)
      if (side == 0)
        vmps::rb_check(rocblas_dgemm(slot.h, rocblas_operation_transpose, rocblas_operation_none, (rocblas_int)r,
                                     (rocblas_int)c, (rocblas_int)r, &one, dG.p, (rocblas_int)r, dA.p,
                                     (rocblas_int)r, &zero, dC.p, (rocblas_int)r),
                       "rocblas_dgemm");
      else
        vmps::rb_check(rocblas_dgemm(slot.h, rocblas_operation_none, rocblas_operation_none, (rocblas_int)r,
                                     (rocblas_int)c, (rocblas_int)c, &one, dA.p, (rocblas_int)r, dG.p,
                                     (rocblas_int)c, &zero, dC.p, (rocblas_int)r),
                       "rocblas_dgemm");
      int info = 0;
      VMPS_HIP_CHECK(hipMemcpyAsync(lambda + lo, dL.p, sizeof(double) * k, hipMemcpyDeviceToHost, s));
      VMPS_HIP_CHECK(hipMemcpyAsync(W + wo, dG.p, sizeof(double) * k * k, hipMemcpyDeviceToHost, s));
      VMPS_HIP_CHECK(hipMemcpyAsync(C + ao, dC.p, sizeof(double) * r * c, hipMemcpyDeviceToHost, s));
      VMPS_HIP_CHECK(hipMemcpyAsync(&info, dInfo.p, sizeof(int), hipMemcpyDeviceToHost, s));
      VMPS_HIP_CHECK(hipStreamSynchronize(s));
      if (info != 0) throw VmpsError(VMPS_ERR_NOCONV, "rocsolver_dsyevd did not converge (info " +
                                                          std::to_string(info) + ")");
      ao += r * c;
      lo += k;
      wo += k * k;
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
