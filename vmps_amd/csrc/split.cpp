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
#include <climits>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <thread>
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

// rocBLAS handles per device, made on first use and kept until the process
// ends, like the device's stream pool: kLanes of them, so that the sectors of
// one split factor concurrently on as many streams (each handle owns its
// workspace, so concurrent work needs one handle per stream).  Calls on one
// device are serialised by the slot's mutex.
constexpr int kLanes = 8;
struct HandleSlot {
  rocblas_handle h[kLanes] = {};
  std::mutex m;
  bool warm[2] = {false, false};  // a factorisation of this side has run here with its lanes on one thread
};
std::mutex g_handles_mutex;
std::map<int, HandleSlot*> g_handles;

HandleSlot& handles_for_current_device() {
  int dev = 0;
  VMPS_HIP_CHECK(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lock(g_handles_mutex);
  HandleSlot*& s = g_handles[dev];
  if (!s) {
    auto* slot = new HandleSlot();
    for (auto& h : slot->h) rb_check(rocblas_create_handle(&h), "rocblas_create_handle");
    s = slot;
  }
  return *s;
}

}  // namespace
}  // namespace vmps

using vmps::VmpsError;

namespace {

// The factorisation of every sector: A is given either on the host (packed
// column-major sectors, copied lane by lane) or already on the device
// (dA_packed, the same layout).  Arguments are validated by the caller.
void factor_sectors(int nblocks, const int64_t* rows, const int64_t* cols, const double* A, const double* dA_packed,
                    int side, double* lambda, double* W, double* C) {
  {
    int64_t maxrc = 0;
    for (int b = 0; b < nblocks; ++b) {
      if (rows[b] < 0 || cols[b] < 0 || rows[b] > (1 << 20) || cols[b] > (1 << 20))
        throw VmpsError(VMPS_ERR_ARG, "bad block dimension");
      const int64_t k = side == 0 ? rows[b] : cols[b];
      if (k > (1 << 16)) throw VmpsError(VMPS_ERR_ARG, "block too large for the eigensolver (side > 65536)");
      maxrc = std::max(maxrc, rows[b] * cols[b]);
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
      return;
    }
    // Sectors go to kLanes streams, largest first, each to the lane with the
    // least work so far (work ~ k^3); every lane runs its sectors in order
    // through its own input buffer, while each sector keeps its own outputs
    // on the device until all lanes are done.
    vmps::HandleSlot& slot = vmps::handles_for_current_device();
    std::lock_guard<std::mutex> hold(slot.m);
    constexpr int NL = vmps::kLanes;
    struct Lanes {  // synchronised and returned to the pool on every exit path
      hipStream_t s[NL] = {};
      ~Lanes() {
        for (auto& x : s)
          if (x) {
            (void)hipStreamSynchronize(x);
            vmps::release_stream(x);
          }
      }
    } lanes;
    for (auto& x : lanes.s) x = vmps::acquire_stream();
    std::vector<int64_t> ao(nblocks), lo(nblocks), wo(nblocks);
    std::vector<int> order;
    int64_t a_acc = 0, l_acc = 0, w_acc = 0;
    for (int b = 0; b < nblocks; ++b) {
      const int64_t k = side == 0 ? rows[b] : cols[b];
      ao[b] = a_acc;
      lo[b] = l_acc;
      wo[b] = w_acc;
      a_acc += rows[b] * cols[b];
      l_acc += k;
      w_acc += k * k;
      if (k == 0) continue;
      if (rows[b] * cols[b] == 0) {  // a k x k zero density matrix
        for (int64_t i = 0; i < k; ++i) lambda[lo[b] + i] = 0.0;
        for (int64_t i = 0; i < k * k; ++i) W[wo[b] + i] = (i % (k + 1) == 0) ? 1.0 : 0.0;
        continue;
      }
      order.push_back(b);
    }
    std::sort(order.begin(), order.end(), [&](int x, int y) {
      const int64_t kx = side == 0 ? rows[x] : cols[x], ky = side == 0 ? rows[y] : cols[y];
      return kx != ky ? kx > ky : x < y;
    });
    std::vector<int> lane(nblocks, -1);
    double work[NL] = {};
    int64_t lane_rc[NL] = {}, lane_k[NL] = {};
    int64_t out_g = 0, out_c = 0;  // device outputs of all sectors
    std::vector<int64_t> g_off(nblocks, 0), c_off(nblocks, 0);
    for (int b : order) {
      const int64_t k = side == 0 ? rows[b] : cols[b];
      int best = 0;
      for (int l = 1; l < NL; ++l)
        if (work[l] < work[best]) best = l;
      lane[b] = best;
      work[best] += (double)k * k * k;
      lane_rc[best] = std::max(lane_rc[best], rows[b] * cols[b]);
      lane_k[best] = std::max(lane_k[best], k);
      g_off[b] = out_g;
      c_off[b] = out_c;
      out_g += k * k;
      out_c += rows[b] * cols[b];
    }
    vmps::DevBuf<double> dA[NL], dE[NL], dG, dC, dL;
    vmps::DevBuf<int> dInfo;
    for (int l = 0; l < NL; ++l) {
      if (lane_rc[l] == 0) continue;
      if (!dA_packed) dA[l].alloc((size_t)lane_rc[l]);
      dE[l].alloc((size_t)std::max<int64_t>(1, lane_k[l]));
    }
    dG.alloc((size_t)std::max<int64_t>(1, out_g));
    dC.alloc((size_t)std::max<int64_t>(1, out_c));
    dL.alloc((size_t)std::max<int64_t>(1, l_acc));
    dInfo.alloc((size_t)std::max(1, nblocks));
    for (int l = 0; l < NL; ++l) vmps::rb_check(rocblas_set_stream(slot.h[l], lanes.s[l]), "rocblas_set_stream");
    // One host thread per lane: rocSOLVER's dsyevd issues thousands of small
    // kernels per sector, so one thread launching for all lanes leaves the
    // device waiting on launches.  Each thread works through its lane's
    // sectors with its own handle and stream and copies their results back.
    int dev = 0;
    VMPS_HIP_CHECK(hipGetDevice(&dev));
    std::string err[NL];
    int err_status[NL] = {};
    auto run_lane = [&](int l) {
      try {
        VMPS_HIP_CHECK(hipSetDevice(dev));
        rocblas_handle h = slot.h[l];
        hipStream_t st = lanes.s[l];
        const double one = 1.0, zero = 0.0;
        for (int b : order) {
          if (lane[b] != l) continue;
          const rocblas_int r = (rocblas_int)rows[b], c = (rocblas_int)cols[b], k = side == 0 ? r : c;
          double* G = dG.p + g_off[b];
          double* Cb = dC.p + c_off[b];
          const double* Ab = dA_packed ? dA_packed + ao[b] : dA[l].p;
          if (!dA_packed)
            VMPS_HIP_CHECK(hipMemcpyAsync(dA[l].p, A + ao[b], sizeof(double) * r * c, hipMemcpyHostToDevice, st));
          // rho = A A^T (r x r) or A^T A (c x c), the full matrix
          if (side == 0)
            vmps::rb_check(rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_transpose, r, r, c, &one, Ab, r,
                                         Ab, r, &zero, G, r),
                           "rocblas_dgemm");
          else
            vmps::rb_check(rocblas_dgemm(h, rocblas_operation_transpose, rocblas_operation_none, c, c, r, &one, Ab, r,
                                         Ab, r, &zero, G, c),
                           "rocblas_dgemm");
          vmps::rb_check(rocsolver_dsyevd(h, rocblas_evect_original, rocblas_fill_upper, k, G, k, dL.p + lo[b],
                                          dE[l].p, dInfo.p + b),
                         "rocsolver_dsyevd");
          // centre: W^T A (k x c) or A W (r x k)
          if (side == 0)
            vmps::rb_check(rocblas_dgemm(h, rocblas_operation_transpose, rocblas_operation_none, r, c, r, &one, G, r,
                                         Ab, r, &zero, Cb, r),
                           "rocblas_dgemm");
          else
            vmps::rb_check(rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_none, r, c, c, &one, Ab, r, G,
                                         c, &zero, Cb, r),
                           "rocblas_dgemm");
        }
        for (int b : order) {
          if (lane[b] != l) continue;
          const int64_t k = side == 0 ? rows[b] : cols[b];
          VMPS_HIP_CHECK(
              hipMemcpyAsync(lambda + lo[b], dL.p + lo[b], sizeof(double) * k, hipMemcpyDeviceToHost, st));
          VMPS_HIP_CHECK(
              hipMemcpyAsync(W + wo[b], dG.p + g_off[b], sizeof(double) * k * k, hipMemcpyDeviceToHost, st));
          VMPS_HIP_CHECK(hipMemcpyAsync(C + ao[b], dC.p + c_off[b], sizeof(double) * rows[b] * cols[b],
                                        hipMemcpyDeviceToHost, st));
        }
        VMPS_HIP_CHECK(hipStreamSynchronize(st));
      } catch (const VmpsError& e) {
        err[l] = e.what();
        err_status[l] = e.status;
      } catch (const std::exception& e) {
        err[l] = e.what();
        err_status[l] = VMPS_ERR_INTERNAL;
      }
    };
    if (!slot.warm[side]) {
      // The first factorisation on a device runs its lanes one after another
      // on this thread: the libraries load their GPU code on first use (the
      // GEMM and eigensolver kernels this side calls), and that loading
      // should not run on eight threads at once.  (A run under the kernel
      // tracer faulted inside such a first launch from a lane thread.)
      for (int l = 0; l < NL; ++l)
        if (lane_rc[l]) run_lane(l);
      slot.warm[side] = true;
    } else {
      std::vector<std::thread> workers;
      for (int l = 1; l < NL; ++l)
        if (lane_rc[l]) workers.emplace_back(run_lane, l);
      if (lane_rc[0]) run_lane(0);
      for (auto& t : workers) t.join();
    }
    for (int l = 0; l < NL; ++l)
      if (err_status[l]) throw VmpsError(err_status[l], err[l]);
    std::vector<int> info(std::max(1, nblocks), 0);
    if (!order.empty()) {
      VMPS_HIP_CHECK(hipMemcpy(info.data(), dInfo.p, sizeof(int) * nblocks, hipMemcpyDeviceToHost));
    }
    for (int b : order)
      if (info[b] != 0)
        throw VmpsError(VMPS_ERR_NOCONV, "rocsolver_dsyevd did not converge (info " + std::to_string(info[b]) + ")");
  }
}

void check_sectors(int nblocks, const int64_t* rows, const int64_t* cols, int side) {
  if (side != 0 && side != 1) throw VmpsError(VMPS_ERR_ARG, "side must be 0 (left) or 1 (right)");
  for (int b = 0; b < nblocks; ++b) {
    if (rows[b] < 0 || cols[b] < 0 || rows[b] > (1 << 20) || cols[b] > (1 << 20))
      throw VmpsError(VMPS_ERR_ARG, "bad block dimension");
    const int64_t k = side == 0 ? rows[b] : cols[b];
    if (k > (1 << 16)) throw VmpsError(VMPS_ERR_ARG, "block too large for the eigensolver (side > 65536)");
  }
}

// Sector matrices from the blocks of a device vector: destination block d is
// sum over its parts p of factor[p] * v[src[p] .. src[p] + r c) (r x c,
// column-major), written at (row0, col0) of its sector.  The work is cut into
// tasks of whole columns of one destination, about 4096 elements each, one
// workgroup per task (grid-strided); parts are added in order.
struct GatherTask {
  int64_t a_off;  // element offset of (row0, col0 + j0) in the packed sectors
  int64_t ld;     // rows of the sector
  int64_t src_off;  // column j0 of the block within a part: j0 * r
  int32_t r, nj;  // block rows, columns of this task
  int64_t p0, p1;  // parts [p0, p1)
};

__global__ __launch_bounds__(256) void k_gather_sectors(const GatherTask* __restrict__ tasks, int ntasks,
                                                        const int64_t* __restrict__ src,
                                                        const double* __restrict__ fac, const double* __restrict__ v,
                                                        double* __restrict__ A) {
  for (int t = blockIdx.x; t < ntasks; t += gridDim.x) {
    const GatherTask g = tasks[t];
    for (int j = 0; j < g.nj; ++j) {
      const int64_t sc = g.src_off + (int64_t)j * g.r;  // column j of the task within a part
      double* __restrict__ dst = A + g.a_off + (int64_t)j * g.ld;
      for (int i = threadIdx.x; i < g.r; i += blockDim.x) {
        double x = 0.0;
        for (int64_t p = g.p0; p < g.p1; ++p) x += fac[p] * v[src[p] + sc + i];
        dst[i] = x;
      }
    }
  }
}

}  // namespace

extern "C" vmps_status vmps_block_eigh_split(int nblocks, const int64_t* rows, const int64_t* cols, const double* A,
                                             int side, double* lambda, double* W, double* C) {
  try {
    if (nblocks < 0 || (nblocks > 0 && (!rows || !cols || !A || !lambda || !W || !C)))
      throw VmpsError(VMPS_ERR_ARG, "bad argument");
    check_sectors(nblocks, rows, cols, side);
    factor_sectors(nblocks, rows, cols, A, nullptr, side, lambda, W, C);
    return VMPS_OK;
  } catch (const VmpsError& e) {
    vmps::set_last_error(e.what());
    return e.status;
  } catch (const std::exception& e) {
    vmps::set_last_error(e.what());
    return VMPS_ERR_INTERNAL;
  }
}

extern "C" vmps_status vmps_block_eigh_split_gather(const double* d_v, int64_t vsize, int nblocks,
                                                    const int64_t* rows, const int64_t* cols, int ndst,
                                                    const int* dst_block, const int64_t* dst_row,
                                                    const int64_t* dst_col, const int64_t* dst_r,
                                                    const int64_t* dst_c, const int64_t* part_begin,
                                                    const int64_t* part_src, const double* part_factor, int side,
                                                    double* lambda, double* W, double* C) {
  try {
    if (nblocks < 0 || ndst < 0 || vsize < 0 || (nblocks > 0 && (!rows || !cols || !lambda || !W || !C)))
      throw VmpsError(VMPS_ERR_ARG, "bad argument");
    if (ndst > 0 && (!d_v || !dst_block || !dst_row || !dst_col || !dst_r || !dst_c || !part_begin))
      throw VmpsError(VMPS_ERR_ARG, "NULL gather table");
    check_sectors(nblocks, rows, cols, side);
    std::vector<int64_t> ao(nblocks + 1, 0);
    for (int b = 0; b < nblocks; ++b) ao[b + 1] = ao[b] + rows[b] * cols[b];
    const int64_t nparts = ndst > 0 ? part_begin[ndst] : 0;
    if (nparts > 0 && (!part_src || !part_factor)) throw VmpsError(VMPS_ERR_ARG, "NULL gather parts");
    // every destination inside its sector, every part inside the vector
    std::vector<GatherTask> g;
    for (int d = 0; d < ndst; ++d) {
      const int b = dst_block[d];
      if (b < 0 || b >= nblocks) throw VmpsError(VMPS_ERR_ARG, "gather: sector index out of range");
      if (dst_r[d] < 0 || dst_c[d] < 0 || dst_row[d] < 0 || dst_col[d] < 0 || dst_row[d] + dst_r[d] > rows[b] ||
          dst_col[d] + dst_c[d] > cols[b])
        throw VmpsError(VMPS_ERR_LAYOUT, "gather: block outside its sector");
      if (part_begin[d] < 0 || part_begin[d + 1] < part_begin[d] || part_begin[d + 1] > nparts)
        throw VmpsError(VMPS_ERR_ARG, "gather: bad part range");
      for (int64_t p = part_begin[d]; p < part_begin[d + 1]; ++p)
        if (part_src[p] < 0 || part_src[p] + dst_r[d] * dst_c[d] > vsize)
          throw VmpsError(VMPS_ERR_LAYOUT, "gather: part outside the vector");
      if (dst_r[d] > INT32_MAX) throw VmpsError(VMPS_ERR_ARG, "gather: block too large");
      if (dst_r[d] * dst_c[d] == 0) continue;
      const int64_t per = std::max<int64_t>(1, 4096 / dst_r[d]);  // columns per task
      for (int64_t j0 = 0; j0 < dst_c[d]; j0 += per)
        g.push_back({ao[b] + dst_row[d] + (dst_col[d] + j0) * rows[b], rows[b], j0 * dst_r[d], (int32_t)dst_r[d],
                     (int32_t)std::min(per, dst_c[d] - j0), part_begin[d], part_begin[d + 1]});
    }
    const int ntasks = (int)g.size();
    const int64_t total = ao[nblocks];
    vmps::DevBuf<double> dA;
    dA.alloc((size_t)std::max<int64_t>(1, total));
    hipStream_t s = vmps::acquire_stream();
    struct Release {
      hipStream_t s;
      ~Release() {
        (void)hipStreamSynchronize(s);
        vmps::release_stream(s);
      }
    } rel{s};
    if (total) VMPS_HIP_CHECK(hipMemsetAsync(dA.p, 0, sizeof(double) * total, s));
    if (ntasks > 0) {
      vmps::DevBuf<GatherTask> dg;
      vmps::DevBuf<int64_t> dsrc;
      vmps::DevBuf<double> dfac;
      dg.upload(g, s);
      dsrc.upload(std::vector<int64_t>(part_src, part_src + std::max<int64_t>(1, nparts)), s);
      dfac.upload(std::vector<double>(part_factor, part_factor + std::max<int64_t>(1, nparts)), s);
      hipLaunchKernelGGL(k_gather_sectors, dim3((unsigned)std::min(ntasks, 65535)), dim3(256), 0, s, dg.p, ntasks,
                         dsrc.p, dfac.p, d_v, dA.p);
      VMPS_HIP_CHECK(hipGetLastError());
      VMPS_HIP_CHECK(hipStreamSynchronize(s));  // the tables are freed on leaving this block
    }
    VMPS_HIP_CHECK(hipStreamSynchronize(s));
    factor_sectors(nblocks, rows, cols, nullptr, dA.p, side, lambda, W, C);
    return VMPS_OK;
  } catch (const VmpsError& e) {
    vmps::set_last_error(e.what());
    return e.status;
  } catch (const std::exception& e) {
    vmps::set_last_error(e.what());
    return VMPS_ERR_INTERNAL;
  }
}
