// Environment held on the device (vmps_denv of include/vmps_hxv.h): the
// sector and block layout on the host, the packed FP64 blocks in one device
// allocation of the device it was made on.
#pragma once
#include <hip/hip_runtime.h>

#include <memory>

#include "pivot.hpp"

namespace vmps {
// Takes ownership of a device allocation made on `device`: the last owner
// (the vmps_denv, or a pivot / contraction that took its Tripod) frees it.
inline std::shared_ptr<const double> own_device(double* p, int device) {
  return std::shared_ptr<const double>(p, [device](const double* q) {
    if (!q) return;
    int cur = -1;
    const bool restore = hipGetDevice(&cur) == hipSuccess && cur != device;
    if (restore) (void)hipSetDevice(device);
    (void)hipFree(const_cast<double*>(q));
    if (restore) (void)hipSetDevice(cur);
  });
}
}  // namespace vmps

struct vmps_denv {
  int nq = 0;
  int device = -1;
  vmps::Tripod T;  // T.dev / T.dev_hold = the device blocks, T.data empty
};
