// Environment held on the device (vmps_denv of include/vmps_hxv.h): the
// sector and block layout on the host, the packed FP64 blocks in one device
// allocation of the device it was made on.
#pragma once
#include <hip/hip_runtime.h>

#include "pivot.hpp"

struct vmps_denv {
  int nq = 0;
  int device = -1;
  vmps::Tripod T;  // T.dev = buf, T.dev_size = size, T.data empty
  double* buf = nullptr;
  ~vmps_denv() {
    if (buf) (void)hipFree(buf);
  }
};
