// Error type carried up to the C ABI, where it becomes a status code plus a
// thread-local message (vmps_last_error).
#pragma once
#include <stdexcept>
#include <string>

#include "vmps_hxv.h"

namespace vmps {

struct VmpsError : std::runtime_error {
  int status;
  VmpsError(int st, const std::string& msg) : std::runtime_error(msg), status(st) {}
};

void set_last_error(const std::string& msg);
const char* last_error();

}  // namespace vmps
