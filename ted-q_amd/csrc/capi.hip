// C ABI plumbing: error reporting and device queries shared by all entry points.
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include "tedq_common.h"

namespace tedq {

static thread_local char g_last_error[512] = {0};

void set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_last_error, sizeof(g_last_error), fmt, ap);
  va_end(ap);
}

int check_launch(const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    set_error("%s: launch failed: %s", what, hipGetErrorString(e));
    return kLaunchFailed;
  }
  return kOk;
}

}  // namespace tedq

extern "C" const char* tedq_last_error(void) { return tedq::g_last_error; }

extern "C" int tedq_abi_version(void) { return 1; }

// Fills props[0..3] = {multiprocessor count, max shared memory per block,
// warp size, gcnArch major*100+minor} for the current device.
extern "C" int tedq_device_info(int* props) {
  if (props == nullptr) {
    tedq::set_error("tedq_device_info: null output");
    return tedq::kBadArgument;
  }
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  hipDeviceProp_t p;
  if (e == hipSuccess) e = hipGetDeviceProperties(&p, dev);
  if (e != hipSuccess) {
    tedq::set_error("tedq_device_info: %s", hipGetErrorString(e));
    return tedq::kLaunchFailed;
  }
  props[0] = p.multiProcessorCount;
  props[1] = (int)p.sharedMemPerBlock;
  props[2] = p.warpSize;
  props[3] = p.major * 100 + p.minor;
  return tedq::kOk;
}
