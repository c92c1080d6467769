// Run-time specialised pass kernels.
//
// The interpreted pass kernels (sv_passes.hip) read the op list of a pass
// from its program and dispatch every op through a switch over target bit and
// control mask.  On gfx950 that costs more VALU instructions than the gate
// arithmetic itself: each switch arm leaves its results in different
// registers, so the join copies them back, and a CNOT becomes a run of
// register moves.  For a circuit structure that is evaluated many times, the
// host generates one kernel per pass with the ops, offsets and matrix
// positions as constants (tedq_amd/engine/jit.py) and compiles it here with
// hiprtc; straight-line code lets the compiler rename registers instead of
// copying them (CNOT costs nothing) and keep only the arithmetic.
//
// The generated kernels use the same device templates (sv_ops.h), pass
// arguments and memory layout as the interpreted ones, so results agree to
// rounding.  Compiling and loading a module allocates host memory and device
// code memory; launches allocate nothing.
#include <hip/hiprtc.h>

#include <algorithm>
#include <atomic>
#include <string>
#include <vector>

#include "sv_ops.h"

using namespace tedq;

// src: the generated source; headers: their texts and include names (the
// library's own sv_ops.h and tedq_common.h).  *module_out receives a
// hipModule_t.  Compiled for gfx950 with the library's own flags.
extern "C" int tedq_jit_compile(const char* src, int n_headers, const char* const* header_srcs,
                                const char* const* header_names, void** module_out) {
  if (src == nullptr || module_out == nullptr || n_headers < 0 ||
      (n_headers > 0 && (header_srcs == nullptr || header_names == nullptr))) {
    set_error("tedq_jit_compile: bad arguments");
    return kBadArgument;
  }
  hiprtcProgram prog;
  if (hiprtcCreateProgram(&prog, src, "tedq_pass_kernels.hip", n_headers, const_cast<const char**>(header_srcs),
                          const_cast<const char**>(header_names)) != HIPRTC_SUCCESS) {
    set_error("tedq_jit_compile: hiprtcCreateProgram failed");
    return kLaunchFailed;
  }
  const char* opts[] = {"--offload-arch=gfx950", "-O3", "-std=c++17",
                        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
                        "-mllvm", "-simplifycfg-sink-common=false"};
  const hiprtcResult rc = hiprtcCompileProgram(prog, (int)(sizeof(opts) / sizeof(opts[0])), opts);
  if (rc != HIPRTC_SUCCESS) {
    size_t n = 0;
    hiprtcGetProgramLogSize(prog, &n);
    std::string log(n, '\0');
    if (n) hiprtcGetProgramLog(prog, &log[0]);
    hiprtcDestroyProgram(&prog);
    set_error("tedq_jit_compile: %s: %.900s", hiprtcGetErrorString(rc), log.c_str());
    return kLaunchFailed;
  }
  size_t size = 0;
  hiprtcGetCodeSize(prog, &size);
  std::vector<char> code(size);
  hiprtcGetCode(prog, code.data());
  hiprtcDestroyProgram(&prog);
  hipModule_t mod;
  const hipError_t e = hipModuleLoadData(&mod, code.data());
  if (e != hipSuccess) {
    set_error("tedq_jit_compile: hipModuleLoadData: %s", hipGetErrorString(e));
    return kLaunchFailed;
  }
  *module_out = mod;
  return kOk;
}

// Compute units of the current device, queried once per device.
static int device_cu_count() {
  static std::atomic<int> cache[64];
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 0;
  int c = cache[dev].load(std::memory_order_relaxed);
  if (c > 0) return c;
  if (hipDeviceGetAttribute(&c, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || c <= 0) return 0;
  cache[dev].store(c, std::memory_order_relaxed);
  return c;
}

extern "C" int tedq_jit_function(void* module, const char* name, void** fn_out) {
  if (module == nullptr || name == nullptr || fn_out == nullptr) {
    set_error("tedq_jit_function: bad arguments");
    return kBadArgument;
  }
  hipFunction_t f;
  const hipError_t e = hipModuleGetFunction(&f, static_cast<hipModule_t>(module), name);
  if (e != hipSuccess) {
    set_error("tedq_jit_function: %s: %s", name, hipGetErrorString(e));
    return kBadArgument;
  }
  *fn_out = f;
  return kOk;
}

extern "C" int tedq_jit_function_scratch(void* fn, int* bytes_out) {
  if (fn == nullptr || bytes_out == nullptr) {
    set_error("tedq_jit_function_scratch: bad arguments");
    return kBadArgument;
  }
  int v = 0;
  const hipError_t e = hipFuncGetAttribute(&v, HIP_FUNC_ATTRIBUTE_LOCAL_SIZE_BYTES, static_cast<hipFunction_t>(fn));
  if (e != hipSuccess) {
    set_error("tedq_jit_function_scratch: %s", hipGetErrorString(e));
    return kLaunchFailed;
  }
  *bytes_out = v;
  return kOk;
}

extern "C" int tedq_jit_unload(void* module) {
  if (module == nullptr) return kOk;
  return hipModuleUnload(static_cast<hipModule_t>(module)) == hipSuccess ? kOk : kLaunchFailed;
}

// One generated pass over every circuit (tiles narrower than the circuit,
// n <= 28).  Arguments as tedq_sv_apply_pass / tedq_sv_adjoint_pass; `fn` is
// the generated kernel of this pass (same k, r and tile bits).
extern "C" int tedq_jit_pass(void* fn, int adjoint, void* psi, void* lam, int64_t batch, int n, int k, int r,
                             uint64_t tile_mask, uint64_t zero_mask, const void* mats, int64_t mat_stride,
                             void* partials, int64_t partial_rows, int n_gslots, const double* seed, int flags,
                             void* stream) {
  if (fn == nullptr || psi == nullptr || (adjoint && (lam == nullptr || (n_gslots > 0 && partials == nullptr))) ||
      batch <= 0 || n < 2 || n > 28 || k < 1 || k >= n || r < 1 || r > 5 || r > k || k - r > 9) {
    set_error("tedq_jit_pass: bad arguments (batch=%lld n=%d k=%d r=%d)", (long long)batch, n, k, r);
    return kBadArgument;
  }
  const uint64_t full = (1ull << n) - 1;
  if (__builtin_popcountll(tile_mask & full) != k || (tile_mask & ~full) != 0) {
    set_error("tedq_jit_pass: tile mask does not have k=%d bits below n=%d", k, n);
    return kBadArgument;
  }
  if ((zero_mask & (tile_mask | ~full)) != 0) {
    set_error("tedq_jit_pass: zero mask 0x%llx holds tile bits or bits above n", (unsigned long long)zero_mask);
    return kBadArgument;
  }
  PassArgs a;
  a.mat_stride = mat_stride;
  a.batch = batch;
  a.rest_mask = full & ~tile_mask & ~zero_mask;
  a.total_tiles = batch << __builtin_popcountll(a.rest_mask);
  a.n = n;
  a.k = k;
  a.tT = k - r;
  a.gT = 0;
  a.np = partial_rows;
  a.n_gslots = n_gslots;
  a.flags = flags;
  if ((flags & ~(kPassOutOfPlace | kPassPsiReadOnly | kPassFromZero | kPassPersistent)) != 0 ||
      (!adjoint && (flags & kPassOutOfPlace) && lam == nullptr) || (adjoint && (flags & (kPassOutOfPlace | kPassFromZero))) ||
      (!adjoint && (flags & (kPassPsiReadOnly | kPassPersistent)))) {
    set_error("tedq_jit_pass: flags %d do not fit a %s pass", flags, adjoint ? "adjoint" : "forward");
    return kBadArgument;
  }
  const unsigned T = 1u << (k - r);
  const size_t lds = (size_t)(1u << k) * sizeof(float2) * (adjoint ? 2 : 1);
  if (lds > 160 * 1024 || a.total_tiles > 0x7fffffffll) {
    set_error("tedq_jit_pass: tile of %zu LDS bytes or %lld tiles is out of range", lds, (long long)a.total_tiles);
    return kBadArgument;
  }
  unsigned grid = (unsigned)a.total_tiles;
  if (flags & kPassPersistent) {
    // as many workgroups as the CUs hold at once by LDS (the kernel loops over the rest)
    const int cus = device_cu_count();
    if (cus <= 0) {
      set_error("tedq_jit_pass: cannot query the device's compute units");
      return kLaunchFailed;
    }
    const int64_t resident = (int64_t)cus * (int64_t)std::max<size_t>(1, (160 * 1024) / lds);
    grid = (unsigned)std::min<int64_t>(a.total_tiles, resident);
  }
  void* args[] = {&psi, &lam, const_cast<void**>(&mats), &partials, &a, const_cast<double**>(&seed)};
  const hipError_t e = hipModuleLaunchKernel(static_cast<hipFunction_t>(fn), grid, 1, 1, T, 1, 1,
                                             (unsigned)lds, static_cast<hipStream_t>(stream), args, nullptr);
  if (e != hipSuccess) {
    set_error("tedq_jit_pass: launch failed: %s", hipGetErrorString(e));
    return kLaunchFailed;
  }
  return kOk;
}
