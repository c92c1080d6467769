// Shared device helpers for the tedq MI355X engine (gfx950, wave64).
#pragma once

#if defined(__HIPCC_RTC__)  // hiprtc: the HIP device runtime is built in, the fixed-width types are not
typedef signed char int8_t;
typedef int int32_t;
typedef unsigned int uint32_t;
typedef long int64_t;
typedef unsigned long uint64_t;
#else
#include <hip/hip_runtime.h>
#include <stdint.h>
#endif

namespace tedq {

// ---------------------------------------------------------------------------
// Status handling for the C ABI: every entry point returns 0 or a negative
// code and records a message retrievable with tedq_last_error().
// ---------------------------------------------------------------------------
enum Status : int {
  kOk = 0,
  kBadArgument = -1,
  kUnsupported = -2,
  kLaunchFailed = -3,
};

void set_error(const char* fmt, ...);
int check_launch(const char* what);

// ---------------------------------------------------------------------------
// complex64 arithmetic on float2 (x = real, y = imag)
// ---------------------------------------------------------------------------
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(fmaf(a.x, b.x, -a.y * b.y), fmaf(a.x, b.y, a.y * b.x));
}
// a*x + b*y
__device__ __forceinline__ float2 cmul2(float2 a, float2 x, float2 b, float2 y) {
  float re = fmaf(a.x, x.x, fmaf(-a.y, x.y, fmaf(b.x, y.x, -b.y * y.y)));
  float im = fmaf(a.x, x.y, fmaf(a.y, x.x, fmaf(b.x, y.y, b.y * y.x)));
  return make_float2(re, im);
}
__device__ __forceinline__ float2 cconj(float2 a) { return make_float2(a.x, -a.y); }
__device__ __forceinline__ float2 cneg(float2 a) { return make_float2(-a.x, -a.y); }

// ---------------------------------------------------------------------------
// complex128 arithmetic for the fp64 table builder
// ---------------------------------------------------------------------------
struct cd {
  double x, y;
};
__device__ __forceinline__ cd mkcd(double x, double y) { return cd{x, y}; }
__device__ __forceinline__ cd cd_mul(cd a, cd b) { return cd{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x}; }
__device__ __forceinline__ cd cd_add(cd a, cd b) { return cd{a.x + b.x, a.y + b.y}; }
__device__ __forceinline__ cd cd_conj(cd a) { return cd{a.x, -a.y}; }

// ---------------------------------------------------------------------------
// LDS swizzle: XOR the low 4 bits of the element index (16 complex64 = 128 B
// = one pass of 16 lanes over the 32 LDS banks) with a linear map of bits
// 4..7 and with bits 8..11, so that the gathers of a register stage (lanes
// differing only in bits above the register bits) hit 16 distinct 8-byte
// slots of a bank row.  Bits 4..7 map through x -> (x & 6) | (x0 ^ x2) |
// parity(x) << 3 (the columns 9, 10, 13, 8): in a model of 16-lane passes over
// 32 banks, which reproduces the measured conflict rates, it removes the
// conflicts left by the identity map (15% / 19% of the forward / adjoint LDS
// cycles at 12-bit tiles) on every tile shape tried.  An XOR-linear involution.
// ---------------------------------------------------------------------------
__device__ __forceinline__ int swz(int e) {
  const int x = (e >> 4) & 15;
  const int lo = (x & 6) | ((x ^ (x >> 2)) & 1) | ((__builtin_popcount(x) & 1) << 3);
  return e ^ (lo ^ ((e >> 8) & 15));
}

// Deposit the low bits of v into the set bits of mask (ascending), like x86 pdep.
__device__ __forceinline__ uint64_t pdep64(uint64_t v, uint64_t mask) {
  uint64_t out = 0;
  while (mask) {
    uint64_t low = mask & (~mask + 1);
    if (v & 1) out |= low;
    v >>= 1;
    mask &= mask - 1;
  }
  return out;
}

}  // namespace tedq
