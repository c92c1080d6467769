// Device pieces shared by the pass kernels compiled into the library
// (csrc/sv_passes.hip) and the per-circuit pass kernels generated at run time
// (tedq_amd/engine/jit.py, compiled with hiprtc): op codes, pass arguments,
// register-level gate templates, buffer access, tile placement and the wave
// reduction of gradient bilinears.
#pragma once

#include "tedq_common.h"

namespace tedq {

enum OpCode : int {
  kOpU1 = 1, kOpD1 = 2, kOpCU1 = 3, kOpCX = 4, kOpCZ = 5,  // 6: unused
  kOpGrad1 = 7, kOpCGrad1 = 8, kOpCD1 = 9,
  // real 2x2 (imaginary parts of the table entries are zero), optionally controlled
  kOpR1 = 10,
  // one gradient value of a one-parameter rotation whose derivative is dU U^dagger = i * (Pauli or
  // projector) / const: words {11, bit, 0, cmask, 0, -1, kind | negate << 4, component}; the value is
  // added to component `component` of the pending gradient group
  kOpPauliGrad = 11,
  // write the pending gradient group (8 values, reduced over the tile) to slot word[5] and clear it
  kOpGradFlush = 12,
  // amplitudes whose bit is 1 (and all control bits) times the slot's entry 1 (its conjugate with the
  // dagger flag): a diagonal gate without its global phase (the slot table stores p = d1 conj(d0) there)
  kOpPhase1 = 13,
};

// kinds of kOpPauliGrad: the value equals 2 Re sum_ab A_ab C_ab for A = dU U^dagger of RX (X), RY (Y),
// RZ (Z) and PhaseShift (P), with C_ab = sum over amplitude pairs of conj(lambda_a) psi_b
enum PauliGradKind : int { kGradX = 0, kGradY = 1, kGradZ = 2, kGradP = 3 };

constexpr int kHdrTileBits = 4;
constexpr int kHdrHiOff = 36;
constexpr int kHdrStages = 100;
constexpr int kStageWords = 48;

struct PassArgs {
  int64_t mat_stride;
  int64_t batch;
  int64_t total_tiles;      // batch * 2^popcount(rest_mask)
  uint64_t rest_mask;       // physical bits (< n) that are not tile bits and over which tiles run
  int n;
  int k;
  int tT;                   // log2(threads per tile)
  int gT;                   // log2(tiles per workgroup)
  int64_t np;               // adjoint: partials are [batch][n_gslots][np][8] (partial rows per slot per circuit)
  int n_gslots;
  int flags;                // generated passes: kPassOutOfPlace, kPassPsiReadOnly, kPassFromZero,
                            // kPassPersistent
};

// Generated forward pass: read psi, write the result to the second state pointer (a checkpoint).
constexpr int kPassOutOfPlace = 1;
// Generated adjoint pass: psi holds the forward state at the end of this pass (a checkpoint) and is only
// read; the state at its start is the previous checkpoint, so psi is not written back.
constexpr int kPassPsiReadOnly = 2;
// Generated forward pass of a sweep from |0..0> (with the zero mask of tile skipping): tile bits that no
// earlier pass held are 0 wherever the state is nonzero, so those amplitudes are taken as zero instead of
// loaded (the first pass, whose tile bits are all new, builds the |0..0> tile without reading psi).
constexpr int kPassFromZero = 4;
// Generated adjoint pass whose workgroups loop over tiles (blockIdx.x, + gridDim.x, ...), loading the next
// tile into registers while processing the current one: launched with at most as many workgroups as the
// device holds at once.
constexpr int kPassPersistent = 8;

// ---------------------------------------------------------------------------
// register-level gate kernels.  B is the target register bit and CM a control
// mask over register bits, both compile-time, so every pair selection below
// folds away: an op costs only its arithmetic (no per-pair scalar tests).
// Ops with several control bits (rare: Toffoli) use the *_rt runtime forms.
// ---------------------------------------------------------------------------
template <int R, int B, int CM>
struct U1Op {  // general 2x2 on bit B where all CM bits are 1
  static __device__ __forceinline__ void run(float2 (&v)[R], float2 m0, float2 m1, float2 m2, float2 m3) {
#pragma unroll
    for (int i = 0; i < R; ++i) {
      if ((i & (1 << B)) || (i & CM) != CM) continue;
      const int j = i | (1 << B);
      const float2 x = v[i], y = v[j];
      v[i] = cmul2(m0, x, m1, y);
      v[j] = cmul2(m2, x, m3, y);
    }
  }
};

template <int R, int B, int CM>
struct DiagOp {  // diag(d0, d1) on bit B where all CM bits are 1
  static __device__ __forceinline__ void run(float2 (&v)[R], float2 d0, float2 d1) {
#pragma unroll
    for (int i = 0; i < R; ++i) {
      if ((i & CM) != CM) continue;
      v[i] = cmul((i >> B) & 1 ? d1 : d0, v[i]);
    }
  }
};

template <int R, int B, int CM>
struct CXOp {  // X on bit B where all CM bits are 1
  static __device__ __forceinline__ void run(float2 (&v)[R]) {
#pragma unroll
    for (int i = 0; i < R; ++i) {
      if ((i & (1 << B)) || (i & CM) != CM) continue;
      const int j = i | (1 << B);
      const float2 t = v[i];
      v[i] = v[j];
      v[j] = t;
    }
  }
};

template <int R, int B, int CM>
struct CZOp {  // -1 where bit B and all CM bits are 1
  static __device__ __forceinline__ void run(float2 (&v)[R]) {
#pragma unroll
    for (int i = 0; i < R; ++i)
      if ((i & (1 << B)) && (i & CM) == CM) v[i] = cneg(v[i]);
  }
};

template <int R, int B, int CM>
struct R1Op {  // real 2x2 on bit B where all CM bits are 1 (half the arithmetic of U1Op)
  static __device__ __forceinline__ void run(float2 (&v)[R], float m0, float m1, float m2, float m3) {
#pragma unroll
    for (int i = 0; i < R; ++i) {
      if ((i & (1 << B)) || (i & CM) != CM) continue;
      const int j = i | (1 << B);
      const float2 x = v[i], y = v[j];
      v[i] = make_float2(fmaf(m0, x.x, m1 * y.x), fmaf(m0, x.y, m1 * y.y));
      v[j] = make_float2(fmaf(m2, x.x, m3 * y.x), fmaf(m2, x.y, m3 * y.y));
    }
  }
};

// Gradient value of a one-parameter rotation (see PauliGradKind) over the pairs of bit B where all CM
// bits are 1; NEG flips the sign (a daggered gate).  X: Im C01 + Im C10, Y: Re C10 - Re C01,
// Z: Im C00 - Im C11, P: -2 Im C11.  Re conj(a) b = a.x b.x + a.y b.y, Im conj(a) b = a.x b.y - a.y b.x.
template <int R, int B, int CM, int KIND, bool NEG>
struct PauliGradOp {
  static __device__ __forceinline__ void run(const float2 (&p)[R], const float2 (&l)[R], float& acc) {
    const float s = NEG ? -1.f : 1.f;
#pragma unroll
    for (int i = 0; i < R; ++i) {
      if ((i & (1 << B)) || (i & CM) != CM) continue;
      const int j = i | (1 << B);
      if (KIND == kGradX) {
        acc = fmaf(s * l[i].x, p[j].y, fmaf(-s * l[i].y, p[j].x, acc));
        acc = fmaf(s * l[j].x, p[i].y, fmaf(-s * l[j].y, p[i].x, acc));
      } else if (KIND == kGradY) {
        acc = fmaf(s * l[j].x, p[i].x, fmaf(s * l[j].y, p[i].y, acc));
        acc = fmaf(-s * l[i].x, p[j].x, fmaf(-s * l[i].y, p[j].y, acc));
      } else if (KIND == kGradZ) {
        acc = fmaf(s * l[i].x, p[i].y, fmaf(-s * l[i].y, p[i].x, acc));
        acc = fmaf(-s * l[j].x, p[j].y, fmaf(s * l[j].y, p[j].x, acc));
      } else {
        acc = fmaf(-2.f * s * l[j].x, p[j].y, fmaf(2.f * s * l[j].y, p[j].x, acc));
      }
    }
  }
};

// bilinears c_ab = sum conj(l_a) p_b over the pairs of bit B where all CM bits are 1
template <int R, int B, int CM>
struct GradOp {
  static __device__ __forceinline__ void run(const float2 (&p)[R], const float2 (&l)[R], float (&acc)[8]) {
#pragma unroll
    for (int i = 0; i < R; ++i) {
      if ((i & (1 << B)) || (i & CM) != CM) continue;
      const int j = i | (1 << B);
      acc[0] = fmaf(l[i].x, p[i].x, fmaf(l[i].y, p[i].y, acc[0]));
      acc[1] = fmaf(l[i].x, p[i].y, fmaf(-l[i].y, p[i].x, acc[1]));
      acc[2] = fmaf(l[i].x, p[j].x, fmaf(l[i].y, p[j].y, acc[2]));
      acc[3] = fmaf(l[i].x, p[j].y, fmaf(-l[i].y, p[j].x, acc[3]));
      acc[4] = fmaf(l[j].x, p[i].x, fmaf(l[j].y, p[i].y, acc[4]));
      acc[5] = fmaf(l[j].x, p[i].y, fmaf(-l[j].y, p[i].x, acc[5]));
      acc[6] = fmaf(l[j].x, p[j].x, fmaf(l[j].y, p[j].y, acc[6]));
      acc[7] = fmaf(l[j].x, p[j].y, fmaf(-l[j].y, p[j].x, acc[7]));
    }
  }
};

// ---------------------------------------------------------------------------
// buffer access: the resource (tile base, uniform per workgroup) lives in
// SGPRs, each lane adds a 32-bit offset and each register row a scalar offset
// ---------------------------------------------------------------------------
typedef int tedq_v2i __attribute__((ext_vector_type(2)));

__device__ __forceinline__ __amdgpu_buffer_rsrc_t make_rsrc(const void* base, uint32_t bytes) {
  // raw buffer, stride 0, num_records in bytes; 0x00020000 is the gfx9 data-format word
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, (int)bytes, 0x00020000);
}
__device__ __forceinline__ float2 bload(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff) {
  return __builtin_bit_cast(float2, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0));
}
__device__ __forceinline__ void bstore(float2 v, __amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff) {
  __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(tedq_v2i, v), r, voff, soff, 0);
}

// Circuit index and global element offset of tile `tile` of a pass whose
// tiles are narrower than the circuit.  Workgroups are dealt round-robin to
// the 8 XCDs (each with its own L2).  When the tile's contiguous runs are
// shorter than a 128-byte line, the tiles differing in the lowest rest bit
// share lines; swapping that bit with rest bit 3 makes those partners
// blockIdx b and b^8, which land on the same XCD at about the same time, so
// the second read can hit L2.
// Tiles of a circuit are numbered over the bits of rest_mask (2^popcount of them: every non-tile bit,
// or fewer when the caller leaves out bits on which the state vanishes).
__device__ __forceinline__ int tiles_log2(const PassArgs& a) { return __popcll(a.rest_mask); }

__device__ __forceinline__ void tile_origin(const PassArgs& a, int64_t tile, int64_t* circuit, int64_t* gbase) {
  const int64_t tpc_log = tiles_log2(a);
  *circuit = tile >> tpc_log;
  uint64_t within = (uint64_t)(tile & ((1ll << tpc_log) - 1));
  if (tpc_log >= 4) within = (within & ~9ull) | ((within & 1ull) << 3) | ((within >> 3) & 1ull);
  *gbase = (*circuit << a.n) + (int64_t)pdep64(within, a.rest_mask);
}

// swizzled LDS index of this thread's register 0 in a stage
__device__ __forceinline__ int stage_base(int tid, int tT, int k, int nonreg) {
  const int g = tid >> tT;
  int tl = tid & ((1 << tT) - 1);
  int base = 0;
  int mask = nonreg;
  while (mask) {
    const int low = mask & (-mask);
    if (tl & 1) base |= low;
    tl >>= 1;
    mask &= mask - 1;
  }
  return swz((g << k) + base);
}

// Cross-lane moves without LDS: DPP within rows of 16 lanes, and the gfx950
// permlane swaps across rows (16) and halves (32) of the wave.
template <int CTRL>
__device__ __forceinline__ float dpp_mov(float x) {
  return __builtin_bit_cast(float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, x), CTRL, 0xf, 0xf, false));
}

// The two results of a permlane swap must stay distinct values: this compiler
// folds `s[0] + s[1]` of __builtin_amdgcn_permlane{16,32}_swap into s[0] + s[0]
// (checked in the generated ISA).  An empty asm that takes both values as
// read-write VGPR operands hides them from that fold; it emits no instruction.
__device__ __forceinline__ void opaque2(unsigned& a, unsigned& b) { asm volatile("" : "+v"(a), "+v"(b)); }

// Sum the 8 bilinears over the 64 lanes of a wave by recursive halving.  A
// permlane32 swap exchanges the upper half of one value with the lower half
// of another, so one add per pair leaves value j on lanes 0..31 and value j+4
// on lanes 32..63; permlane16 does the same across 16-lane rows, and a DPP
// rotate by 8 within rows finishes lane bit 3.  Then lane L holds value
// (L >> 3) & 7 summed over lane bits 3..5, and DPP quad permutes plus a
// half-row mirror sum lane bits 0..2.  All moves are VALU (no LDS round
// trips).  Returns the total for value (lane >> 3) & 7, valid on every lane;
// the wave must be fully active.
__device__ __forceinline__ float reduce8_wave(float (&v)[8], int lane) {
  float h4[4], h2[2];
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const auto s = __builtin_amdgcn_permlane32_swap(__builtin_bit_cast(unsigned, v[j]),
                                                    __builtin_bit_cast(unsigned, v[4 + j]), false, false);
    unsigned x0 = s[0], x1 = s[1];
    opaque2(x0, x1);
    h4[j] = __builtin_bit_cast(float, x0) + __builtin_bit_cast(float, x1);
  }
#pragma unroll
  for (int j = 0; j < 2; ++j) {
    const auto s = __builtin_amdgcn_permlane16_swap(__builtin_bit_cast(unsigned, h4[j]),
                                                    __builtin_bit_cast(unsigned, h4[2 + j]), false, false);
    unsigned x0 = s[0], x1 = s[1];
    opaque2(x0, x1);
    h2[j] = __builtin_bit_cast(float, x0) + __builtin_bit_cast(float, x1);
  }
  const bool up3 = lane & 8;
  float x = (up3 ? h2[1] : h2[0]) + dpp_mov<0x128>(up3 ? h2[0] : h2[1]);  // row_ror:8
  x += dpp_mov<0xB1>(x);   // quad_perm [1,0,3,2]
  x += dpp_mov<0x4E>(x);   // quad_perm [2,3,0,1]
  x += dpp_mov<0x141>(x);  // row_half_mirror: the other quad of the 8-lane group
  return x;
}

}  // namespace tedq
