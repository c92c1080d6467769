// Fused state-vector passes: forward gate application and the adjoint
// (reverse) sweep with gradient bilinears.
//
// Replaces the reference SV hot loop (tedq/backends/pytorch_backend.py:358-380:
// one torch.tensordot + permute per gate, i.e. about two HBM round trips of
// the whole state per gate) and torch autograd's backward through it.
//
// One launch = one pass.  Each workgroup owns G tiles of 2^k amplitudes
// (G = 1 unless a whole circuit fits in a tile); a tile is selected by the
// physical bit set of the pass (low bits always included, so every wave
// reads contiguous runs).  The tile is staged in LDS; each *stage* gathers R
// amplitudes per thread into registers along the stage's register bits,
// applies all gates of the stage in registers, and scatters back.
//
// Pass program (int32 words, built by tedq_amd/engine/program.py):
//   [0] k  [1] r  [2] nstages  [3] tT = log2(threads per tile) = k - r
//   [4..36)   physical position of tile bit i (ascending)
//   [36..100) 64-bit physical offset of register-row m in the load layout
//   [100 + 48 s ...] stage s: [0] nops [1] op word offset [2] non-register
//             local-bit mask [3] unused [4..36) local offset of register m
//   ops: 8 words {type | dagger<<8, b0, b1, cmask, mat_off, gslot, 0, 0}
#include "tedq_common.h"

namespace tedq {

enum OpCode : int {
  kOpU1 = 1, kOpD1 = 2, kOpCU1 = 3, kOpCX = 4, kOpCZ = 5, kOpU2 = 6,
  kOpGrad1 = 7, kOpCGrad1 = 8, kOpCD1 = 9,
};

constexpr int kHdrTileBits = 4;
constexpr int kHdrHiOff = 36;
constexpr int kHdrStages = 100;
constexpr int kStageWords = 48;

struct PassArgs {
  const int* prog;
  const float2* mats;
  int64_t mat_stride;
  int64_t batch;
  int64_t total_tiles;      // batch * 2^(n-k)
  uint64_t rest_mask;       // physical bits (< n) that are not tile bits
  int n;
  int k;
  int tT;                   // log2(threads per tile)
  int gT;                   // log2(tiles per workgroup)
  float* partials;          // adjoint only: [batch][n_gslots][np][8]
  int64_t np;               // partial rows per gradient slot per circuit
  int n_gslots;
};

// ---------------------------------------------------------------------------
// register-level gate kernels; B, B0, B1 are register-bit indices
// ---------------------------------------------------------------------------
template <int R, int B>
__device__ __forceinline__ void u1_bit(float2 (&v)[R], float2 m0, float2 m1, float2 m2, float2 m3, int cmask) {
#pragma unroll
  for (int i = 0; i < R; ++i) {
    if (i & (1 << B)) continue;
    if ((i & cmask) != cmask) continue;
    const int j = i | (1 << B);
    const float2 x = v[i], y = v[j];
    v[i] = cmul2(m0, x, m1, y);
    v[j] = cmul2(m2, x, m3, y);
  }
}

template <int R, int B>
__device__ __forceinline__ void cx_bit(float2 (&v)[R], int cmask) {
#pragma unroll
  for (int i = 0; i < R; ++i) {
    if (i & (1 << B)) continue;
    if ((i & cmask) != cmask) continue;
    const int j = i | (1 << B);
    const float2 t = v[i];
    v[i] = v[j];
    v[j] = t;
  }
}

template <int R, int B0, int B1>
__device__ __forceinline__ void u2_bits(float2 (&v)[R], const float2* m, bool dag) {
#pragma unroll
  for (int i = 0; i < R; ++i) {
    if (i & ((1 << B0) | (1 << B1))) continue;
    const int i0 = i, i1 = i | (1 << B1), i2 = i | (1 << B0), i3 = i | (1 << B0) | (1 << B1);
    const float2 x0 = v[i0], x1 = v[i1], x2 = v[i2], x3 = v[i3];
    float2 y[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      // row r of M (or of M^dagger: conj of column r)
      const float2 a0 = dag ? cconj(m[0 * 4 + r]) : m[r * 4 + 0];
      const float2 a1 = dag ? cconj(m[1 * 4 + r]) : m[r * 4 + 1];
      const float2 a2 = dag ? cconj(m[2 * 4 + r]) : m[r * 4 + 2];
      const float2 a3 = dag ? cconj(m[3 * 4 + r]) : m[r * 4 + 3];
      const float2 s0 = cmul2(a0, x0, a1, x1);
      const float2 s1 = cmul2(a2, x2, a3, x3);
      y[r] = make_float2(s0.x + s1.x, s0.y + s1.y);
    }
    v[i0] = y[0];
    v[i1] = y[1];
    v[i2] = y[2];
    v[i3] = y[3];
  }
}

// bilinears c_ab = sum conj(l_a) p_b over the pairs of register bit B
template <int R, int B>
__device__ __forceinline__ void grad_bit(const float2 (&p)[R], const float2 (&l)[R], int cmask, float (&acc)[8]) {
#pragma unroll
  for (int i = 0; i < R; ++i) {
    if (i & (1 << B)) continue;
    if ((i & cmask) != cmask) continue;
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

#define TEDQ_BIT_SWITCH(b, CALL)                        \
  switch (b) {                                          \
    case 0: if constexpr (R > 1) { CALL(0); } break;    \
    case 1: if constexpr (R > 2) { CALL(1); } break;    \
    case 2: if constexpr (R > 4) { CALL(2); } break;    \
    case 3: if constexpr (R > 8) { CALL(3); } break;    \
    case 4: if constexpr (R > 16) { CALL(4); } break;   \
    default: break;                                     \
  }

template <int R>
__device__ __forceinline__ void op_u1(float2 (&v)[R], int b, int cmask, float2 m0, float2 m1, float2 m2, float2 m3) {
#define CALL(B) u1_bit<R, B>(v, m0, m1, m2, m3, cmask)
  TEDQ_BIT_SWITCH(b, CALL)
#undef CALL
}

template <int R>
__device__ __forceinline__ void op_cx(float2 (&v)[R], int b, int cmask) {
#define CALL(B) cx_bit<R, B>(v, cmask)
  TEDQ_BIT_SWITCH(b, CALL)
#undef CALL
}

template <int R>
__device__ __forceinline__ void op_grad(const float2 (&p)[R], const float2 (&l)[R], int b, int cmask,
                                        float (&acc)[8]) {
#define CALL(B) grad_bit<R, B>(p, l, cmask, acc)
  TEDQ_BIT_SWITCH(b, CALL)
#undef CALL
}

template <int R, int B0>
__device__ __forceinline__ void op_u2_b1(float2 (&v)[R], int b1, const float2* m, bool dag) {
#define CALL(B1) if constexpr (B1 != B0) u2_bits<R, B0, B1>(v, m, dag)
  TEDQ_BIT_SWITCH(b1, CALL)
#undef CALL
}

template <int R>
__device__ __forceinline__ void op_u2(float2 (&v)[R], int b0, int b1, const float2* m, bool dag) {
#define CALL(B0) op_u2_b1<R, B0>(v, b1, m, dag)
  TEDQ_BIT_SWITCH(b0, CALL)
#undef CALL
}

template <int R>
__device__ __forceinline__ void op_diag(float2 (&v)[R], int b, int cmask, float2 d0, float2 d1) {
#pragma unroll
  for (int i = 0; i < R; ++i) {
    if ((i & cmask) != cmask) continue;
    v[i] = cmul(((i >> b) & 1) ? d1 : d0, v[i]);
  }
}

template <int R>
__device__ __forceinline__ void op_cz(float2 (&v)[R], int mask) {
#pragma unroll
  for (int i = 0; i < R; ++i)
    if ((i & mask) == mask) v[i] = cneg(v[i]);
}

// Apply one unitary op (optionally daggered) to one register set, or to two
// (psi and lambda in the adjoint sweep) sharing the matrix load.
template <int R, bool TWO>
__device__ __forceinline__ void apply_op(float2 (&v)[R], float2 (&w)[R], const int* op, const float2* mb) {
  const int type = op[0] & 0xff;
  const bool dag = (op[0] >> 8) & 1;
  switch (type) {
    case kOpU1:
    case kOpCU1: {
      const float2* m = mb + op[4];
      float2 m0 = m[0], m1 = m[1], m2 = m[2], m3 = m[3];
      if (dag) {
        const float2 t = m1;
        m0 = cconj(m0);
        m1 = cconj(m2);
        m2 = cconj(t);
        m3 = cconj(m3);
      }
      op_u1<R>(v, op[1], op[3], m0, m1, m2, m3);
      if (TWO) op_u1<R>(w, op[1], op[3], m0, m1, m2, m3);
      break;
    }
    case kOpD1:
    case kOpCD1: {
      const float2* m = mb + op[4];
      float2 d0 = m[0], d1 = m[3];
      if (dag) {
        d0 = cconj(d0);
        d1 = cconj(d1);
      }
      op_diag<R>(v, op[1], op[3], d0, d1);
      if (TWO) op_diag<R>(w, op[1], op[3], d0, d1);
      break;
    }
    case kOpCX:
      op_cx<R>(v, op[1], op[3]);
      if (TWO) op_cx<R>(w, op[1], op[3]);
      break;
    case kOpCZ:
      op_cz<R>(v, op[3]);
      if (TWO) op_cz<R>(w, op[3]);
      break;
    case kOpU2:
      op_u2<R>(v, op[1], op[2], mb + op[4], dag);
      if (TWO) op_u2<R>(w, op[1], op[2], mb + op[4], dag);
      break;
    default: break;
  }
}

// ---------------------------------------------------------------------------
// tile addressing
// ---------------------------------------------------------------------------
struct TileCtx {
  int64_t circuit;     // circuit index of this thread's tile
  int64_t gbase;       // global element offset of local index 0 (G == 1 layout)
  uint64_t lo;         // physical offset contributed by the thread-id bits
  bool valid;
};

template <int R, bool MULTI>
__device__ __forceinline__ void load_tiles(const float2* __restrict__ src, float2* lds, const PassArgs& a,
                                           int tid, int T, const TileCtx& c) {
  if (MULTI) {
    const int64_t wg_base = (int64_t)blockIdx.x * T * R;
    const int64_t limit = a.batch << a.n;
#pragma unroll
    for (int m = 0; m < R; ++m) {
      const int e = tid + m * T;
      const int64_t gi = wg_base + e;
      lds[swz(e)] = gi < limit ? src[gi] : make_float2(0.f, 0.f);
    }
  } else {
    const int64_t* hi = reinterpret_cast<const int64_t*>(a.prog + kHdrHiOff);
#pragma unroll
    for (int m = 0; m < R; ++m) lds[swz(tid + m * T)] = src[c.gbase + c.lo + hi[m]];
  }
}

template <int R, bool MULTI>
__device__ __forceinline__ void store_tiles(float2* __restrict__ dst, const float2* lds, const PassArgs& a,
                                            int tid, int T, const TileCtx& c) {
  if (MULTI) {
    const int64_t wg_base = (int64_t)blockIdx.x * T * R;
    const int64_t limit = a.batch << a.n;
#pragma unroll
    for (int m = 0; m < R; ++m) {
      const int e = tid + m * T;
      const int64_t gi = wg_base + e;
      if (gi < limit) dst[gi] = lds[swz(e)];
    }
  } else {
    const int64_t* hi = reinterpret_cast<const int64_t*>(a.prog + kHdrHiOff);
#pragma unroll
    for (int m = 0; m < R; ++m) dst[c.gbase + c.lo + hi[m]] = lds[swz(tid + m * T)];
  }
}

__device__ __forceinline__ TileCtx make_ctx(const PassArgs& a, int tid, bool multi) {
  TileCtx c;
  if (multi) {
    const int64_t tile = ((int64_t)blockIdx.x << a.gT) + (tid >> a.tT);
    c.circuit = tile;  // whole-circuit tiles
    c.valid = tile < a.batch;
    c.gbase = 0;
    c.lo = 0;
  } else {
    const int64_t tpc_log = a.n - a.k;
    const int64_t tile = blockIdx.x;
    c.circuit = tile >> tpc_log;
    c.valid = true;
    const uint64_t within = (uint64_t)(tile & ((1ll << tpc_log) - 1));
    c.gbase = (c.circuit << a.n) + (int64_t)pdep64(within, a.rest_mask);
    uint64_t lo = 0;
    for (int i = 0; i < a.tT; ++i)
      if ((tid >> i) & 1) lo |= 1ull << a.prog[kHdrTileBits + i];
    c.lo = lo;
  }
  return c;
}

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
  return (g << k) + base;
}

// ---------------------------------------------------------------------------
// forward pass
// ---------------------------------------------------------------------------
template <int R, bool MULTI>
__global__ void __launch_bounds__(256) k_pass_fwd(float2* __restrict__ state, PassArgs a) {
  extern __shared__ float2 lds[];
  const int tid = threadIdx.x;
  const int T = blockDim.x;
  const TileCtx c = make_ctx(a, tid, MULTI);
  const float2* mb = a.mats + (c.valid ? c.circuit : 0) * a.mat_stride;
  load_tiles<R, MULTI>(state, lds, a, tid, T, c);
  __syncthreads();
  const int nst = a.prog[2];
  for (int s = 0; s < nst; ++s) {
    const int* st = a.prog + kHdrStages + kStageWords * s;
    const int base = stage_base(tid, a.tT, a.k, st[2]);
    float2 v[R];
#pragma unroll
    for (int m = 0; m < R; ++m) v[m] = lds[swz(base + st[4 + m])];
    const int nops = st[0];
    const int* op = a.prog + st[1];
    for (int o = 0; o < nops; ++o, op += 8) apply_op<R, false>(v, v, op, mb);
#pragma unroll
    for (int m = 0; m < R; ++m) lds[swz(base + st[4 + m])] = v[m];
    __syncthreads();
  }
  store_tiles<R, MULTI>(state, lds, a, tid, T, c);
}

// ---------------------------------------------------------------------------
// adjoint pass: walks the (already reversed and daggered) op list over psi and
// lambda; gradient ops reduce their 8 bilinears over the threads of a tile.
// ---------------------------------------------------------------------------
template <int R, bool MULTI>
__global__ void __launch_bounds__(256) k_pass_adj(float2* __restrict__ psi, float2* __restrict__ lam, PassArgs a) {
  extern __shared__ float2 lds[];
  float2* ldp = lds;
  const int T = blockDim.x;
  float2* ldl = lds + T * R;
  const int tid = threadIdx.x;
  const TileCtx c = make_ctx(a, tid, MULTI);
  const float2* mb = a.mats + (c.valid ? c.circuit : 0) * a.mat_stride;
  load_tiles<R, MULTI>(psi, ldp, a, tid, T, c);
  load_tiles<R, MULTI>(lam, ldl, a, tid, T, c);
  __syncthreads();

  // partial rows: W waves per tile when a tile spans several waves
  const int tpt = 1 << a.tT;                 // threads per tile
  const int seg = tpt < 64 ? tpt : 64;       // lanes reduced together
  const int W = tpt >> 6 > 0 ? tpt >> 6 : 1; // partial rows per tile
  const int tl = tid & (tpt - 1);
  int64_t tile_in_circuit = MULTI ? 0 : ((int64_t)blockIdx.x & ((1ll << (a.n - a.k)) - 1));
  const int64_t prow = tile_in_circuit * W + (tl >> 6);
  const bool writer = c.valid && ((tid & (seg - 1)) == 0);

  const int nst = a.prog[2];
  for (int s = 0; s < nst; ++s) {
    const int* st = a.prog + kHdrStages + kStageWords * s;
    const int base = stage_base(tid, a.tT, a.k, st[2]);
    float2 vp[R], vl[R];
#pragma unroll
    for (int m = 0; m < R; ++m) {
      vp[m] = ldp[swz(base + st[4 + m])];
      vl[m] = ldl[swz(base + st[4 + m])];
    }
    const int nops = st[0];
    const int* op = a.prog + st[1];
    for (int o = 0; o < nops; ++o, op += 8) {
      const int type = op[0] & 0xff;
      if (type == kOpGrad1 || type == kOpCGrad1) {
        float acc[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
        op_grad<R>(vp, vl, op[1], op[3], acc);
#pragma unroll
        for (int q = 0; q < 8; ++q)
          for (int off = 1; off < seg; off <<= 1) acc[q] += __shfl_xor(acc[q], off, 64);
        if (writer) {
          float4* dst = reinterpret_cast<float4*>(
              a.partials + (((c.circuit * a.n_gslots + op[5]) * a.np) + prow) * 8);
          dst[0] = make_float4(acc[0], acc[1], acc[2], acc[3]);
          dst[1] = make_float4(acc[4], acc[5], acc[6], acc[7]);
        }
      } else {
        apply_op<R, true>(vp, vl, op, mb);
      }
    }
#pragma unroll
    for (int m = 0; m < R; ++m) {
      ldp[swz(base + st[4 + m])] = vp[m];
      ldl[swz(base + st[4 + m])] = vl[m];
    }
    __syncthreads();
  }
  store_tiles<R, MULTI>(psi, ldp, a, tid, T, c);
  store_tiles<R, MULTI>(lam, ldl, a, tid, T, c);
}

template <int R>
static int launch_pass(bool adjoint, bool multi, dim3 grid, dim3 block, size_t lds, hipStream_t st, float2* psi,
                       float2* lam, const PassArgs& a) {
  if (adjoint) {
    if (multi)
      hipLaunchKernelGGL((k_pass_adj<R, true>), grid, block, lds, st, psi, lam, a);
    else
      hipLaunchKernelGGL((k_pass_adj<R, false>), grid, block, lds, st, psi, lam, a);
  } else {
    if (multi)
      hipLaunchKernelGGL((k_pass_fwd<R, true>), grid, block, lds, st, psi, a);
    else
      hipLaunchKernelGGL((k_pass_fwd<R, false>), grid, block, lds, st, psi, a);
  }
  return check_launch(adjoint ? "tedq_sv_adjoint_pass" : "tedq_sv_apply_pass");
}

static int run_pass(bool adjoint, void* psi, void* lam, int64_t batch, int n, int k, int r, int threads_log2,
                    const int* dev_prog, uint64_t tile_mask, const void* mats, int64_t mat_stride,
                    float* partials, int64_t np, int n_gslots, void* stream) {
  if (psi == nullptr || (adjoint && (lam == nullptr || (n_gslots > 0 && partials == nullptr))) ||
      dev_prog == nullptr || batch <= 0 || n < 1 || n > 40 || k < 1 || k > n || r < 1 || r > 5 || r > k) {
    set_error("sv pass: bad arguments (batch=%lld n=%d k=%d r=%d)", (long long)batch, n, k, r);
    return kBadArgument;
  }
  const int tT = k - r;
  if (threads_log2 < tT || threads_log2 > 8 || (threads_log2 > tT && k != n)) {
    set_error("sv pass: threads_log2=%d incompatible with k=%d r=%d n=%d", threads_log2, k, r, n);
    return kBadArgument;
  }
  const bool multi = (k == n);
  const uint64_t full = (n >= 64) ? ~0ull : ((1ull << n) - 1);
  if (__builtin_popcountll(tile_mask & full) != k || (tile_mask & ~full) != 0) {
    set_error("sv pass: tile mask does not have k=%d bits below n=%d", k, n);
    return kBadArgument;
  }
  PassArgs a;
  a.prog = dev_prog;
  a.mats = static_cast<const float2*>(mats);
  a.mat_stride = mat_stride;
  a.batch = batch;
  a.total_tiles = batch << (n - k);
  a.rest_mask = full & ~tile_mask;
  a.n = n;
  a.k = k;
  a.tT = tT;
  a.gT = threads_log2 - tT;
  a.partials = partials;
  a.np = np;
  a.n_gslots = n_gslots;
  const int T = 1 << threads_log2;
  const int R = 1 << r;
  int64_t blocks;
  if (multi) {
    const int64_t tiles_per_wg = 1ll << a.gT;
    blocks = (batch + tiles_per_wg - 1) / tiles_per_wg;
  } else {
    blocks = a.total_tiles;
  }
  if (blocks > 0x7fffffffll) {
    set_error("sv pass: grid too large (%lld blocks)", (long long)blocks);
    return kBadArgument;
  }
  const size_t lds = (size_t)T * R * sizeof(float2) * (adjoint ? 2 : 1);
  if (lds > 65536) {
    set_error("sv pass: LDS request %zu bytes exceeds 64 KiB", lds);
    return kBadArgument;
  }
  float2* p = static_cast<float2*>(psi);
  float2* l = static_cast<float2*>(lam);
  hipStream_t st = static_cast<hipStream_t>(stream);
  dim3 grid((unsigned)blocks), block((unsigned)T);
  switch (r) {
    case 1: return launch_pass<2>(adjoint, multi, grid, block, lds, st, p, l, a);
    case 2: return launch_pass<4>(adjoint, multi, grid, block, lds, st, p, l, a);
    case 3: return launch_pass<8>(adjoint, multi, grid, block, lds, st, p, l, a);
    case 4: return launch_pass<16>(adjoint, multi, grid, block, lds, st, p, l, a);
    case 5: return launch_pass<32>(adjoint, multi, grid, block, lds, st, p, l, a);
    default: return kBadArgument;
  }
}

}  // namespace tedq

using namespace tedq;

extern "C" int tedq_sv_apply_pass(void* state, int64_t batch, int n_qubits, int tile_bits, int reg_bits,
                                  int threads_log2, const int* dev_prog, uint64_t tile_mask, const void* mats,
                                  int64_t mat_stride, void* stream) {
  return run_pass(false, state, nullptr, batch, n_qubits, tile_bits, reg_bits, threads_log2, dev_prog, tile_mask,
                  mats, mat_stride, nullptr, 0, 0, stream);
}

extern "C" int tedq_sv_adjoint_pass(void* psi, void* lam, int64_t batch, int n_qubits, int tile_bits, int reg_bits,
                                    int threads_log2, const int* dev_prog, uint64_t tile_mask, const void* mats,
                                    int64_t mat_stride, float* partials, int64_t partial_rows, int n_gslots,
                                    void* stream) {
  return run_pass(true, psi, lam, batch, n_qubits, tile_bits, reg_bits, threads_log2, dev_prog, tile_mask, mats,
                  mat_stride, partials, partial_rows, n_gslots, stream);
}
