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
//             local-bit mask [3] register mask [4..36) swizzled local offset of register m
//   ops: 8 words {type | dagger<<8, b0, b1, cmask, mat_off, gslot, w6, w7} (w6/w7: kOpPauliGrad's
//   kind and component, sv_ops.h)
#include "sv_ops.h"

namespace tedq {

// runtime-mask forms for multi-control ops
template <int R, int B>
__device__ __forceinline__ void u1_rt(float2 (&v)[R], float2 m0, float2 m1, float2 m2, float2 m3, int cmask) {
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
__device__ __forceinline__ void cx_rt(float2 (&v)[R], int cmask) {
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
template <int R, int B>
__device__ __forceinline__ void grad_rt(const float2 (&p)[R], const float2 (&l)[R], int cmask, float (&acc)[8]) {
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
template <int R, int B>
__device__ __forceinline__ void r1_rt(float2 (&v)[R], float m0, float m1, float m2, float m3, int cmask) {
#pragma unroll
  for (int i = 0; i < R; ++i) {
    if (i & (1 << B)) continue;
    if ((i & cmask) != cmask) continue;
    const int j = i | (1 << B);
    const float2 x = v[i], y = v[j];
    v[i] = make_float2(fmaf(m0, x.x, m1 * y.x), fmaf(m0, x.y, m1 * y.y));
    v[j] = make_float2(fmaf(m2, x.x, m3 * y.x), fmaf(m2, x.y, m3 * y.y));
  }
}
// gradient value of a one-parameter rotation (PauliGradOp in sv_ops.h) with a runtime kind and mask
template <int R, int B>
__device__ __forceinline__ float pauli_grad_rt(const float2 (&p)[R], const float2 (&l)[R], int cmask, int kind) {
  float acc = 0.f;
#pragma unroll
  for (int i = 0; i < R; ++i) {
    if (i & (1 << B)) continue;
    if ((i & cmask) != cmask) continue;
    const int j = i | (1 << B);
    const float re_ij = fmaf(l[i].x, p[j].x, l[i].y * p[j].y), im_ij = fmaf(l[i].x, p[j].y, -l[i].y * p[j].x);
    const float re_ji = fmaf(l[j].x, p[i].x, l[j].y * p[i].y), im_ji = fmaf(l[j].x, p[i].y, -l[j].y * p[i].x);
    const float im_ii = fmaf(l[i].x, p[i].y, -l[i].y * p[i].x), im_jj = fmaf(l[j].x, p[j].y, -l[j].y * p[j].x);
    acc += kind == kGradX ? im_ij + im_ji : kind == kGradY ? re_ji - re_ij : kind == kGradZ ? im_ii - im_jj
                                                                                             : -2.f * im_jj;
  }
  return acc;
}
template <int R>
__device__ __forceinline__ void diag_rt(float2 (&v)[R], int b, int cmask, float2 d0, float2 d1) {
#pragma unroll
  for (int i = 0; i < R; ++i) {
    if ((i & cmask) != cmask) continue;
    v[i] = cmul(((i >> b) & 1) ? d1 : d0, v[i]);
  }
}
template <int R>
__device__ __forceinline__ void cz_rt(float2 (&v)[R], int mask) {
#pragma unroll
  for (int i = 0; i < R; ++i)
    if ((i & mask) == mask) v[i] = cneg(v[i]);
}

// Dispatch a runtime (target bit, control mask) to Op<R, B, CM> when the mask
// has at most one bit; returns false otherwise.
template <int R, template <int, int, int> class Op, int B, typename... A>
__device__ __forceinline__ bool by_ctrl(int cmask, A&... a) {
  switch (cmask) {
    case 0: Op<R, B, 0>::run(a...); return true;
    case 1: if constexpr (B != 0 && R > 1) { Op<R, B, 1>::run(a...); return true; } break;
    case 2: if constexpr (B != 1 && R > 2) { Op<R, B, 2>::run(a...); return true; } break;
    case 4: if constexpr (B != 2 && R > 4) { Op<R, B, 4>::run(a...); return true; } break;
    case 8: if constexpr (B != 3 && R > 8) { Op<R, B, 8>::run(a...); return true; } break;
    case 16: if constexpr (B != 4 && R > 16) { Op<R, B, 16>::run(a...); return true; } break;
    default: break;
  }
  return false;
}

template <int R, template <int, int, int> class Op, typename... A>
__device__ __forceinline__ bool by_bit(int b, int cmask, A&... a) {
  switch (b) {
    case 0: if constexpr (R > 1) return by_ctrl<R, Op, 0>(cmask, a...); break;
    case 1: if constexpr (R > 2) return by_ctrl<R, Op, 1>(cmask, a...); break;
    case 2: if constexpr (R > 4) return by_ctrl<R, Op, 2>(cmask, a...); break;
    case 3: if constexpr (R > 8) return by_ctrl<R, Op, 3>(cmask, a...); break;
    case 4: if constexpr (R > 16) return by_ctrl<R, Op, 4>(cmask, a...); break;
    default: break;
  }
  return false;
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
  if (by_bit<R, U1Op>(b, cmask, v, m0, m1, m2, m3)) return;
#define CALL(B) u1_rt<R, B>(v, m0, m1, m2, m3, cmask)
  TEDQ_BIT_SWITCH(b, CALL)
#undef CALL
}

template <int R>
__device__ __forceinline__ void op_r1(float2 (&v)[R], int b, int cmask, float m0, float m1, float m2, float m3) {
  if (by_bit<R, R1Op>(b, cmask, v, m0, m1, m2, m3)) return;
#define CALL(B) r1_rt<R, B>(v, m0, m1, m2, m3, cmask)
  TEDQ_BIT_SWITCH(b, CALL)
#undef CALL
}

template <int R>
__device__ __forceinline__ float op_pauli_grad(const float2 (&p)[R], const float2 (&l)[R], int b, int cmask,
                                               int kind) {
  float v = 0.f;
#define CALL(B) v = pauli_grad_rt<R, B>(p, l, cmask, kind)
  TEDQ_BIT_SWITCH(b, CALL)
#undef CALL
  return v;
}

template <int R>
__device__ __forceinline__ void op_cx(float2 (&v)[R], int b, int cmask) {
  if (by_bit<R, CXOp>(b, cmask, v)) return;
#define CALL(B) cx_rt<R, B>(v, cmask)
  TEDQ_BIT_SWITCH(b, CALL)
#undef CALL
}

template <int R>
__device__ __forceinline__ void op_grad(const float2 (&p)[R], const float2 (&l)[R], int b, int cmask,
                                        float (&acc)[8]) {
  if (by_bit<R, GradOp>(b, cmask, p, l, acc)) return;
#define CALL(B) grad_rt<R, B>(p, l, cmask, acc)
  TEDQ_BIT_SWITCH(b, CALL)
#undef CALL
}

template <int R>
__device__ __forceinline__ void op_diag(float2 (&v)[R], int b, int cmask, float2 d0, float2 d1) {
  if (by_bit<R, DiagOp>(b, cmask, v, d0, d1)) return;
  diag_rt<R>(v, b, cmask, d0, d1);
}

template <int R>
__device__ __forceinline__ void op_cz(float2 (&v)[R], int mask) {
  const int top = 31 - __builtin_clz(mask | 1);
  if (by_bit<R, CZOp>(top, mask & ~(1 << top), v)) return;
  cz_rt<R>(v, mask);
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
    case kOpR1: {
      const float2* m = mb + op[4];
      const float m0 = m[0].x, m3 = m[3].x;
      const float m1 = dag ? m[2].x : m[1].x, m2 = dag ? m[1].x : m[2].x;
      op_r1<R>(v, op[1], op[3], m0, m1, m2, m3);
      if (TWO) op_r1<R>(w, op[1], op[3], m0, m1, m2, m3);
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
    case kOpPhase1: {
      const float2 q = dag ? cconj(mb[op[4] + 1]) : mb[op[4] + 1];
      const int cm = op[3] | (1 << op[1]);
      op_diag<R>(v, op[1], cm, q, q);
      if (TWO) op_diag<R>(w, op[1], cm, q, q);
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

// Tiles are moved with buffer instructions: the resource (tile base, uniform
// per workgroup) lives in SGPRs, each lane adds a 32-bit offset and each
// register row a scalar offset, so no 64-bit address is held per element.
// Out-of-range buffer loads return 0 and stores are dropped, which covers the
// partial last workgroup of whole-circuit tiles.  States with more than 2^28
// amplitudes per circuit use plain global addressing.
//
// The swizzle is XOR-linear and a thread's LDS index is (thread part) |
// (register part) with disjoint bits, so swz(base | off) = swz(base) ^
// swz(off): register offsets in the stage header are stored pre-swizzled and
// each element costs one XOR.
template <int R, bool MULTI>
__device__ __forceinline__ void load_tiles(const float2* __restrict__ src, float2* lds, const int* __restrict__ prog,
                                           const PassArgs& a, int tid, int T, const TileCtx& c, int64_t blk) {
  const int sw = swz(tid);
  if (MULTI) {
    const int64_t wg_base = blk * T * R;
    const int64_t left = (a.batch << a.n) - wg_base;
    const int64_t cnt = left < (int64_t)T * R ? left : (int64_t)T * R;
    const __amdgpu_buffer_rsrc_t r = make_rsrc(src + wg_base, (uint32_t)(cnt * 8));
#pragma unroll
    for (int m = 0; m < R; ++m) lds[sw ^ swz(m * T)] = bload(r, (uint32_t)tid * 8u, (uint32_t)(m * T) * 8u);
  } else if (a.n <= 28) {
    const __amdgpu_buffer_rsrc_t r = make_rsrc(src + c.gbase, (uint32_t)(8u << a.n));
    const int64_t* hi = reinterpret_cast<const int64_t*>(prog + kHdrHiOff);
    const uint32_t vo = (uint32_t)c.lo * 8u;
#pragma unroll
    for (int m = 0; m < R; ++m) lds[sw ^ swz(m * T)] = bload(r, vo, (uint32_t)hi[m] * 8u);
  } else {
    const int64_t* hi = reinterpret_cast<const int64_t*>(prog + kHdrHiOff);
    const float2* p = src + c.gbase + c.lo;
#pragma unroll
    for (int m = 0; m < R; ++m) lds[sw ^ swz(m * T)] = p[hi[m]];
  }
}

template <int R, bool MULTI>
__device__ __forceinline__ void store_tiles(float2* __restrict__ dst, const float2* lds, const int* __restrict__ prog,
                                            const PassArgs& a, int tid, int T, const TileCtx& c, int64_t blk) {
  const int sw = swz(tid);
  if (MULTI) {
    const int64_t wg_base = blk * T * R;
    const int64_t left = (a.batch << a.n) - wg_base;
    const int64_t cnt = left < (int64_t)T * R ? left : (int64_t)T * R;
    const __amdgpu_buffer_rsrc_t r = make_rsrc(dst + wg_base, (uint32_t)(cnt * 8));
#pragma unroll
    for (int m = 0; m < R; ++m) bstore(lds[sw ^ swz(m * T)], r, (uint32_t)tid * 8u, (uint32_t)(m * T) * 8u);
  } else if (a.n <= 28) {
    const __amdgpu_buffer_rsrc_t r = make_rsrc(dst + c.gbase, (uint32_t)(8u << a.n));
    const int64_t* hi = reinterpret_cast<const int64_t*>(prog + kHdrHiOff);
    const uint32_t vo = (uint32_t)c.lo * 8u;
#pragma unroll
    for (int m = 0; m < R; ++m) bstore(lds[sw ^ swz(m * T)], r, vo, (uint32_t)hi[m] * 8u);
  } else {
    const int64_t* hi = reinterpret_cast<const int64_t*>(prog + kHdrHiOff);
    float2* p = dst + c.gbase + c.lo;
#pragma unroll
    for (int m = 0; m < R; ++m) p[hi[m]] = lds[sw ^ swz(m * T)];
  }
}

__device__ __forceinline__ TileCtx make_ctx(const PassArgs& a, const int* __restrict__ prog, int tid, bool multi,
                                            int64_t blk) {
  TileCtx c;
  if (multi) {
    const int64_t tile = (blk << a.gT) + (tid >> a.tT);
    c.circuit = tile;  // whole-circuit tiles
    c.valid = tile < a.batch;
    c.gbase = 0;
    c.lo = 0;
  } else {
    c.valid = true;
    tile_origin(a, blk, &c.circuit, &c.gbase);
    uint64_t lo = 0;
    for (int i = 0; i < a.tT; ++i)
      if ((tid >> i) & 1) lo |= 1ull << prog[kHdrTileBits + i];
    c.lo = lo;
  }
  return c;
}

// ---------------------------------------------------------------------------
// forward pass
// ---------------------------------------------------------------------------
// Workgroup `blk` of a forward pass (a tile, or G whole circuits when MULTI).
template <int R, bool MULTI>
__device__ __forceinline__ void pass_fwd_body(float2* __restrict__ state, const int* __restrict__ prog,
                                              const float2* __restrict__ mats, const PassArgs& a, int64_t blk) {
  extern __shared__ float2 lds[];
  const int tid = threadIdx.x;
  const int T = blockDim.x;
  const TileCtx c = make_ctx(a, prog, tid, MULTI, blk);
  const float2* mb = mats + (c.valid ? c.circuit : 0) * a.mat_stride;
  load_tiles<R, MULTI>(state, lds, prog, a, tid, T, c, blk);
  __syncthreads();
  const int nst = prog[2];
  for (int s = 0; s < nst; ++s) {
    const int* st = prog + kHdrStages + kStageWords * s;
    const int sb = stage_base(tid, a.tT, a.k, st[2]);
    float2 v[R];
#pragma unroll
    for (int m = 0; m < R; ++m) v[m] = lds[sb ^ st[4 + m]];
    const int nops = st[0];
    const int* op = prog + st[1];
    for (int o = 0; o < nops; ++o, op += 8) apply_op<R, false>(v, v, op, mb);
#pragma unroll
    for (int m = 0; m < R; ++m) lds[sb ^ st[4 + m]] = v[m];
    __syncthreads();
  }
  store_tiles<R, MULTI>(state, lds, prog, a, tid, T, c, blk);
}

template <int R, bool MULTI>
__global__ void __launch_bounds__(512) k_pass_fwd(float2* __restrict__ state, const int* __restrict__ prog,
                                                  const float2* __restrict__ mats, PassArgs a) {
  pass_fwd_body<R, MULTI>(state, prog, mats, a, blockIdx.x);
}

// ---------------------------------------------------------------------------
// adjoint pass: walks the (already reversed and daggered) op list over psi and
// lambda; gradient ops reduce their 8 bilinears over the threads of a tile.
// ---------------------------------------------------------------------------
template <int R, bool MULTI>
__device__ __forceinline__ void pass_adj_body(float2* __restrict__ psi, float2* __restrict__ lam,
                                              const int* __restrict__ prog, const float2* __restrict__ mats,
                                              float* __restrict__ partials, const PassArgs& a, int64_t blk) {
  extern __shared__ float2 lds[];
  float2* ldp = lds;
  const int T = blockDim.x;
  float2* ldl = lds + T * R;
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const TileCtx c = make_ctx(a, prog, tid, MULTI, blk);
  const float2* mb = mats + (c.valid ? c.circuit : 0) * a.mat_stride;
  load_tiles<R, MULTI>(psi, ldp, prog, a, tid, T, c, blk);
  load_tiles<R, MULTI>(lam, ldl, prog, a, tid, T, c, blk);
  __syncthreads();

  // partial rows: W waves per tile when a tile spans several waves
  const int tpt = 1 << a.tT;                 // threads per tile
  const int seg = tpt < 64 ? tpt : 64;       // lanes reduced together
  const int W = tpt >> 6 > 0 ? tpt >> 6 : 1; // partial rows per tile
  const int tl = tid & (tpt - 1);
  int64_t tile_in_circuit = MULTI ? 0 : (blk & ((1ll << tiles_log2(a)) - 1));
  const int64_t prow = tile_in_circuit * W + (tl >> 6);
  const bool seg_writer = c.valid && ((tid & (seg - 1)) == 0);
  const bool wave_writer = (lane & 7) == 0;
  float* const pbase = partials + (c.circuit * a.n_gslots * a.np + prow) * 8;

  // pending group of one-parameter gradient values (kOpPauliGrad), written by kOpGradFlush
  float sacc[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  const int nst = prog[2];
  for (int s = 0; s < nst; ++s) {
    const int* st = prog + kHdrStages + kStageWords * s;
    const int sb = stage_base(tid, a.tT, a.k, st[2]);
    float2 vp[R], vl[R];
#pragma unroll
    for (int m = 0; m < R; ++m) {
      vp[m] = ldp[sb ^ st[4 + m]];
      vl[m] = ldl[sb ^ st[4 + m]];
    }
    const int nops = st[0];
    const int* op = prog + st[1];
    for (int o = 0; o < nops; ++o, op += 8) {
      const int type = op[0] & 0xff;
      if (type == kOpPauliGrad) {
        float v = op_pauli_grad<R>(vp, vl, op[1], op[3], op[6] & 15);
        if (op[6] >> 4) v = -v;
#pragma unroll
        for (int q = 0; q < 8; ++q)
          if (q == op[7]) sacc[q] += v;
      } else if (type == kOpGrad1 || type == kOpCGrad1 || type == kOpGradFlush) {
        const bool flush = type == kOpGradFlush;
        float acc[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
        if (flush) {
#pragma unroll
          for (int q = 0; q < 8; ++q) {
            acc[q] = sacc[q];
            sacc[q] = 0.f;
          }
        } else {
          op_grad<R>(vp, vl, op[1], op[3], acc);
        }
        float* dst = pbase + (int64_t)op[5] * a.np * 8;
        if (seg == 64) {
          const float tot = reduce8_wave(acc, lane);
          if (wave_writer) dst[(lane >> 3) & 7] = tot;
        } else {
#pragma unroll
          for (int q = 0; q < 8; ++q)
            for (int off = 1; off < seg; off <<= 1) acc[q] += __shfl_xor(acc[q], off, 64);
          if (seg_writer) {
            float4* d4 = reinterpret_cast<float4*>(dst);
            d4[0] = make_float4(acc[0], acc[1], acc[2], acc[3]);
            d4[1] = make_float4(acc[4], acc[5], acc[6], acc[7]);
          }
        }
      } else {
        apply_op<R, true>(vp, vl, op, mb);
      }
    }
#pragma unroll
    for (int m = 0; m < R; ++m) {
      ldp[sb ^ st[4 + m]] = vp[m];
      ldl[sb ^ st[4 + m]] = vl[m];
    }
    __syncthreads();
  }
  store_tiles<R, MULTI>(psi, ldp, prog, a, tid, T, c, blk);
  store_tiles<R, MULTI>(lam, ldl, prog, a, tid, T, c, blk);
}

template <int R, bool MULTI>
__global__ void __launch_bounds__(512) k_pass_adj(float2* __restrict__ psi, float2* __restrict__ lam,
                                                  const int* __restrict__ prog, const float2* __restrict__ mats,
                                                  float* __restrict__ partials, PassArgs a) {
  pass_adj_body<R, MULTI>(psi, lam, prog, mats, partials, a, blockIdx.x);
}

// ---------------------------------------------------------------------------
// ragged passes: circuits of different structures and sizes in one launch
// (a serving batch).  Every circuit has its own state, program, matrices and
// partial rows; all share the tile bits k and register bits r, so one grid of
// 2^(k-r)-thread workgroups covers the tiles of all of them.  Workgroup b
// belongs to the first record whose tile_end exceeds b.
// ---------------------------------------------------------------------------
struct RaggedRec {
  float2* psi;           // the circuit's state [2^n]
  float2* lam;           // adjoint: its lambda [2^n]
  const int* prog;       // its program for this pass (forward or adjoint)
  const float2* mats;    // its slot matrices (one table row)
  float* partials;       // adjoint: its partial rows [n_gslots][2^(n-k) * W][8]
  uint64_t rest_mask;    // physical bits of the circuit that are not tile bits of this pass
  int64_t tile_end;      // tiles of this and all earlier records
  int n;
  int n_gslots;
};
static_assert(sizeof(RaggedRec) == 64, "ragged record layout is part of the C ABI");

__device__ __forceinline__ int ragged_record(const RaggedRec* __restrict__ recs, int nrec, int64_t blk) {
  int lo = 0, hi = nrec - 1;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (recs[mid].tile_end > blk) hi = mid;
    else lo = mid + 1;
  }
  return __builtin_amdgcn_readfirstlane(lo);
}

// The record is the same for the whole workgroup: read every word through readfirstlane so that the
// compiler keeps pointers, masks and sizes in scalar registers (as kernel arguments are).
__device__ __forceinline__ uint64_t uniform64(uint64_t v) {
  const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
  const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
  return ((uint64_t)hi << 32) | lo;
}

// A pointer read from memory is generic (flat); taking it through the global address space lets the
// compiler issue global (and, for uniform read-only words such as the program, scalar) loads, as it does
// for pointer kernel arguments.
template <class T>
__device__ __forceinline__ T* global_ptr(const void* p) {
  return (T*)(__attribute__((address_space(1))) T*)uniform64(reinterpret_cast<uint64_t>(p));
}

__device__ __forceinline__ RaggedRec load_record(const RaggedRec* __restrict__ recs, int c) {
  RaggedRec r;
  const RaggedRec& g = recs[c];
  r.psi = global_ptr<float2>(g.psi);
  r.lam = global_ptr<float2>(g.lam);
  r.prog = global_ptr<const int>(g.prog);
  r.mats = global_ptr<const float2>(g.mats);
  r.partials = global_ptr<float>(g.partials);
  r.rest_mask = uniform64(g.rest_mask);
  r.tile_end = (int64_t)uniform64((uint64_t)g.tile_end);
  r.n = __builtin_amdgcn_readfirstlane(g.n);
  r.n_gslots = __builtin_amdgcn_readfirstlane(g.n_gslots);
  return r;
}

__device__ __forceinline__ PassArgs ragged_args(const RaggedRec& rc, int k, int tT) {
  PassArgs a;
  a.mat_stride = 0;
  a.batch = 1;
  a.total_tiles = 1ll << (rc.n - k);
  a.rest_mask = rc.rest_mask;
  a.n = rc.n;
  a.k = k;
  a.tT = tT;
  a.gT = 0;
  const int W = tT > 6 ? 1 << (tT - 6) : 1;
  a.np = (1ll << (rc.n - k)) * W;
  a.n_gslots = rc.n_gslots;
  a.flags = 0;
  return a;
}

template <int R>
__global__ void __launch_bounds__(512) k_pass_fwd_ragged(const RaggedRec* __restrict__ recs, int nrec, int k,
                                                         int tT) {
  const int c = ragged_record(recs, nrec, blockIdx.x);
  const RaggedRec rc = load_record(recs, c);
  const int64_t first = c ? (int64_t)uniform64((uint64_t)recs[c - 1].tile_end) : 0;
  pass_fwd_body<R, false>(rc.psi, rc.prog, rc.mats, ragged_args(rc, k, tT), (int64_t)blockIdx.x - first);
}

template <int R>
__global__ void __launch_bounds__(512) k_pass_adj_ragged(const RaggedRec* __restrict__ recs, int nrec, int k,
                                                         int tT) {
  const int c = ragged_record(recs, nrec, blockIdx.x);
  const RaggedRec rc = load_record(recs, c);
  const int64_t first = c ? (int64_t)uniform64((uint64_t)recs[c - 1].tile_end) : 0;
  pass_adj_body<R, false>(rc.psi, rc.lam, rc.prog, rc.mats, rc.partials, ragged_args(rc, k, tT),
                          (int64_t)blockIdx.x - first);
}

// gfx950 provides 160 KiB of LDS per workgroup; above 64 KiB the kernel must opt in once.
constexpr size_t kMaxLds = 160 * 1024;

template <typename K>
static bool allow_lds(K kernel, size_t lds) {
  if (lds <= 65536) return true;
  return hipFuncSetAttribute(reinterpret_cast<const void*>(kernel), hipFuncAttributeMaxDynamicSharedMemorySize,
                             (int)lds) == hipSuccess;
}

template <int R>
static int launch_pass(bool adjoint, bool multi, dim3 grid, dim3 block, size_t lds, hipStream_t st, float2* psi,
                       float2* lam, const int* prog, const float2* mats, float* partials, const PassArgs& a) {
  bool ok = true;
  if (adjoint) {
    if (multi) {
      ok = allow_lds(k_pass_adj<R, true>, lds);
      if (ok) hipLaunchKernelGGL((k_pass_adj<R, true>), grid, block, lds, st, psi, lam, prog, mats, partials, a);
    } else {
      ok = allow_lds(k_pass_adj<R, false>, lds);
      if (ok) hipLaunchKernelGGL((k_pass_adj<R, false>), grid, block, lds, st, psi, lam, prog, mats, partials, a);
    }
  } else {
    if (multi) {
      ok = allow_lds(k_pass_fwd<R, true>, lds);
      if (ok) hipLaunchKernelGGL((k_pass_fwd<R, true>), grid, block, lds, st, psi, prog, mats, a);
    } else {
      ok = allow_lds(k_pass_fwd<R, false>, lds);
      if (ok) hipLaunchKernelGGL((k_pass_fwd<R, false>), grid, block, lds, st, psi, prog, mats, a);
    }
  }
  if (!ok) {
    set_error("sv pass: could not enable %zu bytes of LDS", lds);
    return kLaunchFailed;
  }
  return check_launch(adjoint ? "tedq_sv_adjoint_pass" : "tedq_sv_apply_pass");
}

static int run_pass(bool adjoint, void* psi, void* lam, int64_t batch, int n, int k, int r, int threads_log2,
                    const int* dev_prog, uint64_t tile_mask, uint64_t zero_mask, const void* mats,
                    int64_t mat_stride, float* partials, int64_t np, int n_gslots, void* stream) {
  if (psi == nullptr || (adjoint && (lam == nullptr || (n_gslots > 0 && partials == nullptr))) ||
      dev_prog == nullptr || batch <= 0 || n < 1 || n > 40 || k < 1 || k > n || r < 1 || r > 5 || r > k) {
    set_error("sv pass: bad arguments (batch=%lld n=%d k=%d r=%d)", (long long)batch, n, k, r);
    return kBadArgument;
  }
  const int tT = k - r;
  if (threads_log2 < tT || threads_log2 > 9 || (threads_log2 > tT && k != n)) {
    set_error("sv pass: threads_log2=%d incompatible with k=%d r=%d n=%d", threads_log2, k, r, n);
    return kBadArgument;
  }
  const bool multi = (k == n);
  const uint64_t full = (n >= 64) ? ~0ull : ((1ull << n) - 1);
  if (__builtin_popcountll(tile_mask & full) != k || (tile_mask & ~full) != 0) {
    set_error("sv pass: tile mask does not have k=%d bits below n=%d", k, n);
    return kBadArgument;
  }
  if ((zero_mask & (tile_mask | ~full)) != 0) {
    set_error("sv pass: zero mask 0x%llx holds tile bits or bits above n", (unsigned long long)zero_mask);
    return kBadArgument;
  }
  PassArgs a;
  a.mat_stride = mat_stride;
  a.batch = batch;
  a.rest_mask = full & ~tile_mask & ~zero_mask;
  a.total_tiles = batch << __builtin_popcountll(a.rest_mask);
  a.n = n;
  a.k = k;
  a.tT = tT;
  a.gT = threads_log2 - tT;
  a.np = np;
  a.n_gslots = n_gslots;
  a.flags = 0;
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
  if (lds > kMaxLds) {
    set_error("sv pass: LDS request %zu bytes exceeds %zu", lds, kMaxLds);
    return kBadArgument;
  }
  float2* p = static_cast<float2*>(psi);
  float2* l = static_cast<float2*>(lam);
  const float2* mt = static_cast<const float2*>(mats);
  hipStream_t st = static_cast<hipStream_t>(stream);
  dim3 grid((unsigned)blocks), block((unsigned)T);
  switch (r) {
    case 1: return launch_pass<2>(adjoint, multi, grid, block, lds, st, p, l, dev_prog, mt, partials, a);
    case 2: return launch_pass<4>(adjoint, multi, grid, block, lds, st, p, l, dev_prog, mt, partials, a);
    case 3: return launch_pass<8>(adjoint, multi, grid, block, lds, st, p, l, dev_prog, mt, partials, a);
    case 4: return launch_pass<16>(adjoint, multi, grid, block, lds, st, p, l, dev_prog, mt, partials, a);
    case 5: return launch_pass<32>(adjoint, multi, grid, block, lds, st, p, l, dev_prog, mt, partials, a);
    default: return kBadArgument;
  }
}

template <int R>
static int launch_ragged(bool adjoint, const RaggedRec* recs, int nrec, int64_t tiles, int k, int r,
                         hipStream_t st) {
  const int tT = k - r;
  const int T = 1 << tT;
  const size_t lds = (size_t)T * R * sizeof(float2) * (adjoint ? 2 : 1);
  bool ok;
  if (adjoint) {
    ok = allow_lds(k_pass_adj_ragged<R>, lds);
    if (ok) hipLaunchKernelGGL(k_pass_adj_ragged<R>, dim3((unsigned)tiles), dim3(T), lds, st, recs, nrec, k, tT);
  } else {
    ok = allow_lds(k_pass_fwd_ragged<R>, lds);
    if (ok) hipLaunchKernelGGL(k_pass_fwd_ragged<R>, dim3((unsigned)tiles), dim3(T), lds, st, recs, nrec, k, tT);
  }
  if (!ok) {
    set_error("ragged pass: could not enable %zu bytes of LDS", lds);
    return kLaunchFailed;
  }
  return check_launch("tedq_sv_pass_ragged");
}

}  // namespace tedq

using namespace tedq;

extern "C" int tedq_sv_pass_ragged(int adjoint, const void* dev_recs, int n_recs, int64_t total_tiles, int k,
                                   int r, void* stream) {
  if (dev_recs == nullptr || n_recs < 1 || total_tiles < n_recs || total_tiles > 0x7fffffffll || k < 1 ||
      k > 20 || r < 1 || r > 5 || r > k || k - r > 9 || (1ll << (k - r)) * (1ll << r) * 8 * (adjoint ? 2 : 1) >
      (int64_t)kMaxLds) {
    set_error("tedq_sv_pass_ragged: bad arguments (records=%d tiles=%lld k=%d r=%d)", n_recs,
              (long long)total_tiles, k, r);
    return kBadArgument;
  }
  const RaggedRec* recs = static_cast<const RaggedRec*>(dev_recs);
  hipStream_t st = static_cast<hipStream_t>(stream);
  const bool adj = adjoint != 0;
  switch (r) {
    case 1: return launch_ragged<2>(adj, recs, n_recs, total_tiles, k, r, st);
    case 2: return launch_ragged<4>(adj, recs, n_recs, total_tiles, k, r, st);
    case 3: return launch_ragged<8>(adj, recs, n_recs, total_tiles, k, r, st);
    case 4: return launch_ragged<16>(adj, recs, n_recs, total_tiles, k, r, st);
    case 5: return launch_ragged<32>(adj, recs, n_recs, total_tiles, k, r, st);
    default: return kBadArgument;
  }
}

extern "C" int tedq_sv_apply_pass(void* state, int64_t batch, int n_qubits, int tile_bits, int reg_bits,
                                  int threads_log2, const int* dev_prog, uint64_t tile_mask, uint64_t zero_mask,
                                  const void* mats, int64_t mat_stride, void* stream) {
  return run_pass(false, state, nullptr, batch, n_qubits, tile_bits, reg_bits, threads_log2, dev_prog, tile_mask,
                  zero_mask, mats, mat_stride, nullptr, 0, 0, stream);
}

extern "C" int tedq_sv_adjoint_pass(void* psi, void* lam, int64_t batch, int n_qubits, int tile_bits, int reg_bits,
                                    int threads_log2, const int* dev_prog, uint64_t tile_mask, uint64_t zero_mask,
                                    const void* mats, int64_t mat_stride, float* partials, int64_t partial_rows,
                                    int n_gslots, void* stream) {
  return run_pass(true, psi, lam, batch, n_qubits, tile_bits, reg_bits, threads_log2, dev_prog, tile_mask,
                  zero_mask, mats, mat_stride, partials, partial_rows, n_gslots, stream);
}
