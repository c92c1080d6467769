// Pairwise tensor-network contraction step on gfx950.
//
// Replaces one step of the reference's contraction along its frozen order
// (tedq/backends/pytorch_backend.py:276-354, `tree.contract(arrays,
// backend='torch')`, which issues tensordot/einsum = permute + batched complex
// GEMM per pairwise step).  Here the permutation is never materialised: every
// index of every tensor has dimension 2, so a tensor's layout is a map from
// its index bits to address bits, and the kernels gather operands and scatter
// results through those maps while they contract.
//
// One step, per batch element b:
//     C[h, m, n] = sum_k A[h, m, k] * B[h, k, n]
// h: indices kept and shared (Hadamard), m: kept, only in A, n: kept, only in
// B, k: summed.  The caller passes, for each of the nine (tensor, role) pairs,
// the address bit of every role bit (role bit j of h, m, n or k -> bit
// position in A, B or C), so any operand and result layout is allowed.
//
// Kernels:
//   * k_tn_small: output elements in C's address order (coalesced writes), a
//     group of LPO lanes per element strides over k (lanes adjacent in k, so
//     an inner product over A's low bits reads contiguous memory), reduced with
//     wave shuffles; for large k the range is additionally split over S
//     workgroups whose partial sums are added in a fixed order by
//     k_tn_split_sum (deterministic, no atomics).
//   * k_tn_stream: the skinny steps (a large tensor times a gate), 16
//     outputs per thread with table-driven operand offsets;
//   * k_tn_gate: the common case of those (at most two summed and two new
//     indices, nothing kept-shared): each large-operand row read once;
//   * k_tn_env: few outputs summed over a long range (gate environments);
//   * k_tn_permute: a tensor times a scalar into another layout (operand
//     copies), transposed through LDS tiles;
//   * k_tn_gemm: 64x64 output tiles, k in chunks of 16 staged through LDS,
//     complex products on the FP32 matrix cores (four real 16x16x4 MFMAs per
//     complex block and k-step); used when m, n and k are all wide;
//   * k_tn_gemm128: the same on 128x128 tiles with register prefetch of the
//     next k chunk, for steps whose m and n are both at least 128 wide (both
//     GEMM kernels split long sums over workgroups when tiles are few);
//   * k_tn_small_multi (tedq_tn_contract_small_batch): every small step of
//     one level of the contraction tree in one launch;
//   * k_tn_gate_bwd (tedq_tn_gate_backward): both environments of a
//     k_tn_gate step from one read of the result's environment.
// Arithmetic is complex64 with fp32 accumulation, like the reference's
// complex64 torch contraction.
#include <stdlib.h>

#include "tedq_common.h"

namespace tedq {

constexpr int kMaxRank = 30;  // per-batch tensor of 2^30 complex64 = 8 GiB

// host-side copy of a step descriptor, used to validate it and plan the launch
struct TNDesc {
  int nb[4];               // bits of h, m, n, k
  signed char pos[9][32];  // A_h A_m A_k B_h B_k B_n C_h C_m C_n
  int gemm_form;           // descriptor word kDescForm (see below)
};

// The last descriptor word (slot 31 of the C_n role, never a role bit since
// ranks are at most kMaxRank = 30) selects how a GEMM-shaped step runs:
// 0 = complex products as 4 real MFMA products on 128 x 128 tiles where both
// free sizes allow it (the default), 3 = the 3-product (Gauss) form on 64 x 64
// tiles, 64 = 4 products on 64 x 64 tiles only.  The Gauss form issues 25%
// fewer MFMAs but needs a third accumulator set (132 VGPRs against 90) and
// measured slower on the 40-qubit chain (191 ms against 170 ms per batch).
constexpr int kDescWords = 4 + 9 * 32;
constexpr int kDescForm = kDescWords - 1;
enum { AH = 0, AM, AK, BH, BK, BN, CH, CM, CN };

// Role layout as the kernels hold it in LDS (copied from the device
// descriptor at kernel start, so that it costs no registers).
struct TNShared {
  int nb[4];
  signed char pos[9][32];
};

__device__ __forceinline__ void load_desc(TNShared& s, const int32_t* __restrict__ dd) {
  for (int i = threadIdx.x; i < 4 + 9 * 32; i += blockDim.x) {
    if (i < 4) s.nb[i] = dd[i];
    else (&s.pos[0][0])[i - 4] = (signed char)dd[i];
  }
  __syncthreads();
}

// address bits of role value x: bit j of x goes to address bit pos[R][j]
__device__ __forceinline__ uint32_t scat(const TNShared& s, int R, int nbits, uint32_t x) {
  uint32_t off = 0;
  for (int j = 0; j < nbits; ++j) off |= ((x >> j) & 1u) << s.pos[R][j];
  return off;
}

constexpr int kTabBits = 10;  // k offsets for the low k bits come from LDS tables

// Address maps from the output's bits: bit p of a C offset belongs to an index
// that sits at bit mapA[p] of A (h and m roles) and/or mapB[p] of B (h and n
// roles), -1 where it is not there.  Built per workgroup in LDS.
__device__ __forceinline__ void c_bit_maps(const TNShared& d, signed char* mapA, signed char* mapB) {
  const int nH = d.nb[0], nM = d.nb[1], nN = d.nb[2];
  for (int p = threadIdx.x; p < 32; p += blockDim.x) {
    mapA[p] = -1;
    mapB[p] = -1;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int j = 0; j < nH; ++j) {
      mapA[d.pos[CH][j]] = d.pos[AH][j];
      mapB[d.pos[CH][j]] = d.pos[BH][j];
    }
    for (int j = 0; j < nM; ++j) mapA[d.pos[CM][j]] = d.pos[AM][j];
    for (int j = 0; j < nN; ++j) mapB[d.pos[CN][j]] = d.pos[BN][j];
  }
  __syncthreads();
}

__device__ __forceinline__ uint32_t via_map(const signed char* map, int bits, uint32_t o) {
  uint32_t off = 0;
  for (int p = 0; p < bits; ++p)
    if (map[p] >= 0) off |= ((o >> p) & 1u) << map[p];
  return off;
}

// Output elements are enumerated in C's own address order (consecutive
// threads write consecutive addresses); the operand offsets follow through
// the bit maps above.
__global__ void __launch_bounds__(256) k_tn_small(const float2* __restrict__ A, const float2* __restrict__ B,
                                                  float2* __restrict__ C, float2* __restrict__ part,
                                                  const int32_t* __restrict__ dd, int64_t a_bs, int64_t b_bs,
                                                  int64_t c_bs, int lpo_log, int split_log, int64_t n_out) {
  __shared__ uint32_t tka[1 << kTabBits];
  __shared__ uint32_t tkb[1 << kTabBits];
  __shared__ TNShared d;
  __shared__ signed char mapA[32], mapB[32];
  load_desc(d, dd);
  c_bit_maps(d, mapA, mapB);
  const int nH = d.nb[0], nM = d.nb[1], nN = d.nb[2], nK = d.nb[3];
  const int rc = nH + nM + nN;
  const int tb = nK < kTabBits ? nK : kTabBits;
  for (int i = threadIdx.x; i < (1 << tb); i += blockDim.x) {
    tka[i] = scat(d, AK, tb, i);
    tkb[i] = scat(d, BK, tb, i);
  }
  __syncthreads();
  const int64_t b = blockIdx.y;
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int lpo = 1 << lpo_log;
  const int sub = (int)(g & (lpo - 1));
  const int64_t rest = g >> lpo_log;
  const int64_t o = rest >> split_log;                  // output element = C offset
  const int s = (int)(rest & ((1 << split_log) - 1));  // k segment
  const bool valid = o < n_out;
  const uint32_t oo = valid ? (uint32_t)o : 0u;
  const float2* a = A + b * a_bs + via_map(mapA, rc, oo);
  const float2* bp = B + b * b_bs + via_map(mapB, rc, oo);
  const int64_t kcount = 1ll << nK;
  const int64_t seg = kcount >> split_log;
  const int64_t k0 = (int64_t)s * seg;
  float2 acc = make_float2(0.f, 0.f);
  if (valid) {
    const uint32_t lo_mask = (1u << tb) - 1;
    for (int64_t kb = k0; kb < k0 + seg; kb += (int64_t)lo_mask + 1) {
      // high k bits are constant over one table span
      uint32_t ha = 0, hb = 0;
      if (nK > tb) {
        const uint32_t hi = (uint32_t)(kb >> tb);
        for (int j = 0; j < nK - tb; ++j) {
          ha |= ((hi >> j) & 1u) << d.pos[AK][tb + j];
          hb |= ((hi >> j) & 1u) << d.pos[BK][tb + j];
        }
      }
      const int64_t span = (seg < (int64_t)lo_mask + 1) ? seg : (int64_t)lo_mask + 1;
      const uint32_t lo0 = (uint32_t)(kb & lo_mask);
      for (int64_t i = sub; i < span; i += lpo) {
        const uint32_t kl = lo0 + (uint32_t)i;
        const float2 x = a[ha | tka[kl]];
        const float2 y = bp[hb | tkb[kl]];
        acc.x = fmaf(x.x, y.x, fmaf(-x.y, y.y, acc.x));
        acc.y = fmaf(x.x, y.y, fmaf(x.y, y.x, acc.y));
      }
    }
  }
  for (int off = 1; off < lpo; off <<= 1) {
    acc.x += __shfl_xor(acc.x, off, 64);
    acc.y += __shfl_xor(acc.y, off, 64);
  }
  if (!valid || sub != 0) return;
  if (split_log == 0) {
    C[b * c_bs + o] = acc;
  } else {
    part[(b << split_log | s) * n_out + o] = acc;
  }
}

// Streaming form for the common skinny step (a large tensor times a gate:
// at most 16 summed terms, at least 2^12 outputs).  Workgroup g owns C
// offsets g*4096 .. +4095; thread t handles o = g*4096 + j*256 + t for
// j < 16, so the operand offsets split into a per-thread part (bits 0..7),
// a per-j part (bits 8..11, 16-entry LDS tables) and a per-workgroup part
// (bits >= 12): two ORs per element instead of a loop over the output bits.
__global__ void __launch_bounds__(256) k_tn_stream(const float2* __restrict__ A, const float2* __restrict__ B,
                                                   float2* __restrict__ C, const int32_t* __restrict__ dd,
                                                   int64_t a_bs, int64_t b_bs, int64_t c_bs) {
  __shared__ TNShared d;
  __shared__ signed char mapA[32], mapB[32];
  __shared__ uint32_t midA[16], midB[16], tka[16], tkb[16];
  load_desc(d, dd);
  c_bit_maps(d, mapA, mapB);
  const int nH = d.nb[0], nM = d.nb[1], nN = d.nb[2], nK = d.nb[3];
  const int rc = nH + nM + nN;
  const int t = threadIdx.x;
  if (t < 16) {
    uint32_t ma = 0, mb = 0;
    for (int p = 8; p < 12; ++p) {
      const uint32_t bit = (t >> (p - 8)) & 1u;
      if (mapA[p] >= 0) ma |= bit << mapA[p];
      if (mapB[p] >= 0) mb |= bit << mapB[p];
    }
    midA[t] = ma;
    midB[t] = mb;
  } else if (t >= 32 && t < 32 + (1 << nK)) {
    tka[t - 32] = scat(d, AK, nK, t - 32);
    tkb[t - 32] = scat(d, BK, nK, t - 32);
  }
  __syncthreads();
  const uint32_t g = blockIdx.x;
  uint32_t hiA = 0, hiB = 0, loA = 0, loB = 0;
  for (int p = 12; p < rc; ++p) {
    const uint32_t bit = (g >> (p - 12)) & 1u;
    if (mapA[p] >= 0) hiA |= bit << mapA[p];
    if (mapB[p] >= 0) hiB |= bit << mapB[p];
  }
  for (int p = 0; p < 8; ++p) {
    const uint32_t bit = ((uint32_t)t >> p) & 1u;
    if (mapA[p] >= 0) loA |= bit << mapA[p];
    if (mapB[p] >= 0) loB |= bit << mapB[p];
  }
  const int64_t b = blockIdx.y;
  const float2* a = A + b * a_bs + (hiA | loA);
  const float2* bp = B + b * b_bs + (hiB | loB);
  float2* c = C + b * c_bs + ((int64_t)g << 12) + t;
  const int kc = 1 << nK;
#pragma unroll
  for (int j = 0; j < 16; ++j) {
    const uint32_t oa = midA[j], ob = midB[j];
    float2 acc = make_float2(0.f, 0.f);
    for (int k = 0; k < kc; ++k) {
      const float2 x = a[oa | tka[k]];
      const float2 y = bp[ob | tkb[k]];
      acc.x = fmaf(x.x, y.x, fmaf(-x.y, y.y, acc.x));
      acc.y = fmaf(x.x, y.y, fmaf(x.y, y.x, acc.y));
    }
    c[j << 8] = acc;
  }
}

// A large tensor times a gate-sized one without kept-shared indices (the
// gate applications of state-like intermediates): every thread owns four rows
// p of the large operand X (its kept indices), reads X[p, k] for all 2^NK
// summed values once and writes C[p, q] = sum_k G[q, k] X[p, k] for all 2^NQ
// gate outputs, so X is read once and C written once (the general streaming
// form reads X once per output).  Rows are numbered in C's address order; row
// r = (g * 4 + j) * 256 + t takes its offsets from a per-thread part (row bits
// 0..7), a four-entry LDS table (bits 8..9) and a per-workgroup part.  XBIG:
// A is the large operand (else B; the product commutes).  Needs nH = 0 and at
// least 2^10 rows.
template <int NK, int NQ>
__global__ void __launch_bounds__(256) k_tn_gate(const float2* __restrict__ A, const float2* __restrict__ B,
                                                 float2* __restrict__ C, const int32_t* __restrict__ dd,
                                                 int64_t a_bs, int64_t b_bs, int64_t c_bs, int a_big) {
  constexpr int KN = 1 << NK, QN = 1 << NQ;
  __shared__ TNShared d;
  __shared__ uint32_t midX[4], midC[4];
  load_desc(d, dd);
  const int XP = a_big ? AM : BN, XK = a_big ? AK : BK, GQ = a_big ? BN : AM, GK = a_big ? BK : AK;
  const int CP = a_big ? CM : CN, CQ = a_big ? CN : CM;
  const int nP = d.nb[a_big ? 1 : 2];
  const int t = threadIdx.x;
  if (t < 4) {
    uint32_t ox = 0, oc = 0;
    for (int j = 8; j < 10; ++j) {
      const uint32_t bit = ((uint32_t)t >> (j - 8)) & 1u;
      ox |= bit << d.pos[XP][j];
      oc |= bit << d.pos[CP][j];
    }
    midX[t] = ox;
    midC[t] = oc;
  }
  __syncthreads();
  const float2* X = a_big ? A + blockIdx.y * a_bs : B + blockIdx.y * b_bs;
  const float2* G = a_big ? B + blockIdx.y * b_bs : A + blockIdx.y * a_bs;
  float2 g[QN][KN];
  uint32_t xk[KN], cq[QN];
#pragma unroll
  for (int k = 0; k < KN; ++k) xk[k] = scat(d, XK, NK, k);
#pragma unroll
  for (int q = 0; q < QN; ++q) {
    cq[q] = scat(d, CQ, NQ, q);
#pragma unroll
    for (int k = 0; k < KN; ++k) g[q][k] = G[scat(d, GQ, NQ, q) | scat(d, GK, NK, k)];
  }
  uint32_t loX = 0, loC = 0, hiX = 0, hiC = 0;
  for (int j = 0; j < 8; ++j) {
    const uint32_t bit = ((uint32_t)t >> j) & 1u;
    loX |= bit << d.pos[XP][j];
    loC |= bit << d.pos[CP][j];
  }
  const uint32_t blk = blockIdx.x;
  for (int j = 10; j < nP; ++j) {
    const uint32_t bit = (blk >> (j - 10)) & 1u;
    hiX |= bit << d.pos[XP][j];
    hiC |= bit << d.pos[CP][j];
  }
  float2* c = C + blockIdx.y * c_bs;
  float2 x[4][KN];
#pragma unroll
  for (int j = 0; j < 4; ++j)
#pragma unroll
    for (int k = 0; k < KN; ++k) x[j][k] = X[hiX | midX[j] | loX | xk[k]];
#pragma unroll
  for (int j = 0; j < 4; ++j)
#pragma unroll
    for (int q = 0; q < QN; ++q) {
      float2 acc = make_float2(0.f, 0.f);
#pragma unroll
      for (int k = 0; k < KN; ++k) {
        acc.x = fmaf(g[q][k].x, x[j][k].x, fmaf(-g[q][k].y, x[j][k].y, acc.x));
        acc.y = fmaf(g[q][k].x, x[j][k].y, fmaf(g[q][k].y, x[j][k].x, acc.y));
      }
      c[hiC | midC[j] | loC | cq[q]] = acc;
    }
}

// A tensor times a scalar into another index layout (the copies of GEMM
// operands into k-contiguous layouts): a permutation of address bits.  Each
// workgroup moves tiles of 2^10 elements: the tile bits are the index bits at
// the five lowest source and the five lowest destination address bits (topped
// up with the next destination bits), so the loads (in source order) and the
// stores (in destination order) both run along contiguous addresses, and LDS
// carries the transpose.  The remaining bits number the tiles.  The host
// derives the bit lists once per launch (TNPermPlan, passed by value).
constexpr int kPermBits = 10;
constexpr int kPermTilesPerGroup = 4;

struct TNPermPlan {
  int n_rest;                    // index bits outside the tile
  int a_tensor;                  // A is the tensor, B the scalar (else the reverse)
  unsigned char src_t[kPermBits];  // source address bit of tile bit q, tile bits in source order
  unsigned char dst_t[kPermBits];  // destination address bit of tile bit q, tile bits in destination order
  unsigned char s_of_d[kPermBits];  // source-order number of the destination-order tile bit q
  unsigned char src_r[32], dst_r[32];  // source / destination address bit of rest bit q
};

__global__ void __launch_bounds__(256) k_tn_permute(const float2* __restrict__ A, const float2* __restrict__ B,
                                                    float2* __restrict__ C, int64_t a_bs, int64_t b_bs,
                                                    int64_t c_bs, TNPermPlan P) {
  constexpr int N = 1 << kPermBits;
  __shared__ float2 tile[N + N / 32];  // one pad word per 32 elements
  __shared__ uint32_t tabA[N];         // source offset part of the i-th element in source order
  __shared__ uint32_t tabC[N];         // destination offset part of the o-th element in destination order
  __shared__ uint16_t tabI[N];         // source-order number of the o-th element in destination order
  const int t = threadIdx.x;
  for (int e = t; e < N; e += 256) {
    uint32_t sa = 0, oc = 0, is = 0;
#pragma unroll
    for (int q = 0; q < kPermBits; ++q) {
      const uint32_t bs = (e >> q) & 1u;
      sa |= bs << P.src_t[q];
      oc |= bs << P.dst_t[q];
      is |= bs << P.s_of_d[q];
    }
    tabA[e] = sa;
    tabC[e] = oc;
    tabI[e] = (uint16_t)is;
  }
  __syncthreads();
  const float2* x = P.a_tensor ? A + blockIdx.y * a_bs : B + blockIdx.y * b_bs;
  const float2 s = P.a_tensor ? B[blockIdx.y * b_bs] : A[blockIdx.y * a_bs];
  float2* c = C + blockIdx.y * c_bs;
  for (int u = 0; u < kPermTilesPerGroup; ++u) {
    const uint32_t g = blockIdx.x * kPermTilesPerGroup + u;
    uint32_t baseA = 0, baseC = 0;
    for (int q = 0; q < P.n_rest; ++q) {
      const uint32_t bit = (g >> q) & 1u;
      baseA |= bit << P.src_r[q];
      baseC |= bit << P.dst_r[q];
    }
#pragma unroll
    for (int j = 0; j < N / 256; ++j) {
      const int i = t + 256 * j;
      tile[i + (i >> 5)] = x[baseA | tabA[i]];
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < N / 256; ++j) {
      const int o = t + 256 * j;
      const int i = tabI[o];
      const float2 v = tile[i + (i >> 5)];
      c[baseC | tabC[o]] = make_float2(v.x * s.x - v.y * s.y, v.x * s.y + v.y * s.x);
    }
    __syncthreads();
  }
}

// Tile and rest bits of a permutation step (role bits of the tensor operand:
// role bit j sits at source bit src[j] and destination bit dst[j]).
static TNPermPlan perm_plan(const TNDesc& d, int a_tensor) {
  TNPermPlan P;
  const int XR = a_tensor ? AM : BN, CR = a_tensor ? CM : CN;
  const int r = d.nb[a_tensor ? 1 : 2];
  int role_at_src[32], role_at_dst[32];
  for (int j = 0; j < r; ++j) {
    role_at_src[(int)d.pos[XR][j]] = j;
    role_at_dst[(int)d.pos[CR][j]] = j;
  }
  uint32_t in_tile = 0;
  for (int p = 0; p < 5; ++p) in_tile |= (1u << role_at_src[p]) | (1u << role_at_dst[p]);
  for (int p = 5; p < r && __builtin_popcount(in_tile) < kPermBits; ++p) in_tile |= 1u << role_at_dst[p];
  int nt = 0, nr = 0, ns = 0;
  int s_number[32];
  for (int p = 0; p < r; ++p) {
    const int j = role_at_src[p];
    if ((in_tile >> j) & 1u) {
      s_number[j] = ns;
      P.src_t[ns++] = (unsigned char)p;
    }
  }
  for (int p = 0; p < r; ++p) {
    const int j = role_at_dst[p];
    if ((in_tile >> j) & 1u) {
      P.s_of_d[nt] = (unsigned char)s_number[j];
      P.dst_t[nt++] = (unsigned char)p;
    } else {
      P.src_r[nr] = (unsigned char)d.pos[XR][j];
      P.dst_r[nr++] = (unsigned char)p;
    }
  }
  P.n_rest = nr;
  P.a_tensor = a_tensor;
  return P;
}

// Many small steps in one launch (the gate-merging steps of a contraction
// path: every operand a leaf or the result of another small step).  Workgroup
// (r, b) runs step r of the list for batch element b: outputs in C's address
// order over the threads, each summed over all 2^nK terms.  A record is eight
// int64: descriptor row, A base id, A element offset, B base id, B offset,
// C offset (in base 2), output count, reserved; a base id selects one of the
// three (pointer, batch stride) pairs passed by value.
struct TNBases {
  const float2* p[3];
  int64_t stride[3];
};

__global__ void __launch_bounds__(256) k_tn_small_multi(TNBases bases, float2* __restrict__ out_base,
                                                        const int64_t* __restrict__ recs,
                                                        const int32_t* __restrict__ descs) {
  __shared__ TNShared d;
  __shared__ signed char mapA[32], mapB[32];
  __shared__ uint32_t tka[64], tkb[64];
  const int64_t* rec = recs + 8 * (int64_t)blockIdx.x;
  const int row = (int)rec[0];
  load_desc(d, descs + (int64_t)row * (4 + 9 * 32));
  c_bit_maps(d, mapA, mapB);
  const int nH = d.nb[0], nM = d.nb[1], nN = d.nb[2], nK = d.nb[3];
  const int rc = nH + nM + nN;
  const int t = threadIdx.x;
  if (t < (1 << nK)) {
    tka[t] = scat(d, AK, nK, t);
    tkb[t] = scat(d, BK, nK, t);
  }
  __syncthreads();
  const int64_t b = blockIdx.y;
  const int ia = (int)rec[1], ib = (int)rec[3];
  const float2* a = bases.p[ia] + b * bases.stride[ia] + rec[2];
  const float2* bp = bases.p[ib] + b * bases.stride[ib] + rec[4];
  float2* c = out_base + b * bases.stride[2] + rec[5];
  const int n_out = (int)rec[6];
  const int kc = 1 << nK;
  for (int o = t; o < n_out; o += 256) {
    const uint32_t oa = via_map(mapA, rc, (uint32_t)o), ob = via_map(mapB, rc, (uint32_t)o);
    float2 acc = make_float2(0.f, 0.f);
    for (int k = 0; k < kc; ++k) {
      const float2 x = a[oa | tka[k]];
      const float2 y = bp[ob | tkb[k]];
      acc.x = fmaf(x.x, y.x, fmaf(-x.y, y.y, acc.x));
      acc.y = fmaf(x.x, y.y, fmaf(x.y, y.x, acc.y));
    }
    c[o] = acc;
  }
}

// Backward of a k_tn_gate step C[p, q] = sum_k X[p, k] G[q, k] in one pass:
// from the result's environment E_C (C's layout) it writes
// E_X[p, k] = sum_q E_C[p, q] G[q, k] (X's layout) and sums
// E_G[q, k] = sum_p E_C[p, q] X[p, k] (G's layout), reading E_C and X once
// (the two environment steps done separately read E_C twice).  Rows as in
// k_tn_gate; every workgroup leaves its partial E_G in part[] (the partials of
// one element contiguous, elements at G's offsets) and k_tn_env_sum adds the
// workgroups in a fixed order.
template <int NK, int NQ>
__global__ void __launch_bounds__(256) k_tn_gate_bwd(const float2* __restrict__ EC, const float2* __restrict__ X,
                                                     const float2* __restrict__ G, float2* __restrict__ EX,
                                                     float2* __restrict__ part, const int32_t* __restrict__ dd,
                                                     int64_t ec_bs, int64_t x_bs, int64_t g_bs, int64_t ex_bs,
                                                     int a_big) {
  constexpr int KN = 1 << NK, QN = 1 << NQ;
  __shared__ TNShared d;
  __shared__ uint32_t midX[4], midC[4];
  __shared__ float red[4][2 * QN * KN];
  load_desc(d, dd);
  const int XP = a_big ? AM : BN, XK = a_big ? AK : BK, GQ = a_big ? BN : AM, GK = a_big ? BK : AK;
  const int CP = a_big ? CM : CN, CQ = a_big ? CN : CM;
  const int nP = d.nb[a_big ? 1 : 2];
  const int t = threadIdx.x;
  if (t < 4) {
    uint32_t ox = 0, oc = 0;
    for (int j = 8; j < 10; ++j) {
      const uint32_t bit = ((uint32_t)t >> (j - 8)) & 1u;
      ox |= bit << d.pos[XP][j];
      oc |= bit << d.pos[CP][j];
    }
    midX[t] = ox;
    midC[t] = oc;
  }
  __syncthreads();
  const int64_t b = blockIdx.y;
  const float2* ec = EC + b * ec_bs;
  const float2* x = X + b * x_bs;
  const float2* g = G + b * g_bs;
  float2* ex = EX + b * ex_bs;
  float2 gm[QN][KN];
  uint32_t xk[KN], cq[QN], go[QN][KN];
#pragma unroll
  for (int k = 0; k < KN; ++k) xk[k] = scat(d, XK, NK, k);
#pragma unroll
  for (int q = 0; q < QN; ++q) {
    cq[q] = scat(d, CQ, NQ, q);
#pragma unroll
    for (int k = 0; k < KN; ++k) {
      go[q][k] = scat(d, GQ, NQ, q) | scat(d, GK, NK, k);
      gm[q][k] = g[go[q][k]];
    }
  }
  uint32_t loX = 0, loC = 0, hiX = 0, hiC = 0;
  for (int j = 0; j < 8; ++j) {
    const uint32_t bit = ((uint32_t)t >> j) & 1u;
    loX |= bit << d.pos[XP][j];
    loC |= bit << d.pos[CP][j];
  }
  const uint32_t blk = blockIdx.x;
  for (int j = 10; j < nP; ++j) {
    const uint32_t bit = (blk >> (j - 10)) & 1u;
    hiX |= bit << d.pos[XP][j];
    hiC |= bit << d.pos[CP][j];
  }
  float2 acc[QN][KN];
#pragma unroll
  for (int q = 0; q < QN; ++q)
#pragma unroll
    for (int k = 0; k < KN; ++k) acc[q][k] = make_float2(0.f, 0.f);
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const uint32_t xr = hiX | midX[j] | loX, cr = hiC | midC[j] | loC;
    float2 e[QN], v[KN];
#pragma unroll
    for (int q = 0; q < QN; ++q) e[q] = ec[cr | cq[q]];
#pragma unroll
    for (int k = 0; k < KN; ++k) v[k] = x[xr | xk[k]];
#pragma unroll
    for (int k = 0; k < KN; ++k) {
      float2 s = make_float2(0.f, 0.f);
#pragma unroll
      for (int q = 0; q < QN; ++q) {
        s.x = fmaf(e[q].x, gm[q][k].x, fmaf(-e[q].y, gm[q][k].y, s.x));
        s.y = fmaf(e[q].x, gm[q][k].y, fmaf(e[q].y, gm[q][k].x, s.y));
        acc[q][k].x = fmaf(e[q].x, v[k].x, fmaf(-e[q].y, v[k].y, acc[q][k].x));
        acc[q][k].y = fmaf(e[q].x, v[k].y, fmaf(e[q].y, v[k].x, acc[q][k].y));
      }
      ex[xr | xk[k]] = s;
    }
  }
  const int lane = t & 63, w = t >> 6;
#pragma unroll
  for (int q = 0; q < QN; ++q)
#pragma unroll
    for (int k = 0; k < KN; ++k) {
      float re = acc[q][k].x, im = acc[q][k].y;
#pragma unroll
      for (int o = 32; o >= 1; o >>= 1) {
        re += __shfl_xor(re, o, 64);
        im += __shfl_xor(im, o, 64);
      }
      if (lane == 0) {
        red[w][2 * (q * KN + k)] = re;
        red[w][2 * (q * KN + k) + 1] = im;
      }
    }
  __syncthreads();
  if (t < QN * KN) {
    const int q = t / KN, k = t % KN;
    float2 s;
    s.x = (red[0][2 * t] + red[1][2 * t]) + (red[2][2 * t] + red[3][2 * t]);
    s.y = (red[0][2 * t + 1] + red[1][2 * t + 1]) + (red[2][2 * t + 1] + red[3][2 * t + 1]);
    part[(b * (QN * KN) + go[q][k]) * gridDim.x + blk] = s;
  }
}

__global__ void k_tn_split_sum(const float2* __restrict__ part, float2* __restrict__ C, int64_t c_bs, int split_log,
                               int64_t n_out) {
  const int64_t o = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (o >= n_out) return;
  const int64_t b = blockIdx.y;
  float2 acc = make_float2(0.f, 0.f);
  const int S = 1 << split_log;
  for (int s = 0; s < S; ++s) {
    const float2 v = part[(b * S + s) * n_out + o];
    acc.x += v.x;
    acc.y += v.y;
  }
  C[b * c_bs + o] = acc;
}

// Environment-sized results (at most 16 products per summed term, at least
// 2^12 summed terms): the gradient contractions of gate tensors, where a
// large environment and a large intermediate are summed over everything but
// the gate's few indices.  Every thread reads its k elements of A (for all
// 2^LA values of its (h, m) bits) and of B (all 2^LB values of (h, n)) once
// and accumulates every output product in registers, so each operand byte is
// read once (a gather per output element would re-read A 2^nN times and B
// 2^nM times).  Workgroup s owns summed terms s * 2^(nK - split_log) .. +;
// thread t takes k = (s << kw) | (it << 8) | t, so the k offsets come from a
// per-thread part (bits 0..7), an LDS table over the iteration (bits 8..kw-1)
// and a per-workgroup part.  With one workgroup per batch element
// (split_log 0) the sums go straight to C; otherwise to part[] (the partials
// of one output contiguous, outputs in C's address order), and k_tn_env_sum
// adds the workgroups in a fixed order.
template <int LA, int LB>
__global__ void __launch_bounds__(256) k_tn_env(const float2* __restrict__ A, const float2* __restrict__ B,
                                                float2* __restrict__ C, float2* __restrict__ part,
                                                const int32_t* __restrict__ dd, int64_t a_bs, int64_t b_bs,
                                                int64_t c_bs, int split_log, int64_t n_out) {
  constexpr int NA = 1 << LA, NB = 1 << LB;
  __shared__ TNShared d;
  __shared__ uint32_t titA[256], titB[256];
  __shared__ uint32_t offA[NA], offB[NB], offC[NA * NB];
  __shared__ float red[4][2 * NA * NB];
  load_desc(d, dd);
  const int nH = d.nb[0], nM = d.nb[1], nN = d.nb[2], nK = d.nb[3];
  const int t = threadIdx.x;
  const int kw = nK - split_log;      // bits of k one workgroup covers (>= 12)
  const int iters = 1 << (kw - 8);    // <= 256
  const uint32_t hmask = (1u << nH) - 1;
  if (t < iters) {
    uint32_t oa = 0, ob = 0;
    for (int j = 8; j < kw; ++j) {
      const uint32_t bit = ((uint32_t)t >> (j - 8)) & 1u;
      oa |= bit << d.pos[AK][j];
      ob |= bit << d.pos[BK][j];
    }
    titA[t] = oa;
    titB[t] = ob;
  }
  if (t < NA) offA[t] = scat(d, AH, nH, t & hmask) | scat(d, AM, nM, t >> nH);
  if (t >= 64 && t < 64 + NB) offB[t - 64] = scat(d, BH, nH, (t - 64) & hmask) | scat(d, BN, nN, (t - 64) >> nH);
  if (t >= 128 && t < 128 + NA * NB) {
    const uint32_t i = (t - 128) / NB, j = (t - 128) % NB;
    offC[t - 128] = scat(d, CH, nH, i & hmask) | scat(d, CM, nM, i >> nH) | scat(d, CN, nN, j >> nH);
  }
  __syncthreads();
  const uint32_t s = blockIdx.x;
  uint32_t kA = 0, kB = 0;
  for (int j = 0; j < 8; ++j) {
    const uint32_t bit = ((uint32_t)t >> j) & 1u;
    kA |= bit << d.pos[AK][j];
    kB |= bit << d.pos[BK][j];
  }
  for (int j = kw; j < nK; ++j) {
    const uint32_t bit = (s >> (j - kw)) & 1u;
    kA |= bit << d.pos[AK][j];
    kB |= bit << d.pos[BK][j];
  }
  const int64_t b = blockIdx.y;
  const float2* a = A + b * a_bs;
  const float2* bp = B + b * b_bs;
  uint32_t oa[NA], ob[NB];
#pragma unroll
  for (int i = 0; i < NA; ++i) oa[i] = offA[i] | kA;
#pragma unroll
  for (int j = 0; j < NB; ++j) ob[j] = offB[j] | kB;
  float2 acc[NA][NB];
#pragma unroll
  for (int i = 0; i < NA; ++i)
#pragma unroll
    for (int j = 0; j < NB; ++j) acc[i][j] = make_float2(0.f, 0.f);
  for (int it = 0; it < iters; ++it) {
    const uint32_t ta = titA[it], tb = titB[it];
    float2 x[NA], y[NB];
#pragma unroll
    for (int i = 0; i < NA; ++i) x[i] = a[oa[i] | ta];
#pragma unroll
    for (int j = 0; j < NB; ++j) y[j] = bp[ob[j] | tb];
#pragma unroll
    for (int i = 0; i < NA; ++i)
#pragma unroll
      for (int j = 0; j < NB; ++j) {
        acc[i][j].x = fmaf(x[i].x, y[j].x, fmaf(-x[i].y, y[j].y, acc[i][j].x));
        acc[i][j].y = fmaf(x[i].x, y[j].y, fmaf(x[i].y, y[j].x, acc[i][j].y));
      }
  }
  // wave sums, then the four waves in LDS
  const int lane = t & 63, w = t >> 6;
#pragma unroll
  for (int i = 0; i < NA; ++i)
#pragma unroll
    for (int j = 0; j < NB; ++j) {
      float re = acc[i][j].x, im = acc[i][j].y;
#pragma unroll
      for (int o = 32; o >= 1; o >>= 1) {
        re += __shfl_xor(re, o, 64);
        im += __shfl_xor(im, o, 64);
      }
      if (lane == 0) {
        red[w][2 * (i * NB + j)] = re;
        red[w][2 * (i * NB + j) + 1] = im;
      }
    }
  __syncthreads();
  if (t < NA * NB) {
    const uint32_t i = t / NB, j = t % NB;
    if ((i & hmask) == (j & hmask)) {
      float2 v;
      v.x = (red[0][2 * t] + red[1][2 * t]) + (red[2][2 * t] + red[3][2 * t]);
      v.y = (red[0][2 * t + 1] + red[1][2 * t + 1]) + (red[2][2 * t + 1] + red[3][2 * t + 1]);
      if (split_log == 0)
        C[b * c_bs + offC[t]] = v;
      else
        part[((b * n_out + offC[t]) << split_log) + s] = v;
    }
  }
}

// Sum of the k_tn_env workgroup partials: one workgroup per (output element,
// batch element), 256 threads over the 2^split_log partials, then a fixed
// tree in LDS (deterministic; a thread per output would be a serial chain).
__global__ void __launch_bounds__(256) k_tn_env_sum(const float2* __restrict__ part, float2* __restrict__ C,
                                                    int64_t c_bs, int split_log, int64_t n_out) {
  __shared__ float2 red[256];
  const int t = threadIdx.x;
  const int64_t o = blockIdx.x, b = blockIdx.y;
  const int S = 1 << split_log;
  float2 acc = make_float2(0.f, 0.f);
  const float2* p = part + (b * n_out + o) * S;  // one output's partials are contiguous
  for (int s = t; s < S; s += 256) {
    const float2 v = p[s];
    acc.x += v.x;
    acc.y += v.y;
  }
  red[t] = acc;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if (t < h) {
      red[t].x += red[t + h].x;
      red[t].y += red[t + h].y;
    }
    __syncthreads();
  }
  if (t == 0) C[b * c_bs + o] = red[0];
}

// 64x64 output tile of one (batch, h); k staged 16 at a time through LDS; MFMA.
constexpr int kTile = 64;
constexpr int kChunk = 16;
typedef float f32x4 __attribute__((ext_vector_type(4)));

// With split_log > 0 the k range is split over 2^split_log workgroups per
// tile (few tiles, long sums); their partial tiles go to part[] in C's
// address order and k_tn_split_sum adds them in a fixed order.
template <bool THREE_M>
__global__ void __launch_bounds__(256) k_tn_gemm(const float2* __restrict__ A, const float2* __restrict__ B,
                                                 float2* __restrict__ C, const int32_t* __restrict__ dd,
                                                 int64_t a_bs, int64_t b_bs, int64_t c_bs, int m_tiles,
                                                 float2* __restrict__ part, int split_log, int64_t n_out) {
  __shared__ float2 As[kChunk][kTile + 20];  // row stride 168 words, 8 mod 32 banks (see kBigPad)
  __shared__ float2 Bs[kChunk][kTile + 20];
  __shared__ uint32_t rowA[kTile], rowC[kTile], colB[kTile], colC[kTile];
  __shared__ uint32_t ka[kChunk], kb[kChunk];
  __shared__ TNShared d;
  load_desc(d, dd);
  const int nH = d.nb[0], nM = d.nb[1], nN = d.nb[2], nK = d.nb[3];
  const int t = threadIdx.x;
  const int64_t b = blockIdx.z;
  const uint32_t h = blockIdx.y;
  const uint32_t split = blockIdx.x & ((1u << split_log) - 1), tile = blockIdx.x >> split_log;
  const int mt = tile % m_tiles, nt = tile / m_tiles;
  const uint32_t m0 = (uint32_t)mt * kTile, n0 = (uint32_t)nt * kTile;
  const uint32_t msz = 1u << nM, nsz = 1u << nN;
  if (t < kTile) {
    const uint32_t m = m0 + t;
    rowA[t] = scat(d, AH, nH, h) | scat(d, AM, nM, m < msz ? m : 0);
    rowC[t] = scat(d, CH, nH, h) | scat(d, CM, nM, m < msz ? m : 0);
  } else if (t < 2 * kTile) {
    const uint32_t n = n0 + (t - kTile);
    colB[t - kTile] = scat(d, BH, nH, h) | scat(d, BN, nN, n < nsz ? n : 0);
    colC[t - kTile] = scat(d, CN, nN, n < nsz ? n : 0);
  }
  const float2* a = A + b * a_bs;
  const float2* bp = B + b * b_bs;
  // MFMA layout (v_mfma_f32_16x16x4_f32): lane l supplies A[l % 16][l / 16]
  // and B[l / 16][l % 16] and receives D[4 (l / 16) + v][l % 16] in register v.
  // Wave w owns output rows 32 (w & 1) .. +32 and columns 32 (w >> 1) .. +32
  // as 2 x 2 blocks of 16 x 16; complex products as four real MFMAs
  // (re += ar br - ai bi, im += ar bi + ai br).
  const int lane = t & 63, w = t >> 6;
  const int wm = (w & 1) * 32, wn = (w >> 1) * 32;
  const int lr = lane & 15, lk = lane >> 4;
  // 4M: cr, ci are the real and imaginary parts; 3M: cr, ci, c3 hold p1, p2, p3
  f32x4 cr[2][2], ci[2][2], c3[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      cr[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
      ci[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
      c3[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
    }
  const uint32_t kseg = (1u << nK) >> split_log, kbeg = split * kseg, ksz = kbeg + kseg;
  for (uint32_t k0 = kbeg; k0 < ksz; k0 += kChunk) {
    if (t >= 2 * kTile && t < 2 * kTile + kChunk) {
      const uint32_t k = k0 + (t - 2 * kTile);
      ka[t - 2 * kTile] = scat(d, AK, nK, k < ksz ? k : 0);
    } else if (t >= 2 * kTile + kChunk && t < 2 * kTile + 2 * kChunk) {
      const uint32_t k = k0 + (t - 2 * kTile - kChunk);
      kb[t - 2 * kTile - kChunk] = scat(d, BK, nK, k < ksz ? k : 0);
    }
    __syncthreads();
    // k runs fastest across threads: the host lays GEMM operands out with the
    // summed indices as their lowest address bits, so 16 consecutive threads
    // read one contiguous 128-byte run of a row
#pragma unroll
    for (int r = 0; r < kTile * kChunk / 256; ++r) {
      const int e = t + 256 * r;
      const int kk = e & (kChunk - 1), mi = e >> 4;
      const bool ok = (m0 + mi < msz) && (k0 + kk < ksz);
      As[kk][mi] = ok ? a[rowA[mi] | ka[kk]] : make_float2(0.f, 0.f);
      const bool okb = (n0 + mi < nsz) && (k0 + kk < ksz);
      Bs[kk][mi] = okb ? bp[colB[mi] | kb[kk]] : make_float2(0.f, 0.f);
    }
    __syncthreads();
#pragma unroll
    for (int ks = 0; ks < kChunk; ks += 4) {
      float2 x[2], y[2];
#pragma unroll
      for (int i = 0; i < 2; ++i) x[i] = As[ks + lk][wm + 16 * i + lr];
#pragma unroll
      for (int j = 0; j < 2; ++j) y[j] = Bs[ks + lk][wn + 16 * j + lr];
      if (THREE_M) {
        // 3M: p1 += ar br, p2 += ai bi, p3 += (ar + ai)(br + bi); re = p1 - p2, im = p3 - p1 - p2
        float xs[2], ys[2];
#pragma unroll
        for (int i = 0; i < 2; ++i) xs[i] = x[i].x + x[i].y;
#pragma unroll
        for (int j = 0; j < 2; ++j) ys[j] = y[j].x + y[j].y;
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
          for (int j = 0; j < 2; ++j) cr[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(x[i].x, y[j].x, cr[i][j], 0, 0, 0);
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
          for (int j = 0; j < 2; ++j) ci[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(x[i].y, y[j].y, ci[i][j], 0, 0, 0);
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
          for (int j = 0; j < 2; ++j) c3[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(xs[i], ys[j], c3[i][j], 0, 0, 0);
      } else {
        // 4M, one term at a time over the four blocks, so that consecutive
        // MFMAs never wait on each other's accumulator
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
          for (int j = 0; j < 2; ++j) cr[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(x[i].x, y[j].x, cr[i][j], 0, 0, 0);
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
          for (int j = 0; j < 2; ++j) ci[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(x[i].x, y[j].y, ci[i][j], 0, 0, 0);
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
          for (int j = 0; j < 2; ++j) cr[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(-x[i].y, y[j].y, cr[i][j], 0, 0, 0);
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
          for (int j = 0; j < 2; ++j) ci[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(x[i].y, y[j].x, ci[i][j], 0, 0, 0);
      }
    }
    __syncthreads();
  }
  float2* c = C + b * c_bs;
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const uint32_t mi = wm + 16 * i + 4 * lk + v;
        const uint32_t ni = wn + 16 * j + lr;
        if (m0 + mi < msz && n0 + ni < nsz) {
          const float re = THREE_M ? cr[i][j][v] - ci[i][j][v] : cr[i][j][v];
          const float im = THREE_M ? c3[i][j][v] - cr[i][j][v] - ci[i][j][v] : ci[i][j][v];
          if (split_log == 0)
            c[rowC[mi] | colC[ni]] = make_float2(re, im);
          else
            part[((b << split_log) + split) * n_out + (rowC[mi] | colC[ni])] = make_float2(re, im);
        }
      }
}

// 128 x 128 output tiles for steps with m and n at least 128 wide (whole
// tiles, no bounds checks): twice the operand reuse of the 64 x 64 form (half
// the bytes read per multiply-add) and 256 MFMAs per wave per k chunk, with
// the next chunk's operands fetched into registers while the current one is
// multiplied (double-buffered in LDS).  Wave w owns rows 64 (w & 1) .. +64
// and columns 64 (w >> 1) .. +64 as 4 x 4 blocks of 16 x 16; measured
// 107-114 TFLOP/s (0.68-0.72 of the FP32 matrix peak) on 2^10-2^12 cubes.
// Each thread loads one k column (t % 16) of eight rows (t / 16 + 16 r) of
// both operands per chunk.  KID: the summed indices are the lowest address
// bits of both operands in ascending order (the layout the host gives GEMM
// operands), so a chunk's k offsets need no scatter.  Requires nM, nN >= 7
// and nK >= 4.  Splits k over workgroups as k_tn_gemm does.
constexpr int kBig = 128;
// LDS row stride of 148 complex = 296 words.  A 16-lane pass of the tile stores covers the 16 k
// rows of one m: with the earlier stride of 288 words (0 mod 32 banks) all of them fell on one
// bank pair and the kernel ran at 114 TFLOP/s on a 4096^3 product; at 296 words (8 mod 32) it runs
// at 138.  Odd strides (258, 262 words: conflict-free stores and reads in that model) measured
// within 3% of this one.
constexpr int kBigPad = kBig + 20;

template <bool KID>
__global__ void __launch_bounds__(256) k_tn_gemm128(const float2* __restrict__ A, const float2* __restrict__ B,
                                                    float2* __restrict__ C, const int32_t* __restrict__ dd,
                                                    int64_t a_bs, int64_t b_bs, int64_t c_bs, int m_tiles,
                                                    float2* __restrict__ part, int split_log, int64_t n_out) {
  __shared__ float2 As2[2][kChunk][kBigPad];
  __shared__ float2 Bs2[2][kChunk][kBigPad];
  __shared__ uint32_t rowC[kBig], colC[kBig];
  __shared__ TNShared d;
  load_desc(d, dd);
  const int nH = d.nb[0], nM = d.nb[1], nN = d.nb[2], nK = d.nb[3];
  const int t = threadIdx.x;
  const int64_t b = blockIdx.z;
  const uint32_t h = blockIdx.y;
  const uint32_t split = blockIdx.x & ((1u << split_log) - 1), tile = blockIdx.x >> split_log;
  const int mt = tile % m_tiles, nt = tile / m_tiles;
  const uint32_t m0 = (uint32_t)mt * kBig, n0 = (uint32_t)nt * kBig;
  const uint32_t kseg = (1u << nK) >> split_log, kbeg = split * kseg, kend = kbeg + kseg;
  if (t < kBig)
    rowC[t] = scat(d, CH, nH, h) | scat(d, CM, nM, m0 + t);
  else
    colC[t - kBig] = scat(d, CN, nN, n0 + (t - kBig));
  const float2* a = A + b * a_bs;
  const float2* bp = B + b * b_bs;
  const int kk = t & (kChunk - 1), r0 = t >> 4;
  uint32_t ra[8], rb[8];
  const uint32_t ha = scat(d, AH, nH, h), hb = scat(d, BH, nH, h);
#pragma unroll
  for (int r = 0; r < 8; ++r) {
    ra[r] = ha | scat(d, AM, nM, m0 + r0 + 16 * r);
    rb[r] = hb | scat(d, BN, nN, n0 + r0 + 16 * r);
  }
  float2 pa[8], pb[8];
  auto fetch = [&](uint32_t k0) {
    const uint32_t ka = KID ? k0 + kk : scat(d, AK, nK, k0 + kk);
    const uint32_t kb = KID ? k0 + kk : scat(d, BK, nK, k0 + kk);
#pragma unroll
    for (int r = 0; r < 8; ++r) {
      pa[r] = a[ra[r] | ka];
      pb[r] = bp[rb[r] | kb];
    }
  };
  const int lane = t & 63, w = t >> 6;
  const int wm = (w & 1) * 64, wn = (w >> 1) * 64;
  const int lr = lane & 15, lk = lane >> 4;
  f32x4 cr[4][4], ci[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      cr[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
      ci[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
    }
  // LDS holds two chunks: chunk i is multiplied from buffer i % 2 while chunk
  // i + 1 (fetched into registers meanwhile) goes into the other buffer, so
  // one barrier per chunk suffices
  fetch(kbeg);
#pragma unroll
  for (int r = 0; r < 8; ++r) {
    As2[0][kk][r0 + 16 * r] = pa[r];
    Bs2[0][kk][r0 + 16 * r] = pb[r];
  }
  __syncthreads();
  int cur = 0;
  for (uint32_t k0 = kbeg; k0 < kend; k0 += kChunk) {
    // unconditional (the last chunk re-reads the first, unused), so that the
    // loads stay in the same block as, and ahead of, the MFMAs below
    fetch(k0 + kChunk < kend ? k0 + kChunk : kbeg);
    asm volatile("" ::: "memory");  // issue the loads before the MFMAs, not after them
    float2(*As)[kBigPad] = As2[cur];
    float2(*Bs)[kBigPad] = Bs2[cur];
#pragma unroll
    for (int ks = 0; ks < kChunk; ks += 4) {
      float2 x[4], y[4];
      float nx[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        x[i] = As[ks + lk][wm + 16 * i + lr];
        nx[i] = -x[i].y;
      }
#pragma unroll
      for (int j = 0; j < 4; ++j) y[j] = Bs[ks + lk][wn + 16 * j + lr];
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) cr[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(x[i].x, y[j].x, cr[i][j], 0, 0, 0);
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) ci[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(x[i].x, y[j].y, ci[i][j], 0, 0, 0);
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) cr[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(nx[i], y[j].y, cr[i][j], 0, 0, 0);
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) ci[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(x[i].y, y[j].x, ci[i][j], 0, 0, 0);
    }
    cur ^= 1;
#pragma unroll
    for (int r = 0; r < 8; ++r) {
      As2[cur][kk][r0 + 16 * r] = pa[r];
      Bs2[cur][kk][r0 + 16 * r] = pb[r];
    }
    __syncthreads();
  }
  float2* c = C + b * c_bs;
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const uint32_t o = rowC[wm + 16 * i + 4 * lk + v] | colC[wn + 16 * j + lr];
        if (split_log == 0)
          c[o] = make_float2(cr[i][j][v], ci[i][j][v]);
        else
          part[((b << split_log) + split) * n_out + o] = make_float2(cr[i][j][v], ci[i][j][v]);
      }
}

// ---------------------------------------------------------------------------
// launch planning shared by the workspace query and the launch
// ---------------------------------------------------------------------------
struct TNLaunch {
  bool gemm;
  bool stream;
  bool env;
  int lpo_log, split_log;
  int64_t n_out;
};

static int parse_desc(const int32_t* desc, TNDesc* d) {
  d->gemm_form = desc[kDescForm];
  if (d->gemm_form != 0 && d->gemm_form != 3 && d->gemm_form != 64) return -1;
  for (int r = 0; r < 4; ++r) {
    if (desc[r] < 0 || desc[r] > kMaxRank) return -1;
    d->nb[r] = desc[r];
  }
  const int nH = d->nb[0], nM = d->nb[1], nN = d->nb[2], nK = d->nb[3];
  const int roles[9] = {nH, nM, nK, nH, nK, nN, nH, nM, nN};
  const int rankA = nH + nM + nK, rankB = nH + nK + nN, rankC = nH + nM + nN;
  if (rankA > kMaxRank || rankB > kMaxRank || rankC > kMaxRank) return -1;
  const int ranks[3] = {rankA, rankB, rankC};
  for (int t = 0; t < 3; ++t) {
    uint32_t seen = 0;
    for (int r = 0; r < 3; ++r) {
      const int q = 3 * t + r;
      for (int j = 0; j < 32; ++j) {
        const int p = desc[4 + q * 32 + j];
        if (j < roles[q]) {
          if (p < 0 || p >= ranks[t] || (seen >> p) & 1u) return -1;
          seen |= 1u << p;
          d->pos[q][j] = (signed char)p;
        } else {
          d->pos[q][j] = 0;
        }
      }
    }
    if (ranks[t] < 32 && seen != ((1u << ranks[t]) - 1)) return -1;
  }
  return 0;
}

static int gemm_tile(const TNDesc& d) { return (d.gemm_form == 0 && d.nb[1] >= 7 && d.nb[2] >= 7) ? 128 : 64; }

static TNLaunch plan_launch(const TNDesc& d, int64_t batch) {
  TNLaunch L;
  const int nH = d.nb[0], nM = d.nb[1], nN = d.nb[2], nK = d.nb[3];
  L.n_out = 1ll << (nH + nM + nN);
  L.gemm = nM >= 6 && nN >= 6 && nK >= 4;
  // a large tensor times a gate-sized one (the host keeps the large tensor's layout for such steps)
  L.stream = !L.gemm && nK <= 4 && nH + nM + nN >= 12 && (nH + nK + nN <= 8 || nH + nM + nK <= 8);
  L.lpo_log = 0;
  L.split_log = 0;
  if (L.gemm) {
    // fewer tiles than about one round of workgroups (the 128 x 128 kernel holds two per CU, the
    // 64 x 64 one five) and a long sum: split k until there are that many (segments of >= 64 terms);
    // the environment steps of the backward have such shapes (256 x 128 results summed over 2^17)
    const int64_t tile = gemm_tile(d);
    const int64_t round = tile == 128 ? 512 : 2048;
    const int64_t wgs = (((1ll << nM) + tile - 1) / tile) * (((1ll << nN) + tile - 1) / tile) * (1ll << nH) * batch;
    while ((wgs << L.split_log) < round && L.split_log < 11 && nK - L.split_log > 6) ++L.split_log;
  }
  if (L.gemm || L.stream) {
    L.env = false;
    return L;
  }
  // a gate-sized result summed over a large rest (k_tn_env): workgroups of 2^12 .. 2^16 summed terms,
  // at most 2^11 of them
  L.env = (nH + nM) + (nH + nN) <= 4 && nK >= 12;
  if (L.env) {
    const int lo = nK - 16, hi = nK - 12 < 11 ? nK - 12 : 11;
    L.split_log = lo > hi ? lo : hi;
    return L;
  }
  // aim for >= 2^19 threads in flight and <= 2^12 k iterations per thread
  const int64_t want = 1ll << 19;
  int64_t threads = L.n_out * batch;
  while (L.lpo_log < 6 && L.lpo_log < nK && threads < want) {
    ++L.lpo_log;
    threads *= 2;
  }
  while (L.split_log < nK - L.lpo_log - kTabBits && (threads < want || nK - L.lpo_log - L.split_log > 12)) {
    ++L.split_log;
    threads *= 2;
  }
  return L;
}

}  // namespace tedq

using namespace tedq;

// desc: 4 + 9*32 int32 = bit counts of (h, m, n, k), then the address bit of
// every role bit for A_h A_m A_k B_h B_k B_n C_h C_m C_n (32 slots per role).
// tedq_tn_contract takes it twice: in host memory (validated, used to size
// the launch) and as an identical copy in device memory (read by the kernels).
extern "C" int64_t tedq_tn_workspace_bytes(const int32_t* desc, int64_t batch) {
  TNDesc d;
  if (desc == nullptr || batch <= 0 || parse_desc(desc, &d) != 0) return -1;
  const TNLaunch L = plan_launch(d, batch);
  if (L.split_log == 0) return 0;
  return (int64_t)sizeof(float2) * batch * (L.n_out << L.split_log);
}

extern "C" int tedq_tn_contract(const void* a, int64_t a_bstride, const void* b, int64_t b_bstride, void* c,
                                int64_t c_bstride, int64_t batch, const int32_t* desc, const int32_t* desc_dev,
                                void* workspace, int64_t workspace_bytes, void* stream) {
  TNDesc d;
  if (a == nullptr || b == nullptr || c == nullptr || desc == nullptr || desc_dev == nullptr || batch <= 0 ||
      batch > 65535) {
    set_error("tedq_tn_contract: bad arguments (batch=%lld)", (long long)batch);
    return kBadArgument;
  }
  if (parse_desc(desc, &d) != 0) {
    set_error("tedq_tn_contract: inconsistent index layout (ranks above %d or bits not a permutation)", kMaxRank);
    return kBadArgument;
  }
  const TNLaunch L = plan_launch(d, batch);
  hipStream_t st = static_cast<hipStream_t>(stream);
  const float2* A = static_cast<const float2*>(a);
  const float2* Bp = static_cast<const float2*>(b);
  float2* Cp = static_cast<float2*>(c);
  if (L.gemm) {
    const int form = d.gemm_form;
    const int tile = gemm_tile(d);
    const int64_t need = L.split_log ? (int64_t)sizeof(float2) * batch * (L.n_out << L.split_log) : 0;
    if (need > 0 && (workspace == nullptr || workspace_bytes < need)) {
      set_error("tedq_tn_contract: workspace of %lld bytes required, %lld given", (long long)need,
                (long long)workspace_bytes);
      return kBadArgument;
    }
    const int m_tiles = (int)(((1ll << d.nb[1]) + tile - 1) / tile);
    const int64_t n_tiles = ((1ll << d.nb[2]) + tile - 1) / tile;
    const int64_t gx = m_tiles * n_tiles;
    const int64_t gy = 1ll << d.nb[0];
    if ((gx << L.split_log) > 0x7fffffffll || gy > 65535) {
      set_error("tedq_tn_contract: grid too large (%lld x %lld)", (long long)gx, (long long)gy);
      return kBadArgument;
    }
    bool kid = true;
    for (int j = 0; j < d.nb[3]; ++j) kid = kid && d.pos[AK][j] == j && d.pos[BK][j] == j;
    if (tile == kBig && kid)
      hipLaunchKernelGGL(k_tn_gemm128<true>, dim3((unsigned)(gx << L.split_log), (unsigned)gy, (unsigned)batch),
                         dim3(256), 0, st, A, Bp, Cp, desc_dev, a_bstride, b_bstride, c_bstride, m_tiles,
                         static_cast<float2*>(workspace), L.split_log, L.n_out);
    else if (tile == kBig)
      hipLaunchKernelGGL(k_tn_gemm128<false>, dim3((unsigned)(gx << L.split_log), (unsigned)gy, (unsigned)batch),
                         dim3(256), 0, st, A, Bp, Cp, desc_dev, a_bstride, b_bstride, c_bstride, m_tiles,
                         static_cast<float2*>(workspace), L.split_log, L.n_out);
    else if (form != 3)
      hipLaunchKernelGGL(k_tn_gemm<false>, dim3((unsigned)(gx << L.split_log), (unsigned)gy, (unsigned)batch),
                         dim3(256), 0, st, A, Bp, Cp, desc_dev, a_bstride, b_bstride, c_bstride, m_tiles,
                         static_cast<float2*>(workspace), L.split_log, L.n_out);
    else
      hipLaunchKernelGGL(k_tn_gemm<true>, dim3((unsigned)(gx << L.split_log), (unsigned)gy, (unsigned)batch),
                         dim3(256), 0, st, A, Bp, Cp, desc_dev, a_bstride, b_bstride, c_bstride, m_tiles,
                         static_cast<float2*>(workspace), L.split_log, L.n_out);
    int rc = check_launch("tedq_tn_contract(gemm)");
    if (rc != kOk || L.split_log == 0) return rc;
    const int64_t sb = (L.n_out + 255) / 256;
    hipLaunchKernelGGL(k_tn_split_sum, dim3((unsigned)sb, (unsigned)batch), dim3(256), 0, st,
                       static_cast<const float2*>(workspace), Cp, c_bstride, L.split_log, L.n_out);
    return check_launch("tedq_tn_contract(split sum)");
  }
  if (d.nb[0] == 0 && d.nb[3] == 0 && (d.nb[1] == 0 || d.nb[2] == 0) && d.nb[1] + d.nb[2] >= kPermBits + 2) {
    // a tensor times a scalar: a layout change through LDS tiles (k_tn_permute)
    const int a_tensor = d.nb[1] > 0 ? 1 : 0;
    const int r = d.nb[1] + d.nb[2];
    const TNPermPlan P = perm_plan(d, a_tensor);
    const unsigned groups = (1u << (r - kPermBits)) / kPermTilesPerGroup;
    hipLaunchKernelGGL(k_tn_permute, dim3(groups, (unsigned)batch), dim3(256), 0, st, A, Bp, Cp, a_bstride,
                       b_bstride, c_bstride, P);
    return check_launch("tedq_tn_contract(permute)");
  }
  if (L.stream && d.nb[0] == 0 && d.nb[3] <= 2 && ((d.nb[2] <= 2 && d.nb[1] >= 10) || (d.nb[1] <= 2 && d.nb[2] >= 10))) {
    // a gate of at most two summed and two new indices on a large tensor: k_tn_gate
    const int a_big = d.nb[1] >= 10 ? 1 : 0;
    const int nq = a_big ? d.nb[2] : d.nb[1], nk = d.nb[3], np = a_big ? d.nb[1] : d.nb[2];
    const dim3 grid(1u << (np - 10), (unsigned)batch);
#define TEDQ_GATE_CASE(K, Q)                                                                                  \
  else if (nk == K && nq == Q) hipLaunchKernelGGL((k_tn_gate<K, Q>), grid, dim3(256), 0, st, A, Bp, Cp, desc_dev, \
                                                  a_bstride, b_bstride, c_bstride, a_big);
    if (false) {
    }
    TEDQ_GATE_CASE(0, 0) TEDQ_GATE_CASE(0, 1) TEDQ_GATE_CASE(0, 2)
    TEDQ_GATE_CASE(1, 0) TEDQ_GATE_CASE(1, 1) TEDQ_GATE_CASE(1, 2)
    TEDQ_GATE_CASE(2, 0) TEDQ_GATE_CASE(2, 1) TEDQ_GATE_CASE(2, 2)
#undef TEDQ_GATE_CASE
    return check_launch("tedq_tn_contract(gate)");
  }
  if (L.stream) {
    const int64_t blocks = L.n_out >> 12;
    if (blocks > 0x7fffffffll) {
      set_error("tedq_tn_contract: grid too large (%lld blocks)", (long long)blocks);
      return kBadArgument;
    }
    hipLaunchKernelGGL(k_tn_stream, dim3((unsigned)blocks, (unsigned)batch), dim3(256), 0, st, A, Bp, Cp, desc_dev,
                       a_bstride, b_bstride, c_bstride);
    return check_launch("tedq_tn_contract(stream)");
  }
  const int64_t need = L.split_log ? (int64_t)sizeof(float2) * batch * (L.n_out << L.split_log) : 0;
  if (need > 0 && (workspace == nullptr || workspace_bytes < need)) {
    set_error("tedq_tn_contract: workspace of %lld bytes required, %lld given", (long long)need,
              (long long)workspace_bytes);
    return kBadArgument;
  }
  float2* part = static_cast<float2*>(workspace);
  if (L.env) {
    const int la = d.nb[0] + d.nb[1], lb = d.nb[0] + d.nb[2];
    const dim3 grid(1u << L.split_log, (unsigned)batch);
#define TEDQ_ENV_CASE(X, Y)                                                                                       \
  else if (la == X && lb == Y) hipLaunchKernelGGL((k_tn_env<X, Y>), grid, dim3(256), 0, st, A, Bp, Cp, part,      \
                                                  desc_dev, a_bstride, b_bstride, c_bstride, L.split_log, L.n_out);
    if (false) {
    }
    TEDQ_ENV_CASE(0, 0) TEDQ_ENV_CASE(0, 1) TEDQ_ENV_CASE(0, 2) TEDQ_ENV_CASE(0, 3) TEDQ_ENV_CASE(0, 4)
    TEDQ_ENV_CASE(1, 0) TEDQ_ENV_CASE(1, 1) TEDQ_ENV_CASE(1, 2) TEDQ_ENV_CASE(1, 3)
    TEDQ_ENV_CASE(2, 0) TEDQ_ENV_CASE(2, 1) TEDQ_ENV_CASE(2, 2)
    TEDQ_ENV_CASE(3, 0) TEDQ_ENV_CASE(3, 1)
    TEDQ_ENV_CASE(4, 0)
#undef TEDQ_ENV_CASE
    int rc = check_launch("tedq_tn_contract(env)");
    if (rc != kOk || L.split_log == 0) return rc;
    hipLaunchKernelGGL(k_tn_env_sum, dim3((unsigned)L.n_out, (unsigned)batch), dim3(256), 0, st, part, Cp, c_bstride,
                       L.split_log, L.n_out);
    return check_launch("tedq_tn_contract(env sum)");
  }
  const int64_t threads = L.n_out << (L.lpo_log + L.split_log);
  const int64_t blocks = (threads + 255) / 256;
  if (blocks > 0x7fffffffll) {
    set_error("tedq_tn_contract: grid too large (%lld blocks)", (long long)blocks);
    return kBadArgument;
  }
  hipLaunchKernelGGL(k_tn_small, dim3((unsigned)blocks, (unsigned)batch), dim3(256), 0, st, A, Bp, Cp, part, desc_dev,
                     a_bstride, b_bstride, c_bstride, L.lpo_log, L.split_log, L.n_out);
  int rc = check_launch("tedq_tn_contract(small)");
  if (rc != kOk || L.split_log == 0) return rc;
  const int64_t sb = (L.n_out + 255) / 256;
  hipLaunchKernelGGL(k_tn_split_sum, dim3((unsigned)sb, (unsigned)batch), dim3(256), 0, st, part, Cp, c_bstride,
                     L.split_log, L.n_out);
  return check_launch("tedq_tn_contract(split sum)");
}

// Small steps of one tree level in one launch (see k_tn_small_multi).  bases /
// strides / lens: three operand arenas (device pointers, per-batch strides in
// elements, elements valid per batch row); base 2 is also where every result
// goes.  records / records_dev: n_records records of 8 int64, host and device
// copies; descs / descs_dev: the descriptor table (n_descs rows), host and
// device copies.  Each step must have rank <= 12 operands and results, at most
// 2^6 summed terms, and fit its arenas (all checked on the host).
extern "C" int tedq_tn_contract_small_batch(const void* const* bases, const int64_t* strides, const int64_t* lens,
                                            const int64_t* records, const int64_t* records_dev, int n_records,
                                            int64_t batch, const int32_t* descs, const int32_t* descs_dev,
                                            int n_descs, void* stream) {
  if (bases == nullptr || strides == nullptr || lens == nullptr || records == nullptr || records_dev == nullptr ||
      descs == nullptr || descs_dev == nullptr || n_records < 0 || n_descs <= 0 || batch <= 0 || batch > 65535) {
    set_error("tedq_tn_contract_small_batch: bad arguments");
    return kBadArgument;
  }
  if (n_records == 0) return kOk;
  TNBases tb;
  for (int i = 0; i < 3; ++i) {
    if (bases[i] == nullptr || strides[i] < 0 || lens[i] < 0) {
      set_error("tedq_tn_contract_small_batch: bad arena %d", i);
      return kBadArgument;
    }
    tb.p[i] = static_cast<const float2*>(bases[i]);
    tb.stride[i] = strides[i];
  }
  for (int r = 0; r < n_records; ++r) {
    const int64_t* rec = records + 8 * (int64_t)r;
    TNDesc d;
    if (rec[0] < 0 || rec[0] >= n_descs || parse_desc(descs + rec[0] * (4 + 9 * 32), &d) != 0) {
      set_error("tedq_tn_contract_small_batch: record %d: bad descriptor row %lld", r, (long long)rec[0]);
      return kBadArgument;
    }
    const int nH = d.nb[0], nM = d.nb[1], nN = d.nb[2], nK = d.nb[3];
    const int64_t sa = 1ll << (nH + nM + nK), sb = 1ll << (nH + nN + nK), sc = 1ll << (nH + nM + nN);
    const bool shapes = nH + nM + nK <= 12 && nH + nN + nK <= 12 && nH + nM + nN <= 12 && nK <= 6 && rec[6] == sc;
    const bool in_a = rec[1] >= 0 && rec[1] < 3 && rec[2] >= 0 && rec[2] + sa <= lens[rec[1]];
    const bool in_b = rec[3] >= 0 && rec[3] < 3 && rec[4] >= 0 && rec[4] + sb <= lens[rec[3]];
    const bool in_c = rec[5] >= 0 && rec[5] + sc <= lens[2];
    if (!shapes || !in_a || !in_b || !in_c) {
      set_error("tedq_tn_contract_small_batch: record %d does not fit (shape or arena bounds)", r);
      return kBadArgument;
    }
  }
  hipLaunchKernelGGL(k_tn_small_multi, dim3((unsigned)n_records, (unsigned)batch), dim3(256), 0,
                     static_cast<hipStream_t>(stream), tb, const_cast<float2*>(tb.p[2]), records_dev, descs_dev);
  return check_launch("tedq_tn_contract_small_batch");
}

// Eligibility of a forward step for tedq_tn_gate_backward (the k_tn_gate shape
// rule) and the launch shape: returns the partial-sum workspace bytes, or -1.
static int64_t gate_bwd_plan(const TNDesc& d, int64_t batch, int* a_big, int* nk, int* nq, int* np) {
  const int nH = d.nb[0], nM = d.nb[1], nN = d.nb[2], nK = d.nb[3];
  if (nH != 0 || nK > 2) return -1;
  if (nN <= 2 && nM >= 10) {
    *a_big = 1;
    *nq = nN;
    *np = nM;
  } else if (nM <= 2 && nN >= 10) {
    *a_big = 0;
    *nq = nM;
    *np = nN;
  } else {
    return -1;
  }
  *nk = nK;
  return (int64_t)sizeof(float2) * batch * (1ll << (*np - 10)) * (1ll << (nK + *nq));
}

extern "C" int64_t tedq_tn_gate_backward_workspace(const int32_t* desc, int64_t batch) {
  TNDesc d;
  int a_big, nk, nq, np;
  if (desc == nullptr || batch <= 0 || parse_desc(desc, &d) != 0) return -1;
  return gate_bwd_plan(d, batch, &a_big, &nk, &nq, &np);
}

// Both environments of a gate step (forward descriptor desc: C = A B with one
// operand X large and the other G gate-sized, as tedq_tn_gate_backward_workspace
// accepts): ex = E_X in X's layout, eg = E_G in G's layout, from ec = E_C in
// C's layout; x/g are the step's X and G operands.
extern "C" int tedq_tn_gate_backward(const void* ec, int64_t ec_bs, const void* x, int64_t x_bs, const void* g,
                                     int64_t g_bs, void* ex, int64_t ex_bs, void* eg, int64_t eg_bs, int64_t batch,
                                     const int32_t* desc, const int32_t* desc_dev, void* workspace,
                                     int64_t workspace_bytes, void* stream) {
  TNDesc d;
  int a_big, nk, nq, np;
  if (ec == nullptr || x == nullptr || g == nullptr || ex == nullptr || eg == nullptr || desc == nullptr ||
      desc_dev == nullptr || batch <= 0 || batch > 65535 || parse_desc(desc, &d) != 0) {
    set_error("tedq_tn_gate_backward: bad arguments");
    return kBadArgument;
  }
  const int64_t need = gate_bwd_plan(d, batch, &a_big, &nk, &nq, &np);
  if (need < 0) {
    set_error("tedq_tn_gate_backward: the step is not a gate on a large tensor");
    return kBadArgument;
  }
  if (workspace == nullptr || workspace_bytes < need) {
    set_error("tedq_tn_gate_backward: workspace of %lld bytes required, %lld given", (long long)need,
              (long long)workspace_bytes);
    return kBadArgument;
  }
  hipStream_t st = static_cast<hipStream_t>(stream);
  const float2* EC = static_cast<const float2*>(ec);
  const float2* X = static_cast<const float2*>(x);
  const float2* G = static_cast<const float2*>(g);
  float2* EX = static_cast<float2*>(ex);
  float2* part = static_cast<float2*>(workspace);
  const int split_log = np - 10;
  const dim3 grid(1u << split_log, (unsigned)batch);
#define TEDQ_GBWD_CASE(K, Q)                                                                                   \
  else if (nk == K && nq == Q) hipLaunchKernelGGL((k_tn_gate_bwd<K, Q>), grid, dim3(256), 0, st, EC, X, G, EX, part, \
                                                  desc_dev, ec_bs, x_bs, g_bs, ex_bs, a_big);
  if (false) {
  }
  TEDQ_GBWD_CASE(0, 0) TEDQ_GBWD_CASE(0, 1) TEDQ_GBWD_CASE(0, 2)
  TEDQ_GBWD_CASE(1, 0) TEDQ_GBWD_CASE(1, 1) TEDQ_GBWD_CASE(1, 2)
  TEDQ_GBWD_CASE(2, 0) TEDQ_GBWD_CASE(2, 1) TEDQ_GBWD_CASE(2, 2)
#undef TEDQ_GBWD_CASE
  int rc = check_launch("tedq_tn_gate_backward");
  if (rc != kOk) return rc;
  hipLaunchKernelGGL(k_tn_env_sum, dim3(1u << (nk + nq), (unsigned)batch), dim3(256), 0, st, part,
                     static_cast<float2*>(eg), eg_bs, split_log, (int64_t)(1ll << (nk + nq)));
  return check_launch("tedq_tn_gate_backward(sum)");
}

