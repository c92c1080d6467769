// State initialisation, measurements and gradient reductions.
//
// Replaces the reference measurement loop (tedq/backends/pytorch_backend.py:393-498):
// there every expectation value costs a tensordot, a permute and a full
// inner product (three passes over the state per observable) in complex64.
// Here all diagonal (Z-string) observables of a circuit are reduced in ONE
// pass, accumulated in fp64 per thread and combined across workgroups with
// Neumaier-compensated fp64 summation; the mask 0 observable is the norm.
#include <mutex>

#include "tedq_common.h"

namespace tedq {

constexpr int kMaxZ = 32;          // Z-string observables per launch
constexpr int kChunkLog = 14;      // amplitudes per workgroup in reductions
constexpr int kRedThreads = 256;

__device__ __forceinline__ double wave_sum_d(double v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Neumaier compensated accumulation
__device__ __forceinline__ void neumaier(double& s, double& c, double x) {
  const double t = s + x;
  if (fabs(s) >= fabs(x))
    c += (s - t) + x;
  else
    c += (x - t) + s;
  s = t;
}

// ---------------------------------------------------------------------------
// |0...0> and user initial states
// ---------------------------------------------------------------------------
__global__ void k_init_zero(float2* __restrict__ st, int64_t total, int n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= total) return;
  st[i] = make_float2((i & ((1ll << n) - 1)) == 0 ? 1.f : 0.f, 0.f);
}

__global__ void k_init_copy(float2* __restrict__ st, const float2* __restrict__ init, int64_t total, int n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= total) return;
  st[i] = init[i & ((1ll << n) - 1)];
}

// ---------------------------------------------------------------------------
// <Z-string> expectation values, all masks in one pass
// partial[b][chunk][m] = sum over the chunk of |psi_x|^2 * (-1)^popc(x & mask_m)
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(kRedThreads) k_expval_z_partial(const float2* __restrict__ st, int n,
                                                                   const uint64_t* __restrict__ masks, int M,
                                                                   double* __restrict__ partial, int64_t nchunk) {
  __shared__ double red[kRedThreads / 64][kMaxZ];
  const int64_t b = blockIdx.y;
  const int64_t chunk = blockIdx.x;
  const int64_t csize = 1ll << (n < kChunkLog ? n : kChunkLog);
  const int64_t base = (b << n) + chunk * csize;
  double acc[kMaxZ];
#pragma unroll
  for (int m = 0; m < kMaxZ; ++m) acc[m] = 0.0;
  for (int64_t i = threadIdx.x; i < csize; i += kRedThreads) {
    const float2 a = st[base + i];
    const double p = (double)a.x * a.x + (double)a.y * a.y;
    const uint64_t x = (uint64_t)(chunk * csize + i);
#pragma unroll
    for (int m = 0; m < kMaxZ; ++m)
      if (m < M) acc[m] += (__popcll(x & masks[m]) & 1) ? -p : p;
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int m = 0; m < kMaxZ; ++m) {
    if (m < M) {
      const double v = wave_sum_d(acc[m]);
      if (lane == 0) red[wave][m] = v;
    }
  }
  __syncthreads();
  if (threadIdx.x < M) {
    double s = 0.0;
    for (int w = 0; w < kRedThreads / 64; ++w) s += red[w][threadIdx.x];
    partial[(b * nchunk + chunk) * M + threadIdx.x] = s;
  }
}

// Same partial sums for n >= 14, organised so that the parity work is not per
// amplitude: thread t of chunk c handles the 64 amplitudes x = c*2^14 + s*2^12
// + j*2^8 + t (s < 4, j < 16).  For each s the 16 probabilities p_j are
// Walsh-Hadamard transformed in registers, W[f] = sum_j p_j (-1)^popc(j & f),
// and then mask m contributes (-1)^popc(hi & mask) * W[(mask >> 8) & 15] with hi
// the bits of x outside j.  Per amplitude that is about 4 fp64 adds instead of
// a popcount, select and add for every mask.  Each thread keeps its W in its
// own LDS column, so the per-mask lookup is an indexed LDS read.
constexpr int kWhtChunkLog = 14;
static_assert(kWhtChunkLog == kChunkLog, "the transformed kernels use the reduction chunks");

__global__ void __launch_bounds__(kRedThreads) k_expval_z_wht(const float2* __restrict__ st, int n,
                                                               const uint64_t* __restrict__ masks, int M,
                                                               double* __restrict__ partial, int64_t nchunk) {
  __shared__ double wl[16][kRedThreads];
  __shared__ double red[kRedThreads / 64][kMaxZ];
  const int64_t b = blockIdx.y;
  const int64_t chunk = blockIdx.x;
  const int t = threadIdx.x;
  const int64_t base = (b << n) + (chunk << kWhtChunkLog);
  double acc[kMaxZ];
#pragma unroll
  for (int m = 0; m < kMaxZ; ++m) acc[m] = 0.0;
  // all 64 amplitudes of this thread are requested before any is used, so that a wave keeps 64 loads in
  // flight (issued per quarter, the chunk's reads ran at 2.9 TB/s)
  float2 av[4][16];
#pragma unroll
  for (int s = 0; s < 4; ++s)
#pragma unroll
    for (int j = 0; j < 16; ++j) av[s][j] = st[base + (s << 12) + (j << 8) + t];
#pragma unroll
  for (int s = 0; s < 4; ++s) {
    double w[16];
#pragma unroll
    for (int j = 0; j < 16; ++j) {
      const float2 a = av[s][j];
      w[j] = (double)a.x * a.x + (double)a.y * a.y;
    }
#pragma unroll
    for (int h = 1; h < 16; h <<= 1)
#pragma unroll
      for (int j = 0; j < 16; ++j)
        if (!(j & h)) {
          const double u = w[j], v = w[j | h];
          w[j] = u + v;
          w[j | h] = u - v;
        }
#pragma unroll
    for (int f = 0; f < 16; ++f) wl[f][t] = w[f];
    const uint64_t hi = ((uint64_t)chunk << kWhtChunkLog) | ((uint64_t)s << 12) | (uint64_t)t;
#pragma unroll
    for (int m = 0; m < kMaxZ; ++m) {
      if (m < M) {
        const uint64_t mk = masks[m];
        const double v = wl[(mk >> 8) & 15][t];
        acc[m] += (__popcll(hi & mk) & 1) ? -v : v;
      }
    }
  }
  const int lane = t & 63, wave = t >> 6;
#pragma unroll
  for (int m = 0; m < kMaxZ; ++m) {
    if (m < M) {
      const double v = wave_sum_d(acc[m]);
      if (lane == 0) red[wave][m] = v;
    }
  }
  __syncthreads();
  if (t < M) {
    double sum = 0.0;
    for (int w2 = 0; w2 < kRedThreads / 64; ++w2) sum += red[w2][t];
    partial[(b * nchunk + chunk) * M + t] = sum;
  }
}

// out[b][m] = compensated sum over chunks; one thread per (b, m)
template <typename OT>
__global__ void k_sum_partials(const double* __restrict__ partial, int64_t batch, int64_t nchunk, int M,
                               OT* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= batch * M) return;
  const int64_t b = i / M;
  const int m = (int)(i - b * M);
  double s = 0.0, c = 0.0;
  for (int64_t ch = 0; ch < nchunk; ++ch) neumaier(s, c, partial[(b * nchunk + ch) * M + m]);
  out[i] = (OT)(s + c);
}

// lambda_x = (sum_m dy[b][m] (-1)^popc(x & mask_m)) psi_x      (M may be 0: lambda = 0)
__global__ void k_lambda_z(const float2* __restrict__ psi, float2* __restrict__ lam, int64_t total, int n,
                           const uint64_t* __restrict__ masks, int M, const double* __restrict__ dy) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= total) return;
  const int64_t b = i >> n;
  const uint64_t x = (uint64_t)(i & ((1ll << n) - 1));
  double w = 0.0;
  for (int m = 0; m < M; ++m) {
    const double d = dy[b * M + m];
    w += (__popcll(x & masks[m]) & 1) ? -d : d;
  }
  const float2 a = psi[i];
  lam[i] = make_float2((float)(w * a.x), (float)(w * a.y));
}

// lambda for n >= 14 with the same layout as k_expval_z_wht: per thread and s,
// E[f] = sum over masks with (mask >> 8) & 15 == f of dy_m (-1)^popc(hi & mask)
// (built in the thread's LDS column), then a Walsh-Hadamard transform gives
// w(x) for the 16 amplitudes x = hi | j*2^8 at once.
__global__ void __launch_bounds__(kRedThreads) k_lambda_z_wht(const float2* __restrict__ psi,
                                                               float2* __restrict__ lam, int n,
                                                               const uint64_t* __restrict__ masks, int M,
                                                               const double* __restrict__ dy) {
  __shared__ double el[16][kRedThreads];
  const int64_t b = blockIdx.y;
  const int64_t chunk = blockIdx.x;
  const int t = threadIdx.x;
  const int64_t base = (b << n) + (chunk << kWhtChunkLog);
  for (int s = 0; s < 4; ++s) {
    const uint64_t hi = ((uint64_t)chunk << kWhtChunkLog) | ((uint64_t)s << 12) | (uint64_t)t;
#pragma unroll
    for (int f = 0; f < 16; ++f) el[f][t] = 0.0;
    for (int m = 0; m < M; ++m) {
      const uint64_t mk = masks[m];
      const double d = dy[b * M + m];
      el[(mk >> 8) & 15][t] += (__popcll(hi & mk) & 1) ? -d : d;
    }
    double e[16];
#pragma unroll
    for (int f = 0; f < 16; ++f) e[f] = el[f][t];
#pragma unroll
    for (int h = 1; h < 16; h <<= 1)
#pragma unroll
      for (int j = 0; j < 16; ++j)
        if (!(j & h)) {
          const double u = e[j], v = e[j | h];
          e[j] = u + v;
          e[j | h] = u - v;
        }
#pragma unroll
    for (int j = 0; j < 16; ++j) {
      const int64_t x = base + (s << 12) + (j << 8) + t;
      const float2 a = psi[x];
      lam[x] = make_float2((float)(e[j] * a.x), (float)(e[j] * a.y));
    }
  }
}

// ---------------------------------------------------------------------------
// Factored observable O = zsign(zmask) * (F_0 (x) F_1 (x) ... ) on q <= 4 bits.
// Each thread owns one group of 2^q amplitudes that differ only in those bits.
// ---------------------------------------------------------------------------
struct FactorObs {
  int q;
  int bits[4];        // physical positions
  float2 f[4][4];     // 2x2 factor per bit, row-major
  uint64_t zmask;
};

__device__ __forceinline__ void apply_factors(float2 (&g)[16], const FactorObs& o, bool dagger) {
  for (int t = 0; t < o.q; ++t) {
    float2 m0 = o.f[t][0], m1 = o.f[t][1], m2 = o.f[t][2], m3 = o.f[t][3];
    if (dagger) {
      const float2 tmp = m1;
      m0 = cconj(m0);
      m1 = cconj(m2);
      m2 = cconj(tmp);
      m3 = cconj(m3);
    }
    const int sz = 1 << o.q;
    for (int i = 0; i < sz; ++i) {
      if (i & (1 << t)) continue;
      const int j = i | (1 << t);
      const float2 x = g[i], y = g[j];
      g[i] = cmul2(m0, x, m1, y);
      g[j] = cmul2(m2, x, m3, y);
    }
  }
}

__device__ __forceinline__ uint64_t group_index(uint64_t gid, const FactorObs& o, int t) {
  // insert the bits of t (bit i -> physical o.bits[i]) into gid with zeros at o.bits
  uint64_t x = gid;
  // bits must be inserted in ascending physical order
  int order[4];
  for (int i = 0; i < o.q; ++i) order[i] = i;
  for (int i = 0; i < o.q; ++i)
    for (int j = i + 1; j < o.q; ++j)
      if (o.bits[order[j]] < o.bits[order[i]]) {
        const int tmp = order[i];
        order[i] = order[j];
        order[j] = tmp;
      }
  for (int ii = 0; ii < o.q; ++ii) {
    const int p = o.bits[order[ii]];
    const uint64_t low = x & ((1ull << p) - 1);
    x = ((x >> p) << (p + 1)) | low;
  }
  for (int i = 0; i < o.q; ++i)
    if ((t >> i) & 1) x |= 1ull << o.bits[i];
  return x;
}

__global__ void __launch_bounds__(kRedThreads) k_expval_factored(const float2* __restrict__ st, int n, FactorObs o,
                                                                  double* __restrict__ partial, int64_t ngroups,
                                                                  int64_t nchunk) {
  __shared__ double red[kRedThreads / 64];
  const int64_t b = blockIdx.y;
  const float2* s = st + (b << n);
  const int64_t per = (ngroups + nchunk - 1) / nchunk;
  const int64_t g0 = blockIdx.x * per, g1 = g0 + per < ngroups ? g0 + per : ngroups;
  double acc = 0.0;
  const int sz = 1 << o.q;
  for (int64_t gid = g0 + threadIdx.x; gid < g1; gid += kRedThreads) {
    float2 g[16], h[16];
    uint64_t idx[16];
    for (int t = 0; t < sz; ++t) {
      idx[t] = group_index((uint64_t)gid, o, t);
      g[t] = s[idx[t]];
      h[t] = g[t];
    }
    apply_factors(h, o, false);
    const double zs = (__popcll(idx[0] & o.zmask) & 1) ? -1.0 : 1.0;
    double part = 0.0;
    for (int t = 0; t < sz; ++t) part += (double)g[t].x * h[t].x + (double)g[t].y * h[t].y;
    acc += zs * part;
  }
  acc = wave_sum_d(acc);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int w = 0; w < kRedThreads / 64; ++w) t += red[w];
    partial[b * nchunk + blockIdx.x] = t;
  }
}

// lambda += coef[b] * zsign * ((F + F^dagger)/2) psi on every group
__global__ void k_lambda_factored(const float2* __restrict__ psi, float2* __restrict__ lam, int n, FactorObs o,
                                  int64_t ngroups, const double* __restrict__ coef, int64_t batch) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= batch * ngroups) return;
  const int64_t b = i / ngroups, gid = i - b * ngroups;
  const float2* s = psi + (b << n);
  float2* l = lam + (b << n);
  const int sz = 1 << o.q;
  float2 g[16], h1[16], h2[16];
  uint64_t idx[16];
  for (int t = 0; t < sz; ++t) {
    idx[t] = group_index((uint64_t)gid, o, t);
    g[t] = s[idx[t]];
    h1[t] = g[t];
    h2[t] = g[t];
  }
  apply_factors(h1, o, false);
  apply_factors(h2, o, true);
  const double zs = (__popcll(idx[0] & o.zmask) & 1) ? -1.0 : 1.0;
  const float c = (float)(0.5 * zs * coef[b]);
  for (int t = 0; t < sz; ++t) {
    float2 cur = l[idx[t]];
    cur.x += c * (h1[t].x + h2[t].x);
    cur.y += c * (h1[t].y + h2[t].y);
    l[idx[t]] = cur;
  }
}

// ---------------------------------------------------------------------------
// probabilities
// ---------------------------------------------------------------------------
__global__ void k_probs_full(const float2* __restrict__ st, float* __restrict__ out, int64_t total) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= total) return;
  const float2 a = st[i];
  out[i] = a.x * a.x + a.y * a.y;
}

__device__ __forceinline__ uint32_t extract_bits(uint64_t x, const int* bits, int k) {
  uint32_t y = 0;
  for (int i = 0; i < k; ++i) y = (y << 1) | (uint32_t)((x >> bits[i]) & 1);
  return y;
}

struct BitList {
  int k;
  int bits[32];  // physical positions, most significant output bit first
};

// Physical offset of outcome y of the listed bits (bit k-1-i of y goes to bits[i]).
__device__ __forceinline__ uint64_t deposit_outcome(uint32_t y, const BitList& bl) {
  uint64_t x = 0;
  for (int i = 0; i < bl.k; ++i) x |= (uint64_t)((y >> (bl.k - 1 - i)) & 1u) << bl.bits[i];
  return x;
}

// Marginal probabilities without atomics, so the summation order is fixed.
// Block (bin * nrc + rc, b) sums |psi_x|^2 in fp64 over 2^clog values r of the
// remaining bits (clog = min(n - k, 14), r in rest chunk rc) of outcome bin, by
// a fixed per-thread stride and a fixed tree, into partial[b][rc][bin].
__global__ void __launch_bounds__(kRedThreads) k_probs_marg(const float2* __restrict__ st, int n, BitList bl,
                                                             uint64_t rest_mask, double* __restrict__ partial,
                                                             int64_t nrc) {
  __shared__ double red[kRedThreads / 64];
  const int64_t b = blockIdx.y;
  const int64_t bin = blockIdx.x / nrc, rc = blockIdx.x - bin * nrc;
  const int rest = n - bl.k;
  const int clog = rest < kChunkLog ? rest : kChunkLog;
  const uint64_t fixed = deposit_outcome((uint32_t)bin, bl);
  const float2* base = st + (b << n);
  double acc = 0.0;
  for (int64_t i = threadIdx.x; i < (1ll << clog); i += kRedThreads) {
    const float2 a = base[fixed | pdep64((uint64_t)((rc << clog) + i), rest_mask)];
    acc += (double)a.x * a.x + (double)a.y * a.y;
  }
  acc = wave_sum_d(acc);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0.0;
    for (int w = 0; w < kRedThreads / 64; ++w) s += red[w];
    partial[(b * nrc + rc) * (1ll << bl.k) + bin] = s;
  }
}

// Post-measurement states of one circuit: out[y][r] = psi[x(y, r)] / sqrt(p[y]) for outcome y of
// the listed bits and r over the remaining bits (C order of the remaining qubits, ascending).
__global__ void k_post_measure(const float2* __restrict__ psi, int n, BitList bl, uint64_t rest_mask,
                               const float* __restrict__ p, float2* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (1ll << n)) return;
  const int rest = n - bl.k;
  const uint32_t y = (uint32_t)(i >> rest);
  const uint64_t r = (uint64_t)(i & ((1ll << rest) - 1));
  const float2 a = psi[deposit_outcome(y, bl) | pdep64(r, rest_mask)];
  const float scale = sqrtf(p[y]);
  out[i] = make_float2(a.x / scale, a.y / scale);
}

// lambda_x += g[b][marg(x)] psi_x   (k == n means the full distribution)
__global__ void k_lambda_probs(const float2* __restrict__ psi, float2* __restrict__ lam, int64_t total, int n,
                               BitList bl, const float* __restrict__ g) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= total) return;
  const int64_t b = i >> n;
  const uint64_t x = (uint64_t)(i & ((1ll << n) - 1));
  const uint32_t y = (bl.k == n) ? (uint32_t)x : extract_bits(x, bl.bits, bl.k);
  const float w = g[(b << bl.k) + y];
  const float2 a = psi[i];
  float2 cur = lam[i];
  cur.x = fmaf(w, a.x, cur.x);
  cur.y = fmaf(w, a.y, cur.y);
  lam[i] = cur;
}

// lambda += scale * G   (gradient of a state output)
__global__ void k_axpy_c(float2* __restrict__ lam, const float2* __restrict__ G, int64_t total, float scale) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= total) return;
  float2 cur = lam[i];
  const float2 gg = G[i];
  cur.x = fmaf(scale, gg.x, cur.x);
  cur.y = fmaf(scale, gg.y, cur.y);
  lam[i] = cur;
}

// ---------------------------------------------------------------------------
// gradient: C[b][g][8] = compensated sum of adjoint partial rows, then
// grad[b][p] = 2 Re sum_ij A[b][p][ij] C[b][gslot(p)][ij]
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(kRedThreads) k_reduce_bilinears(const float* __restrict__ partials, int64_t rows,
                                                                   double* __restrict__ C) {
  __shared__ double red[kRedThreads / 64][8];
  const int64_t bg = blockIdx.x;  // b * n_gslots + g
  const float* src = partials + bg * rows * 8;
  double s[8], c[8];
  for (int q = 0; q < 8; ++q) s[q] = c[q] = 0.0;
  for (int64_t r = threadIdx.x; r < rows; r += kRedThreads) {
    const float4 lo = reinterpret_cast<const float4*>(src + r * 8)[0];
    const float4 hi = reinterpret_cast<const float4*>(src + r * 8)[1];
    neumaier(s[0], c[0], lo.x);
    neumaier(s[1], c[1], lo.y);
    neumaier(s[2], c[2], lo.z);
    neumaier(s[3], c[3], lo.w);
    neumaier(s[4], c[4], hi.x);
    neumaier(s[5], c[5], hi.y);
    neumaier(s[6], c[6], hi.z);
    neumaier(s[7], c[7], hi.w);
  }
  for (int q = 0; q < 8; ++q) {
    const double v = wave_sum_d(s[q] + c[q]);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6][q] = v;
  }
  __syncthreads();
  if (threadIdx.x < 8) {
    double t = 0.0;
    for (int w = 0; w < kRedThreads / 64; ++w) t += red[w][threadIdx.x];
    C[bg * 8 + threadIdx.x] = t;
  }
}

template <typename OT>
__global__ void k_grad_combine(const double* __restrict__ C, const double2* __restrict__ amats,
                               const int* __restrict__ pgslot, int64_t batch, int P, int G, OT* __restrict__ grad) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= batch * P) return;
  const int64_t b = i / P;
  const int p = (int)(i - b * P);
  const int g = pgslot[p];
  if (g < 0) {
    grad[i] = (OT)0;
    return;
  }
  const double* c = C + (b * G + g) * 8;
  const double2* A = amats + i * 4;
  double re = 0.0;
  for (int q = 0; q < 4; ++q) re += A[q].x * c[2 * q] - A[q].y * c[2 * q + 1];
  grad[i] = (OT)(2.0 * re);
}

static inline unsigned nblocks(int64_t total, int threads) { return (unsigned)((total + threads - 1) / threads); }

}  // namespace tedq

using namespace tedq;

extern "C" int tedq_sv_init_zero(void* state, int64_t batch, int n_qubits, void* stream) {
  if (state == nullptr || batch <= 0 || n_qubits < 1 || n_qubits > 40) {
    set_error("tedq_sv_init_zero: bad arguments");
    return kBadArgument;
  }
  const int64_t total = batch << n_qubits;
  hipLaunchKernelGGL(k_init_zero, dim3(nblocks(total, 256)), dim3(256), 0, static_cast<hipStream_t>(stream),
                     static_cast<float2*>(state), total, n_qubits);
  return check_launch("tedq_sv_init_zero");
}

extern "C" int tedq_sv_init_copy(void* state, const void* init, int64_t batch, int n_qubits, void* stream) {
  if (state == nullptr || init == nullptr || batch <= 0 || n_qubits < 1 || n_qubits > 40) {
    set_error("tedq_sv_init_copy: bad arguments");
    return kBadArgument;
  }
  const int64_t total = batch << n_qubits;
  hipLaunchKernelGGL(k_init_copy, dim3(nblocks(total, 256)), dim3(256), 0, static_cast<hipStream_t>(stream),
                     static_cast<float2*>(state), static_cast<const float2*>(init), total, n_qubits);
  return check_launch("tedq_sv_init_copy");
}

extern "C" int64_t tedq_sv_reduce_chunks(int n_qubits) {
  return n_qubits <= kChunkLog ? 1 : (1ll << (n_qubits - kChunkLog));
}

static int expval_z_args_ok(const void* state, int64_t batch, int n_qubits, const uint64_t* dev_masks, int n_masks,
                            const double* out, const char* what) {
  if (state == nullptr || dev_masks == nullptr || out == nullptr || batch <= 0 || n_qubits < 1 || n_qubits > 40 ||
      n_masks < 1 || n_masks > kMaxZ || batch > 65535) {
    set_error("%s: bad arguments (batch=%lld n=%d masks=%d)", what, (long long)batch, n_qubits, n_masks);
    return 0;
  }
  return 1;
}

static int launch_expval_z(const void* state, int64_t batch, int n_qubits, const uint64_t* dev_masks, int n_masks,
                           double* partial, double* out, hipStream_t st) {
  const int64_t nchunk = tedq_sv_reduce_chunks(n_qubits);
  if (n_qubits >= kWhtChunkLog)
    hipLaunchKernelGGL(k_expval_z_wht, dim3((unsigned)nchunk, (unsigned)batch), dim3(kRedThreads), 0, st,
                       static_cast<const float2*>(state), n_qubits, dev_masks, n_masks, partial, nchunk);
  else
    hipLaunchKernelGGL(k_expval_z_partial, dim3((unsigned)nchunk, (unsigned)batch), dim3(kRedThreads), 0, st,
                       static_cast<const float2*>(state), n_qubits, dev_masks, n_masks, partial, nchunk);
  int rc = check_launch("tedq_sv_expval_z(partial)");
  if (rc) return rc;
  hipLaunchKernelGGL(k_sum_partials<double>, dim3(nblocks(batch * n_masks, 128)), dim3(128), 0, st, partial, batch, nchunk,
                     n_masks, out);
  return check_launch("tedq_sv_expval_z(sum)");
}

extern "C" int tedq_sv_expval_z(const void* state, int64_t batch, int n_qubits, const uint64_t* dev_masks,
                                int n_masks, double* partial, double* out, void* stream) {
  if (!expval_z_args_ok(state, batch, n_qubits, dev_masks, n_masks, out, "tedq_sv_expval_z")) return kBadArgument;
  if (partial == nullptr) {
    set_error("tedq_sv_expval_z: bad arguments (null partial buffer)");
    return kBadArgument;
  }
  return launch_expval_z(state, batch, n_qubits, dev_masks, n_masks, partial, out, static_cast<hipStream_t>(stream));
}

// ---------------------------------------------------------------------------
// tedq_sv_expval_z without a caller buffer, for callers that cannot allocate.
// The partial sums live in `scratch`, a file-scope device array: every GPU the
// library runs on holds one copy of it until the process ends.  Host code never
// passes the array's name to a kernel (on the host it is only a stand-in for
// the symbol): the device address of each GPU's copy is looked up once per
// device with hipGetSymbolAddress, with that device current, and kept in
// g_scratch_dev.  The partial kernels write every entry the sum reads, so no
// clearing is needed; calls that overlap on one GPU share the array and race.
// ---------------------------------------------------------------------------
constexpr int64_t kNoAllocScratchBytes = 8ll << 20;
constexpr int kMaxScratchDevices = 64;
static __device__ __align__(256) double scratch[kNoAllocScratchBytes / sizeof(double)];
static double* g_scratch_dev[kMaxScratchDevices];
static std::mutex g_scratch_mu;

static int scratch_for_current_device(double** out) {
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess || dev < 0 || dev >= kMaxScratchDevices) {
    set_error("tedq_sv_expval_z_noalloc: no usable current device (%s)", hipGetErrorString(e));
    return kLaunchFailed;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_dev[dev] == nullptr) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || p == nullptr) {
      set_error("tedq_sv_expval_z_noalloc: hipGetSymbolAddress(scratch) failed on device %d: %s", dev,
                hipGetErrorString(e));
      return kLaunchFailed;
    }
    g_scratch_dev[dev] = static_cast<double*>(p);
  }
  *out = g_scratch_dev[dev];
  return kOk;
}

extern "C" int tedq_sv_expval_z_noalloc(const void* state, int64_t batch, int n_qubits, const uint64_t* dev_masks,
                                        int n_masks, double* out, void* stream) {
  if (!expval_z_args_ok(state, batch, n_qubits, dev_masks, n_masks, out, "tedq_sv_expval_z_noalloc"))
    return kBadArgument;
  const int64_t need = batch * tedq_sv_reduce_chunks(n_qubits) * n_masks * (int64_t)sizeof(double);
  if (need > kNoAllocScratchBytes) {
    set_error("tedq_sv_expval_z_noalloc: needs %lld bytes of partial sums (batch=%lld n=%d masks=%d), "
              "the device scratch holds %lld; use tedq_sv_expval_z with a caller buffer",
              (long long)need, (long long)batch, n_qubits, n_masks, (long long)kNoAllocScratchBytes);
    return kBadArgument;
  }
  double* partial = nullptr;
  int rc = scratch_for_current_device(&partial);
  if (rc) return rc;
  return launch_expval_z(state, batch, n_qubits, dev_masks, n_masks, partial, out, static_cast<hipStream_t>(stream));
}

extern "C" int tedq_sv_lambda_z(const void* psi, void* lam, int64_t batch, int n_qubits, const uint64_t* dev_masks,
                                int n_masks, const double* dy, void* stream) {
  if (psi == nullptr || lam == nullptr || batch <= 0 || n_qubits < 1 || n_qubits > 40 || n_masks < 0 ||
      (n_masks > 0 && (dev_masks == nullptr || dy == nullptr))) {
    set_error("tedq_sv_lambda_z: bad arguments");
    return kBadArgument;
  }
  const int64_t total = batch << n_qubits;
  if (n_qubits >= kWhtChunkLog && batch <= 65535)
    hipLaunchKernelGGL(k_lambda_z_wht, dim3((unsigned)tedq_sv_reduce_chunks(n_qubits), (unsigned)batch),
                       dim3(kRedThreads), 0, static_cast<hipStream_t>(stream), static_cast<const float2*>(psi),
                       static_cast<float2*>(lam), n_qubits, dev_masks, n_masks, dy);
  else
    hipLaunchKernelGGL(k_lambda_z, dim3(nblocks(total, 256)), dim3(256), 0, static_cast<hipStream_t>(stream),
                       static_cast<const float2*>(psi), static_cast<float2*>(lam), total, n_qubits, dev_masks,
                       n_masks, dy);
  return check_launch("tedq_sv_lambda_z");
}

static int make_factor_obs(FactorObs* o, int n, int q, const int* bits, const float* factors, uint64_t zmask) {
  if (q < 0 || q > 4) return kUnsupported;
  o->q = q;
  for (int i = 0; i < 4; ++i) o->bits[i] = 0;
  for (int i = 0; i < q; ++i) {
    if (bits[i] < 0 || bits[i] >= n) return kBadArgument;
    o->bits[i] = bits[i];
    for (int j = 0; j < 4; ++j) o->f[i][j] = make_float2(factors[(i * 4 + j) * 2], factors[(i * 4 + j) * 2 + 1]);
  }
  o->zmask = zmask;
  return kOk;
}

extern "C" int tedq_sv_expval_factored(const void* state, int64_t batch, int n_qubits, int q, const int* bits,
                                       const float* factors, uint64_t zmask, double* partial, double* out,
                                       void* stream) {
  FactorObs o;
  if (state == nullptr || partial == nullptr || out == nullptr || batch <= 0 || batch > 65535 || n_qubits < 1 ||
      n_qubits > 40 || q > n_qubits || make_factor_obs(&o, n_qubits, q, bits, factors, zmask) != kOk) {
    set_error("tedq_sv_expval_factored: bad arguments (q=%d)", q);
    return kBadArgument;
  }
  const int64_t ngroups = 1ll << (n_qubits - q);
  const int64_t nchunk = tedq_sv_reduce_chunks(n_qubits);
  hipStream_t st = static_cast<hipStream_t>(stream);
  hipLaunchKernelGGL(k_expval_factored, dim3((unsigned)nchunk, (unsigned)batch), dim3(kRedThreads), 0, st,
                     static_cast<const float2*>(state), n_qubits, o, partial, ngroups, nchunk);
  int rc = check_launch("tedq_sv_expval_factored(partial)");
  if (rc) return rc;
  hipLaunchKernelGGL(k_sum_partials<double>, dim3(nblocks(batch, 128)), dim3(128), 0, st, partial, batch, nchunk, 1, out);
  return check_launch("tedq_sv_expval_factored(sum)");
}

extern "C" int tedq_sv_lambda_factored(const void* psi, void* lam, int64_t batch, int n_qubits, int q,
                                       const int* bits, const float* factors, uint64_t zmask, const double* coef,
                                       void* stream) {
  FactorObs o;
  if (psi == nullptr || lam == nullptr || coef == nullptr || batch <= 0 || n_qubits < 1 || n_qubits > 40 ||
      q > n_qubits || make_factor_obs(&o, n_qubits, q, bits, factors, zmask) != kOk) {
    set_error("tedq_sv_lambda_factored: bad arguments (q=%d)", q);
    return kBadArgument;
  }
  const int64_t ngroups = 1ll << (n_qubits - q);
  hipLaunchKernelGGL(k_lambda_factored, dim3(nblocks(batch * ngroups, 256)), dim3(256), 0,
                     static_cast<hipStream_t>(stream), static_cast<const float2*>(psi), static_cast<float2*>(lam),
                     n_qubits, o, ngroups, coef, batch);
  return check_launch("tedq_sv_lambda_factored");
}

static int make_bitlist(BitList* bl, int n, int k, const int* bits, const char* what) {
  bl->k = k;
  uint64_t seen = 0;
  for (int i = 0; i < 32; ++i) bl->bits[i] = i < k ? bits[i] : 0;
  for (int i = 0; i < k; ++i) {
    if (bl->bits[i] < 0 || bl->bits[i] >= n || ((seen >> bl->bits[i]) & 1)) {
      set_error("%s: bit %d out of range or repeated", what, bl->bits[i]);
      return kBadArgument;
    }
    seen |= 1ull << bl->bits[i];
  }
  return kOk;
}

static uint64_t rest_mask_of(const BitList& bl, int n) {
  uint64_t m = (n >= 64) ? ~0ull : ((1ull << n) - 1);
  for (int i = 0; i < bl.k; ++i) m &= ~(1ull << bl.bits[i]);
  return m;
}

extern "C" int64_t tedq_sv_probs_chunks(int n_qubits, int k) {
  if (n_qubits < 1 || n_qubits > 40 || k < 1 || k > n_qubits) return -1;
  return n_qubits - k > kChunkLog ? (1ll << (n_qubits - k - kChunkLog)) : 1;
}

extern "C" int tedq_sv_probs(const void* state, int64_t batch, int n_qubits, int k, const int* bits,
                             double* partial, float* out, void* stream) {
  if (state == nullptr || out == nullptr || batch <= 0 || batch > 65535 || n_qubits < 1 || n_qubits > 40 || k < 1 ||
      k > n_qubits || k > 30) {
    set_error("tedq_sv_probs: bad arguments (k=%d n=%d)", k, n_qubits);
    return kBadArgument;
  }
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (k == n_qubits && bits == nullptr) {
    const int64_t total = batch << n_qubits;
    hipLaunchKernelGGL(k_probs_full, dim3(nblocks(total, 256)), dim3(256), 0, st, static_cast<const float2*>(state),
                       out, total);
    return check_launch("tedq_sv_probs(full)");
  }
  if (bits == nullptr || partial == nullptr) {
    set_error("tedq_sv_probs: marginals need the bit list and a partial buffer");
    return kBadArgument;
  }
  BitList bl;
  int rc = make_bitlist(&bl, n_qubits, k, bits, "tedq_sv_probs");
  if (rc) return rc;
  const int64_t nrc = tedq_sv_probs_chunks(n_qubits, k);
  if ((nrc << k) > 0x7fffffffll) {
    set_error("tedq_sv_probs: grid too large");
    return kBadArgument;
  }
  hipLaunchKernelGGL(k_probs_marg, dim3((unsigned)(nrc << k), (unsigned)batch), dim3(kRedThreads), 0, st,
                     static_cast<const float2*>(state), n_qubits, bl, rest_mask_of(bl, n_qubits), partial, nrc);
  rc = check_launch("tedq_sv_probs(partial)");
  if (rc) return rc;
  hipLaunchKernelGGL(k_sum_partials<float>, dim3(nblocks(batch << k, 128)), dim3(128), 0, st, partial, batch, nrc,
                     1 << k, out);
  return check_launch("tedq_sv_probs(sum)");
}

extern "C" int tedq_sv_post_measure(const void* psi, int n_qubits, int k, const int* bits, const float* probs,
                                    void* out, void* stream) {
  if (psi == nullptr || bits == nullptr || probs == nullptr || out == nullptr || n_qubits < 1 || n_qubits > 40 ||
      k < 1 || k > n_qubits || k > 30) {
    set_error("tedq_sv_post_measure: bad arguments (k=%d n=%d)", k, n_qubits);
    return kBadArgument;
  }
  BitList bl;
  const int rc = make_bitlist(&bl, n_qubits, k, bits, "tedq_sv_post_measure");
  if (rc) return rc;
  const int64_t total = 1ll << n_qubits;
  hipLaunchKernelGGL(k_post_measure, dim3(nblocks(total, 256)), dim3(256), 0, static_cast<hipStream_t>(stream),
                     static_cast<const float2*>(psi), n_qubits, bl, rest_mask_of(bl, n_qubits), probs,
                     static_cast<float2*>(out));
  return check_launch("tedq_sv_post_measure");
}

extern "C" int tedq_sv_lambda_probs(const void* psi, void* lam, int64_t batch, int n_qubits, int k, const int* bits,
                                    const float* g, void* stream) {
  if (psi == nullptr || lam == nullptr || g == nullptr || batch <= 0 || n_qubits < 1 || n_qubits > 40 || k < 1 ||
      k > n_qubits || k > 30) {
    set_error("tedq_sv_lambda_probs: bad arguments");
    return kBadArgument;
  }
  BitList bl;
  bl.k = k;
  for (int i = 0; i < 32; ++i) bl.bits[i] = (bits != nullptr && i < k) ? bits[i] : 0;
  const int64_t total = batch << n_qubits;
  hipLaunchKernelGGL(k_lambda_probs, dim3(nblocks(total, 256)), dim3(256), 0, static_cast<hipStream_t>(stream),
                     static_cast<const float2*>(psi), static_cast<float2*>(lam), total, n_qubits, bl, g);
  return check_launch("tedq_sv_lambda_probs");
}

extern "C" int tedq_sv_axpy(void* lam, const void* G, int64_t batch, int n_qubits, float scale, void* stream) {
  if (lam == nullptr || G == nullptr || batch <= 0 || n_qubits < 1 || n_qubits > 40) {
    set_error("tedq_sv_axpy: bad arguments");
    return kBadArgument;
  }
  const int64_t total = batch << n_qubits;
  hipLaunchKernelGGL(k_axpy_c, dim3(nblocks(total, 256)), dim3(256), 0, static_cast<hipStream_t>(stream),
                     static_cast<float2*>(lam), static_cast<const float2*>(G), total, scale);
  return check_launch("tedq_sv_axpy");
}

extern "C" int tedq_sum_partials(const double* partial, int64_t batch, int64_t nchunk, int M, double* out,
                                 void* stream) {
  if (partial == nullptr || out == nullptr || batch <= 0 || nchunk <= 0 || M <= 0) {
    set_error("tedq_sum_partials: bad arguments");
    return kBadArgument;
  }
  hipLaunchKernelGGL(k_sum_partials<double>, dim3(nblocks(batch * M, 128)), dim3(128), 0, static_cast<hipStream_t>(stream),
                     partial, batch, nchunk, M, out);
  return check_launch("tedq_sum_partials");
}

extern "C" int tedq_sv_grad_reduce(const float* partials, int64_t batch, int n_gslots, int64_t rows,
                                   const void* amats, const int* pgslot, int n_params, double* C, void* grad,
                                   int grad_is_f64, void* stream) {
  if (batch <= 0 || n_params < 0 || n_gslots < 0 || (n_gslots > 0 && (partials == nullptr || C == nullptr)) ||
      (n_params > 0 && (grad == nullptr || pgslot == nullptr || amats == nullptr)) || rows <= 0) {
    set_error("tedq_sv_grad_reduce: bad arguments");
    return kBadArgument;
  }
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (n_gslots > 0) {
    hipLaunchKernelGGL(k_reduce_bilinears, dim3((unsigned)(batch * n_gslots)), dim3(kRedThreads), 0, st, partials,
                       rows, C);
    int rc = check_launch("tedq_sv_grad_reduce(bilinears)");
    if (rc) return rc;
  }
  if (n_params == 0) return kOk;
  if (grad_is_f64)
    hipLaunchKernelGGL(k_grad_combine<double>, dim3(nblocks(batch * n_params, 128)), dim3(128), 0, st, C,
                       static_cast<const double2*>(amats), pgslot, batch, n_params, n_gslots,
                       static_cast<double*>(grad));
  else
    hipLaunchKernelGGL(k_grad_combine<float>, dim3(nblocks(batch * n_params, 128)), dim3(128), 0, st, C,
                       static_cast<const double2*>(amats), pgslot, batch, n_params, n_gslots,
                       static_cast<float*>(grad));
  return check_launch("tedq_sv_grad_reduce(combine)");
}
