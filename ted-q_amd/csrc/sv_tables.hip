// Gate-table builder: per (circuit, slot) fused gate matrix and per-parameter
// derivative matrices, computed in fp64 on the device from the parameter rows.
//
// Replaces the per-gate tensor builders of the reference PyTorch backend
// (tedq/backends/pytorch_backend.py:579-1188, e.g. RY at :896-922), which
// build every gate with several tiny torch ops per call.  Matrices follow the
// reference definitions (tedq/QInterpreter/operators/qubit.py: RX :990-997,
// RY :1048-1055, RZ :1106-1112, Rot :1169-1189, PhaseShift :1253-1257).
//
// A slot is the product S = U_m ... U_1 of consecutive single-qubit gates on
// one qubit (or a single controlled block / 4x4 user matrix).  For a
// trainable parameter of component i the table also holds
//     A = S  S_i^dagger  dU_i  S_{i-1}  S^dagger ,   S_i = U_i ... U_1,
// so that with psi = state after the slot and lambda = adjoint state after the
// slot, dL/dtheta = 2 Re <lambda| A |psi>  =  2 Re sum_ij A_ij C_ij  where
// C_ij = sum over amplitude pairs of conj(lambda_i) psi_j.
#include "tedq_common.h"

namespace tedq {

enum GateCode : int {
  kI = 0, kH = 1, kX = 2, kY = 3, kZ = 4, kS = 5, kT = 6, kSX = 7,
  kRX = 14, kRY = 15, kRZ = 16, kRot = 17, kPhaseShift = 18, kUnitary = 23,
};

// scalar: 0 for a slot whose gradients use the full bilinears C_ij of its gslot; c + 1 when the slot's
// one trainable parameter gets its gradient value directly in component c of gslot (kOpPauliGrad): then
// A is the constant that picks that component out of 2 Re sum A_ij C_ij.
struct SlotRec {
  // phase = 1: a diagonal slot diag(d0, d1) also stores p = d1 conj(d0) in its (zero) entry 1, the phase
  // that kOpPhase1 applies to the |1> half (the gate up to its global phase d0)
  int kind, ncomp, comp_start, mat_off, gslot, scalar, phase, pad2;
};
struct CompRec {
  int code, dagger, src0, src1, src2, const_off, pad0, pad1;
};

struct M2 {
  cd a[4];  // row-major 2x2
};

__device__ __forceinline__ M2 m2_mul(const M2& p, const M2& q) {
  M2 r;
  r.a[0] = cd_add(cd_mul(p.a[0], q.a[0]), cd_mul(p.a[1], q.a[2]));
  r.a[1] = cd_add(cd_mul(p.a[0], q.a[1]), cd_mul(p.a[1], q.a[3]));
  r.a[2] = cd_add(cd_mul(p.a[2], q.a[0]), cd_mul(p.a[3], q.a[2]));
  r.a[3] = cd_add(cd_mul(p.a[2], q.a[1]), cd_mul(p.a[3], q.a[3]));
  return r;
}
__device__ __forceinline__ M2 m2_dag(const M2& p) {
  M2 r;
  r.a[0] = cd_conj(p.a[0]);
  r.a[1] = cd_conj(p.a[2]);
  r.a[2] = cd_conj(p.a[1]);
  r.a[3] = cd_conj(p.a[3]);
  return r;
}
__device__ __forceinline__ M2 m2_eye() {
  M2 r;
  r.a[0] = mkcd(1, 0);
  r.a[1] = mkcd(0, 0);
  r.a[2] = mkcd(0, 0);
  r.a[3] = mkcd(1, 0);
  return r;
}
__device__ __forceinline__ M2 m2_make(cd a, cd b, cd c, cd d) {
  M2 r;
  r.a[0] = a;
  r.a[1] = b;
  r.a[2] = c;
  r.a[3] = d;
  return r;
}
__device__ __forceinline__ cd expi(double t) {
  double s, c;
  sincos(t, &s, &c);
  return mkcd(c, s);
}

// Matrix of one component and, when which >= 0, its derivative with respect
// to parameter `which`.  p[] holds the three parameter values.
__device__ M2 component_matrix(int code, const double* p, const double2* cpool, int const_off, int which,
                               bool* ok) {
  const double h = 0.70710678118654752440;
  const cd z = mkcd(0, 0), one = mkcd(1, 0);
  *ok = true;
  switch (code) {
    case kI: return which < 0 ? m2_eye() : m2_make(z, z, z, z);
    case kH: return m2_make(mkcd(h, 0), mkcd(h, 0), mkcd(h, 0), mkcd(-h, 0));
    case kX: return m2_make(z, one, one, z);
    case kY: return m2_make(z, mkcd(0, -1), mkcd(0, 1), z);
    case kZ: return m2_make(one, z, z, mkcd(-1, 0));
    case kS: return m2_make(one, z, z, mkcd(0, 1));
    case kT: return m2_make(one, z, z, mkcd(h, h));
    case kSX: return m2_make(mkcd(0.5, 0.5), mkcd(0.5, -0.5), mkcd(0.5, -0.5), mkcd(0.5, 0.5));
    case kRX: {
      double s, c;
      sincos(0.5 * p[0], &s, &c);
      if (which < 0) return m2_make(mkcd(c, 0), mkcd(0, -s), mkcd(0, -s), mkcd(c, 0));
      return m2_make(mkcd(-0.5 * s, 0), mkcd(0, -0.5 * c), mkcd(0, -0.5 * c), mkcd(-0.5 * s, 0));
    }
    case kRY: {
      double s, c;
      sincos(0.5 * p[0], &s, &c);
      if (which < 0) return m2_make(mkcd(c, 0), mkcd(-s, 0), mkcd(s, 0), mkcd(c, 0));
      return m2_make(mkcd(-0.5 * s, 0), mkcd(-0.5 * c, 0), mkcd(0.5 * c, 0), mkcd(-0.5 * s, 0));
    }
    case kRZ: {
      cd e0 = expi(-0.5 * p[0]), e1 = expi(0.5 * p[0]);
      if (which < 0) return m2_make(e0, z, z, e1);
      return m2_make(cd_mul(mkcd(0, -0.5), e0), z, z, cd_mul(mkcd(0, 0.5), e1));
    }
    case kRot: {
      // Rot(phi, theta, omega) = RZ(omega) RY(theta) RZ(phi)
      const double phi = p[0], th = p[1], om = p[2];
      double s, c;
      sincos(0.5 * th, &s, &c);
      cd a = expi(-0.5 * (phi + om)), b = expi(0.5 * (phi - om));
      cd cc = expi(-0.5 * (phi - om)), d = expi(0.5 * (phi + om));
      if (which < 0)
        return m2_make(mkcd(a.x * c, a.y * c), mkcd(-b.x * s, -b.y * s), mkcd(cc.x * s, cc.y * s),
                       mkcd(d.x * c, d.y * c));
      if (which == 1)
        return m2_make(mkcd(-0.5 * a.x * s, -0.5 * a.y * s), mkcd(-0.5 * b.x * c, -0.5 * b.y * c),
                       mkcd(0.5 * cc.x * c, 0.5 * cc.y * c), mkcd(-0.5 * d.x * s, -0.5 * d.y * s));
      // d/dphi and d/domega multiply each entry by i * (+-1/2)
      const double sa = -0.5, sb = (which == 0) ? 0.5 : -0.5, sc = (which == 0) ? -0.5 : 0.5, sd = 0.5;
      cd e00 = mkcd(a.x * c, a.y * c), e01 = mkcd(-b.x * s, -b.y * s);
      cd e10 = mkcd(cc.x * s, cc.y * s), e11 = mkcd(d.x * c, d.y * c);
      return m2_make(cd_mul(mkcd(0, sa), e00), cd_mul(mkcd(0, sb), e01), cd_mul(mkcd(0, sc), e10),
                     cd_mul(mkcd(0, sd), e11));
    }
    case kPhaseShift: {
      cd e = expi(p[0]);
      if (which < 0) return m2_make(one, z, z, e);
      return m2_make(z, z, z, cd_mul(mkcd(0, 1), e));
    }
    case kUnitary: {
      M2 r;
      for (int i = 0; i < 4; ++i) r.a[i] = mkcd(cpool[const_off + i].x, cpool[const_off + i].y);
      return r;
    }
    default:
      *ok = false;
      return m2_eye();
  }
}

template <typename PT>
__device__ __forceinline__ double param_value(int src, const PT* prow, const double2* cpool) {
  return src >= 0 ? static_cast<double>(prow[src]) : cpool[-1 - src].x;
}

template <typename PT>
__device__ M2 comp_eval(const CompRec& cr, const PT* prow, const double2* cpool, int which, bool* ok) {
  double p[3];
  p[0] = cr.src0 == INT32_MIN ? 0.0 : param_value(cr.src0, prow, cpool);
  p[1] = cr.src1 == INT32_MIN ? 0.0 : param_value(cr.src1, prow, cpool);
  p[2] = cr.src2 == INT32_MIN ? 0.0 : param_value(cr.src2, prow, cpool);
  M2 m = component_matrix(cr.code, p, cpool, cr.const_off, which, ok);
  return cr.dagger ? m2_dag(m) : m;
}

template <typename PT>
__global__ void k_build_tables(const PT* __restrict__ params, int64_t batch, int n_params,
                               const SlotRec* __restrict__ slots, int nslots, const CompRec* __restrict__ comps,
                               const double2* __restrict__ cpool, float2* __restrict__ mats, int64_t mat_stride,
                               double2* __restrict__ amats, int* __restrict__ err) {
  const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= batch * nslots) return;
  const int64_t b = gid / nslots;
  const int s = (int)(gid - b * nslots);
  const SlotRec sr = slots[s];
  const PT* prow = params + b * (int64_t)n_params;
  float2* out = mats + b * mat_stride + sr.mat_off;
  if (sr.kind == 1) {  // 4x4 user matrix, constant
    const CompRec cr = comps[sr.comp_start];
    for (int i = 0; i < 4; ++i)
      for (int j = 0; j < 4; ++j) {
        double2 v = cr.dagger ? cpool[cr.const_off + j * 4 + i] : cpool[cr.const_off + i * 4 + j];
        out[i * 4 + j] = make_float2((float)v.x, (float)(cr.dagger ? -v.y : v.y));
      }
    return;
  }
  bool ok = true, ok1;
  M2 S = m2_eye();
  for (int c = 0; c < sr.ncomp; ++c) {
    S = m2_mul(comp_eval(comps[sr.comp_start + c], prow, cpool, -1, &ok1), S);
    ok = ok && ok1;
  }
  for (int i = 0; i < 4; ++i) out[i] = make_float2((float)S.a[i].x, (float)S.a[i].y);
  if (sr.phase) {
    const cd p = cd_mul(S.a[3], cd_conj(S.a[0]));
    out[1] = make_float2((float)p.x, (float)p.y);
  }
  if (!ok) atomicOr(err, 1);
  if (amats == nullptr || sr.gslot < 0) return;
  if (sr.scalar > 0) {
    // component c = Re (even) or Im (odd) of C_{c/2}: A_{c/2} = 1/2 or -i/2, the other entries 0
    const int c = sr.scalar - 1;
    for (int k = 0; k < sr.ncomp; ++k) {
      const CompRec cr = comps[sr.comp_start + k];
      const int srcs[3] = {cr.src0, cr.src1, cr.src2};
      for (int w = 0; w < 3; ++w) {
        if (srcs[w] == INT32_MIN || srcs[w] < 0) continue;
        double2* dst = amats + (b * (int64_t)n_params + srcs[w]) * 4;
        for (int i = 0; i < 4; ++i)
          dst[i] = i == (c >> 1) ? ((c & 1) ? make_double2(0.0, -0.5) : make_double2(0.5, 0.0))
                                 : make_double2(0.0, 0.0);
      }
    }
    return;
  }
  const M2 Sd = m2_dag(S);
  M2 prev = m2_eye();  // S_{i-1}
  for (int c = 0; c < sr.ncomp; ++c) {
    const CompRec cr = comps[sr.comp_start + c];
    M2 U = comp_eval(cr, prow, cpool, -1, &ok1);
    M2 cur = m2_mul(U, prev);  // S_i
    const int srcs[3] = {cr.src0, cr.src1, cr.src2};
    for (int w = 0; w < 3; ++w) {
      if (srcs[w] == INT32_MIN || srcs[w] < 0) continue;
      M2 dU = comp_eval(cr, prow, cpool, w, &ok1);
      // A = S cur^dag dU prev S^dag
      M2 A = m2_mul(m2_mul(m2_mul(m2_mul(S, m2_dag(cur)), dU), prev), Sd);
      double2* dst = amats + (b * (int64_t)n_params + srcs[w]) * 4;
      for (int i = 0; i < 4; ++i) dst[i] = make_double2(A.a[i].x, A.a[i].y);
    }
    prev = cur;
  }
}


// ---------------------------------------------------------------------------
// Tensor-network leaves: full gate tensors for the ket (U, axes (out..., in...))
// and the bra (U^dagger) of every gate instance, per parameter row, plus the
// derivative 2x2 blocks dV/dtheta of every trainable parameter.  The reference
// builds the same operands with torch ops per gate per call
// (tedq/backends/pytorch_backend.py:524-546 for the conjugated copies).
// ---------------------------------------------------------------------------
// diag = w > 0: the network merged the gate's first w wires (in qubit order) into hyper-indices, the
// wires along which its matrix is diagonal.  w = nq: the tensor is its D diagonal entries; 0 < w < nq
// (the controls of a controlled gate): entry [c][t'][t] = M[(c, t'), (c, t)] over the w merged bits c
// and the remaining out / in bits t', t, 2^w x 4^(nq - w) entries; w = 0: the D x D matrix.
struct TNGateRec {
  int code, src0, src1, src2, const_off, nq, kind, bra, off, d0, d1, d2, dagger, diag, pad2, pad3;
};

template <typename PT>
__global__ void k_tn_gate_tensors(const PT* __restrict__ params, int64_t batch, int n_params,
                                  const TNGateRec* __restrict__ recs, int nrec, const double2* __restrict__ cpool,
                                  float2* __restrict__ leaves, int64_t leaf_stride, float2* __restrict__ dmats,
                                  int64_t dmat_stride, int* __restrict__ err) {
  const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= batch * nrec) return;
  const int64_t b = gid / nrec;
  const TNGateRec r = recs[gid - b * nrec];
  const PT* prow = params + b * (int64_t)n_params;
  float2* out = leaves + b * leaf_stride + r.off;
  const int D = 1 << r.nq;
  if (r.kind == 1) {  // constant matrix, row-major complex128
    if (r.diag == r.nq) {
      for (int i = 0; i < D; ++i) {
        const double2 v = cpool[r.const_off + i * D + i];
        out[i] = make_float2((float)v.x, (float)(r.bra ? -v.y : v.y));
      }
      return;
    }
    if (r.diag > 0) {
      const int f = r.nq - r.diag, F = 1 << f;       // free (unmerged) wires
      for (int e = 0; e < (1 << r.diag) * F * F; ++e) {
        const int c = e / (F * F), tp = (e / F) % F, t = e % F;
        const int i = c * F + tp, j = c * F + t;
        const double2 v = r.bra ? cpool[r.const_off + j * D + i] : cpool[r.const_off + i * D + j];
        out[e] = make_float2((float)v.x, (float)(r.bra ? -v.y : v.y));
      }
      return;
    }
    for (int i = 0; i < D; ++i)
      for (int j = 0; j < D; ++j) {
        const double2 v = r.bra ? cpool[r.const_off + j * D + i] : cpool[r.const_off + i * D + j];
        out[i * D + j] = make_float2((float)v.x, (float)(r.bra ? -v.y : v.y));
      }
    return;
  }
  CompRec cr;
  cr.code = r.code;
  cr.dagger = r.dagger;
  cr.src0 = r.src0;
  cr.src1 = r.src1;
  cr.src2 = r.src2;
  cr.const_off = r.const_off;
  bool ok = true;
  const M2 V = comp_eval(cr, prow, cpool, -1, &ok);
  if (!ok) atomicOr(err, 2);
  // controlled embedding: identity except the block where every control is 1
  if (r.diag == r.nq) {
    for (int i = 0; i < D; ++i) {
      const int bi = i - (D - 2);
      const cd v = bi >= 0 ? (r.bra ? cd_conj(V.a[bi * 3]) : V.a[bi * 3]) : mkcd(1.0, 0.0);
      out[i] = make_float2((float)v.x, (float)v.y);
    }
  } else if (r.diag > 0) {
    // controls merged (r.diag = nq - 1): [c][t'][t], the target block where every control is 1
    const int last = (1 << r.diag) - 1;
    for (int e = 0; e < (4 << r.diag); ++e) {
      const int c = e >> 2, a = (e >> 1) & 1, b2 = e & 1;
      cd v = mkcd(a == b2 ? 1.0 : 0.0, 0.0);
      if (c == last) v = r.bra ? cd_conj(V.a[b2 * 2 + a]) : V.a[a * 2 + b2];
      out[e] = make_float2((float)v.x, (float)v.y);
    }
  }
  for (int i = 0; i < D && !r.diag; ++i)
    for (int j = 0; j < D; ++j) {
      const int bi = i - (D - 2), bj = j - (D - 2);
      cd v;
      if (bi >= 0 && bj >= 0) {
        v = r.bra ? cd_conj(V.a[bj * 2 + bi]) : V.a[bi * 2 + bj];
      } else {
        v = mkcd(i == j ? 1.0 : 0.0, 0.0);
      }
      out[i * D + j] = make_float2((float)v.x, (float)v.y);
    }
  const int ds[3] = {r.d0, r.d1, r.d2};
  for (int w = 0; w < 3; ++w) {
    if (ds[w] < 0) continue;
    const M2 dV = comp_eval(cr, prow, cpool, w, &ok);
    float2* dst = dmats + b * dmat_stride + (int64_t)ds[w] * 4;
    for (int i = 0; i < 4; ++i) dst[i] = make_float2((float)dV.a[i].x, (float)dV.a[i].y);
  }
}

// d value / d theta_p = sum over the gate's ket tensor of E_ket . dU plus over
// its bra tensor of E_bra . dU^dagger, where E is the environment (the
// contraction of everything else).  Only the controlled block of dU is
// non-zero.  grad[b][p] += g[b] * Re(...), accumulated over networks.
struct TNParamRec {
  int ket_off, bra_off, nq, dslot;
};

__global__ void k_tn_param_grad(const float2* __restrict__ env, int64_t env_stride, const float2* __restrict__ dmats,
                                int64_t dmat_stride, const TNParamRec* __restrict__ recs, int np,
                                const float* __restrict__ upstream, double* __restrict__ grad, int64_t batch) {
  const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= batch * np) return;
  const int64_t b = gid / np;
  const int p = (int)(gid - b * np);
  const TNParamRec r = recs[p];
  if (r.dslot < 0) return;
  const float2* dv = dmats + b * dmat_stride + (int64_t)r.dslot * 4;
  const bool diag = r.nq & 16;   // leaves stored as diagonals: entry (i, i) at position i
  const bool ctrl = r.nq & 32;   // controls merged: target block entry (a, c) at (2^(nq-1) - 1) * 4 + 2a + c
  const int D = 1 << (r.nq & 15);
  double acc = 0.0;
  if (diag) {
    for (int a = 0; a < 2; ++a) {
      const int e = D - 2 + a;
      const float2 d = dv[a * 2 + a];
      if (r.ket_off >= 0) {
        const float2 x = env[b * env_stride + r.ket_off + e];
        acc += (double)x.x * d.x - (double)x.y * d.y;
      }
      if (r.bra_off >= 0) {
        const float2 x = env[b * env_stride + r.bra_off + e];
        acc += (double)x.x * d.x + (double)x.y * d.y;
      }
    }
    grad[b * np + p] += (double)upstream[b] * acc;
    return;
  }
  for (int a = 0; a < 2; ++a)
    for (int c = 0; c < 2; ++c) {
      const int e = ctrl ? (D / 2 - 1) * 4 + a * 2 + c : (D - 2 + a) * D + (D - 2 + c);
      const float2 d = dv[a * 2 + c];
      if (r.ket_off >= 0) {
        const float2 x = env[b * env_stride + r.ket_off + e];
        acc += (double)x.x * d.x - (double)x.y * d.y;
      }
      if (r.bra_off >= 0) {
        // bra tensor entry (a, c) is conj(U[c][a]); its derivative is conj(dV[c][a])
        const float2 x = env[b * env_stride + r.bra_off + e];
        const float2 dc = dv[c * 2 + a];
        acc += (double)x.x * dc.x + (double)x.y * dc.y;
      }
    }
  grad[b * np + p] += (double)upstream[b] * acc;
}

}  // namespace tedq

using namespace tedq;

extern "C" int tedq_sv_build_tables(const void* params, int params_are_f64, int64_t batch, int n_params,
                                    const void* slot_recs, int n_slots, const void* comp_recs,
                                    const void* const_pool, void* mats, int64_t mat_stride, void* amats,
                                    int* err_flag, void* stream) {
  if (batch <= 0 || n_slots < 0 || n_params < 0 || mats == nullptr || err_flag == nullptr) {
    set_error("tedq_sv_build_tables: bad arguments (batch=%lld slots=%d params=%d)", (long long)batch, n_slots,
              n_params);
    return kBadArgument;
  }
  if (n_slots == 0) return kOk;
  if (n_params > 0 && params == nullptr) {
    set_error("tedq_sv_build_tables: params pointer is null");
    return kBadArgument;
  }
  const int64_t total = batch * n_slots;
  const int threads = 128;
  const int64_t blocks = (total + threads - 1) / threads;
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (params_are_f64) {
    hipLaunchKernelGGL(k_build_tables<double>, dim3((unsigned)blocks), dim3(threads), 0, st,
                       static_cast<const double*>(params), batch, n_params, static_cast<const SlotRec*>(slot_recs),
                       n_slots, static_cast<const CompRec*>(comp_recs), static_cast<const double2*>(const_pool),
                       static_cast<float2*>(mats), mat_stride, static_cast<double2*>(amats), err_flag);
  } else {
    hipLaunchKernelGGL(k_build_tables<float>, dim3((unsigned)blocks), dim3(threads), 0, st,
                       static_cast<const float*>(params), batch, n_params, static_cast<const SlotRec*>(slot_recs),
                       n_slots, static_cast<const CompRec*>(comp_recs), static_cast<const double2*>(const_pool),
                       static_cast<float2*>(mats), mat_stride, static_cast<double2*>(amats), err_flag);
  }
  return check_launch("tedq_sv_build_tables");
}

extern "C" int tedq_tn_gate_tensors(const void* params, int params_are_f64, int64_t batch, int n_params,
                                    const void* recs, int nrec, const void* const_pool, void* leaves,
                                    int64_t leaf_stride, void* dmats, int64_t dmat_stride, int* err_flag,
                                    void* stream) {
  if (batch <= 0 || nrec < 0 || n_params < 0 || leaves == nullptr || err_flag == nullptr ||
      (n_params > 0 && params == nullptr)) {
    set_error("tedq_tn_gate_tensors: bad arguments (batch=%lld records=%d)", (long long)batch, nrec);
    return kBadArgument;
  }
  if (nrec == 0) return kOk;
  const int threads = 128;
  const int64_t blocks = (batch * nrec + threads - 1) / threads;
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (params_are_f64) {
    hipLaunchKernelGGL(k_tn_gate_tensors<double>, dim3((unsigned)blocks), dim3(threads), 0, st,
                       static_cast<const double*>(params), batch, n_params, static_cast<const TNGateRec*>(recs), nrec,
                       static_cast<const double2*>(const_pool), static_cast<float2*>(leaves), leaf_stride,
                       static_cast<float2*>(dmats), dmat_stride, err_flag);
  } else {
    hipLaunchKernelGGL(k_tn_gate_tensors<float>, dim3((unsigned)blocks), dim3(threads), 0, st,
                       static_cast<const float*>(params), batch, n_params, static_cast<const TNGateRec*>(recs), nrec,
                       static_cast<const double2*>(const_pool), static_cast<float2*>(leaves), leaf_stride,
                       static_cast<float2*>(dmats), dmat_stride, err_flag);
  }
  return check_launch("tedq_tn_gate_tensors");
}

extern "C" int tedq_tn_param_grad(const void* env, int64_t env_stride, const void* dmats, int64_t dmat_stride,
                                  const void* recs, int n_params, const float* upstream, double* grad, int64_t batch,
                                  void* stream) {
  if (env == nullptr || recs == nullptr || upstream == nullptr || grad == nullptr || batch <= 0 || n_params < 0) {
    set_error("tedq_tn_param_grad: bad arguments");
    return kBadArgument;
  }
  if (n_params == 0) return kOk;
  const int threads = 128;
  const int64_t blocks = (batch * n_params + threads - 1) / threads;
  hipLaunchKernelGGL(k_tn_param_grad, dim3((unsigned)blocks), dim3(threads), 0, static_cast<hipStream_t>(stream),
                     static_cast<const float2*>(env), env_stride, static_cast<const float2*>(dmats), dmat_stride,
                     static_cast<const TNParamRec*>(recs), n_params, upstream, grad, batch);
  return check_launch("tedq_tn_param_grad");
}
