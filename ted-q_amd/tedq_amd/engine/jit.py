"""Per-circuit pass kernels: source generation and run-time compilation.

The library's pass kernels interpret a pass program (op list, offsets,
matrix positions) at run time.  For a circuit structure evaluated many
times, this module turns the same encoded programs into one HIP kernel per
pass in which every op, register bit, control mask, LDS offset and matrix
position is a constant, and compiles them with hiprtc through the library
(tedq_jit_compile, csrc/sv_jit.hip).  The generated code calls the device
templates of csrc/sv_ops.h with the same arguments the interpreter derives
from the program words, and reads and writes memory exactly as the
interpreted kernels do (same tile placement, same LDS layout, same partial
rows), so the two agree to rounding.

Limits: tiles narrower than the circuit (k < n), n <= 28, and at most
64 KiB of LDS per workgroup; other programs keep the interpreted kernels.
"""

import ctypes
import os

import numpy as np

from . import hip
from .planner import (OP_U1, OP_D1, OP_CU1, OP_CX, OP_CZ, OP_GRAD1, OP_CGRAD1, OP_CD1, OP_R1, OP_PGRAD, OP_GFLUSH,
                      OP_PH1)
from .program import HDR_TILE_BITS, HDR_HI_OFF, HDR_STAGES, STAGE_WORDS, OP_WORDS, lds_swizzle

_CSRC = os.path.join(os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__)))), "csrc")
_HEADERS = ("tedq_common.h", "sv_ops.h")


# generated adjoint kernels loop over tiles with a register prefetch of the next tile (kPassPersistent)
# where one workgroup's LDS fills a CU (see persistent_pays)
PERSISTENT_ADJOINT = True


def persistent_pays(k, adjoint):
    """Whether generated kernels of this tile size loop over tiles with a register prefetch: only where
    one workgroup's LDS fills a CU (an adjoint tile of 2^13 amplitudes or more takes 128 KiB of the 160).
    There the load of a tile otherwise waits with nothing to overlap it (20q ansatz, adjoint 2^13:
    30.5 -> 27.0 ms); where two workgroups share a CU they overlap each other and the loop only costs
    registers (adjoint 2^12: 30.3 -> 32.1 ms; forward 2^13, 64 KiB: 11.85 -> 12.16 ms)."""
    return PERSISTENT_ADJOINT and adjoint and (1 << k) * 16 > 80 * 1024


def supported(n, k, r, adjoint):
    lds = (1 << k) * 8 * (2 if adjoint else 1)
    return k < n <= 28 and lds <= 160 * 1024 and k - r <= 9 and 1 <= r <= 5


def _bits(mask):
    return [i for i in range(64) if (mask >> i) & 1]


def _pdep_expr(var, positions):
    """C expression depositing the low bits of ``var`` at ``positions`` (ascending bit i -> positions[i])."""
    if not positions:
        return "0u"
    terms = []
    for i, p in enumerate(positions):
        if p >= i:
            terms.append(f"((({var}) & {1 << i}u) << {p - i})")
        else:
            terms.append(f"((({var}) & {1 << i}u) >> {i - p})")
    return " | ".join(terms)


def _u1_args(dag, off):
    m = [f"mb[{off + i}]" for i in range(4)]
    if dag:
        return f"cconj({m[0]}), cconj({m[2]}), cconj({m[1]}), cconj({m[3]})"
    return ", ".join(m)


def _diag_args(dag, off):
    if dag:
        return f"cconj(mb[{off}]), cconj(mb[{off + 3}])"
    return f"mb[{off}], mb[{off + 3}]"


def _reduce_write_lines(var, seg):
    """Sum the 8 values ``var`` over the tile's lanes and write them to ``dst`` (one partial row)."""
    out = []
    if seg == 64:
        out.append(f"  const float tot = reduce8_wave({var}, lane);")
        out.append("  if ((lane & 7) == 0) dst[(lane >> 3) & 7] = tot;")
    else:
        out.append("#pragma unroll")
        out.append("  for (int q = 0; q < 8; ++q)")
        out.append(f"    for (int o = 1; o < {seg}; o <<= 1) {var}[q] += __shfl_xor({var}[q], o, 64);")
        out.append(f"  if ((tid & {seg - 1}) == 0) {{")
        out.append("    float4* d4 = reinterpret_cast<float4*>(dst);")
        out.append(f"    d4[0] = make_float4({var}[0], {var}[1], {var}[2], {var}[3]);")
        out.append(f"    d4[1] = make_float4({var}[4], {var}[5], {var}[6], {var}[7]);")
        out.append("  }")
    return out


def _op_lines(w, R, regs, adjoint, seg):
    """Source lines of one op (8 program words) applied to the register arrays ``regs``."""
    t, dag = int(w[0]) & 0xFF, (int(w[0]) >> 8) & 1
    b, cm, off, gslot = int(w[1]), int(w[3]), int(w[4]), int(w[5])
    out = []
    if t == OP_R1:
        m = [f"mb[{off + i}].x" for i in range(4)]
        if dag:
            m[1], m[2] = m[2], m[1]
        for v in regs:
            out.append(f"R1Op<{R}, {b}, {cm}>::run({v}, {', '.join(m)});")
    elif t == OP_PGRAD:
        if not adjoint:
            raise ValueError("gradient op in a forward program")
        kind, neg, comp = int(w[6]) & 15, int(w[6]) >> 4, int(w[7])
        out.append(f"PauliGradOp<{R}, {b}, {cm}, {kind}, {'true' if neg else 'false'}>::run(vp, vl, sacc[{comp}]);")
    elif t == OP_GFLUSH:
        if not adjoint:
            raise ValueError("gradient op in a forward program")
        out.append("{")
        out.append(f"  float* dst = pbase + (int64_t){gslot} * a.np * 8;")
        out.extend(_reduce_write_lines("sacc", seg))
        out.append("#pragma unroll")
        out.append("  for (int q = 0; q < 8; ++q) sacc[q] = 0.f;")
        out.append("}")
    elif t in (OP_U1, OP_CU1):
        out.append("{")
        for v in regs:
            out.append(f"  U1Op<{R}, {b}, {cm}>::run({v}, {_u1_args(dag, off)});")
        out.append("}")
    elif t in (OP_D1, OP_CD1):
        for v in regs:
            out.append(f"DiagOp<{R}, {b}, {cm}>::run({v}, {_diag_args(dag, off)});")
    elif t == OP_PH1:
        for v in regs:
            p = f"cconj(mb[{off + 1}])" if dag else f"mb[{off + 1}]"
            out.append(f"DiagOp<{R}, {b}, {cm | (1 << b)}>::run({v}, {p}, {p});")
    elif t == OP_CX:
        for v in regs:
            out.append(f"CXOp<{R}, {b}, {cm}>::run({v});")
    elif t == OP_CZ:
        mask = cm
        top = mask.bit_length() - 1
        for v in regs:
            out.append(f"CZOp<{R}, {top}, {mask & ~(1 << top)}>::run({v});")
    elif t in (OP_GRAD1, OP_CGRAD1):
        if not adjoint:
            raise ValueError("gradient op in a forward program")
        out.append("{")
        out.append("  float acc[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};")
        out.append(f"  GradOp<{R}, {b}, {cm}>::run(vp, vl, acc);")
        out.append(f"  float* dst = pbase + (int64_t){gslot} * a.np * 8;")
        out.extend(_reduce_write_lines("acc", seg))
        out.append("}")
    else:
        raise ValueError(f"op type {t} has no generated form")
    return out


def _seed_lines(hi, R, masks):
    """Per-register lambda / psi factors d_m = sum_j c_j (-1)^popc(x & mask_j) for x = rest bits | lane
    bits | register m's bits: the rest and lane parities per thread, the register parities as constants."""
    out = ["  const double* __restrict__ cb = seed + circ * %d;" % len(masks),
           "  const uint64_t xr = (uint64_t)(gbase - (circ << a.n)) | (uint64_t)lo;"]
    for j, mk in enumerate(masks):
        out.append(f"  const float t{j} = (__popcll(xr & {mk}ull) & 1) ? -(float)cb[{j}] : (float)cb[{j}];")
    for m in range(R):
        terms = []
        for j, mk in enumerate(masks):
            neg = bin(int(hi[m]) & mk).count("1") & 1
            terms.append(("- " if neg else "+ ") + f"t{j}")
        expr = " ".join(terms).lstrip("+ ") if terms else "0.f"
        if expr.startswith("- "):
            expr = "-" + expr[2:]
        out.append(f"  const float d{m} = {expr};")
    return out


def pass_kernel_source(name, words, adjoint, seed_masks=None, fresh=0, persistent=False):
    """HIP source of one generated pass kernel from its program words (program.py layout).

    ``seed_masks`` (adjoint only): instead of loading lambda, form it from the loaded psi as
    lambda(x) = sum_j seed[circuit][j] (-1)^popc(x & mask_j) psi(x) -- the adjoint seed of Z-string
    expectation values -- for the first pass of the sweep.
    ``fresh`` (forward only): the tile bits no earlier pass held.  Under kPassFromZero (a sweep from
    |0..0>) the state is zero wherever one of them is 1, so those amplitudes are set instead of loaded;
    when every tile bit is fresh (the first pass) the tile is |0..0> itself and nothing is loaded.
    ``persistent`` (adjoint only; launched with kPassPersistent): each workgroup loops over the tiles
    blockIdx.x, blockIdx.x + gridDim.x, ... and loads the next tile's psi and lambda into registers
    while the current one is processed, so the load of a tile overlaps the arithmetic of the previous
    one even where a single workgroup fills a CU's LDS."""
    words = np.asarray(words, dtype=np.int32)
    k, r, nst, tT = (int(x) for x in words[:4])
    R, T = 1 << r, 1 << tT
    tile_bits = [int(x) for x in words[HDR_TILE_BITS:HDR_TILE_BITS + k]]
    hi = words[HDR_HI_OFF:HDR_HI_OFF + 64].view(np.int64)
    seg = min(T, 64)
    W = max(T >> 6, 1)
    persistent = persistent and adjoint
    seeded = adjoint and seed_masks is not None
    tile = "TILE" if persistent else "(int64_t)blockIdx.x"
    L = [f'extern "C" __global__ void __launch_bounds__({T}) {name}(',
         "    float2* __restrict__ psi, float2* __restrict__ lam, const float2* __restrict__ mats,",
         "    float* __restrict__ partials, tedq::PassArgs a, const double* __restrict__ seed) {",
         "  using namespace tedq;",
         "  extern __shared__ float2 lds[];"]
    thread_lines = [f"  const uint32_t lo = {_pdep_expr('(uint32_t)tid', tile_bits[:tT])};",
                    "  const int sw = swz(tid);"]
    if adjoint:
        thread_lines += [f"  float2* ldl = lds + {T * R};", "  const int lane = tid & 63;", "  (void)lane;"]
    if not persistent:
        L += ["  const int tid = threadIdx.x;"] + thread_lines
    else:
        # registers holding the next tile (psi, and lambda unless the seed forms it from psi)
        regs_in = [("qp", "psi")] + ([] if seeded else [("ql", "lam")])

        def prefetch(expr, lo_var):
            out = [f"  {{ int64_t c_, g_; tile_origin(a, {expr}, &c_, &g_); (void)c_;"]
            for v, buf in regs_in:
                out.append(f"    const __amdgpu_buffer_rsrc_t f{v} = make_rsrc({buf} + g_, (uint32_t)(8u << a.n));")
                out += [f"    {v}[{m}] = bload(f{v}, {lo_var} * 8u, {int(hi[m]) * 8}u);" for m in range(R)]
            return out + ["  }"]
        L += [f"  float2 {v}[{R}];" for v, _ in regs_in]
        L.append(f"  const uint32_t lo0 = {_pdep_expr('(uint32_t)threadIdx.x', tile_bits[:tT])};")
        L += prefetch("(int64_t)blockIdx.x", "lo0")
        L.append("  for (int64_t TILE = blockIdx.x; TILE < a.total_tiles; TILE += gridDim.x) {")
        # per-thread offsets are recomputed in every iteration: hoisted out of the loop, the LDS offsets of
        # all stages would stay live across it (about 90 more VGPRs, and spills with the prefetch)
        L += ["  int tid = threadIdx.x;", '  asm volatile("" : "+v"(tid));'] + thread_lines
    L += ["  int64_t circ, gbase;",
          f"  tile_origin(a, {tile}, &circ, &gbase);",
          "  const float2* __restrict__ mb = mats + circ * a.mat_stride;",
          "  const __amdgpu_buffer_rsrc_t rp = make_rsrc(psi + gbase, (uint32_t)(8u << a.n));"]
    if adjoint:
        L += ["  const __amdgpu_buffer_rsrc_t rl = make_rsrc(lam + gbase, (uint32_t)(8u << a.n));",
              f"  const int64_t prow = ({tile} & ((1ll << tiles_log2(a)) - 1)) * {W} + (tid >> 6);",
              "  float* const pbase = partials + (circ * a.n_gslots * a.np + prow) * 8;",
              "  float sacc[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};",
              "  (void)sacc;"]
    if seeded:
        L += _seed_lines(hi, R, seed_masks)
        for m in range(R):
            x = lds_swizzle(m * T)
            src = f"qp[{m}]" if persistent else f"bload(rp, lo * 8u, {int(hi[m]) * 8}u)"
            L.append(f"  {{ const float2 v = {src}; lds[sw ^ {x}] = v;"
                     f" ldl[sw ^ {x}] = make_float2(d{m} * v.x, d{m} * v.y); }}")
    elif persistent:
        for m in range(R):
            x = lds_swizzle(m * T)
            L.append(f"  lds[sw ^ {x}] = qp[{m}]; ldl[sw ^ {x}] = ql[{m}];")
    L.append("  (void)seed;")
    tile_mask = sum(1 << b for b in tile_bits)
    if not adjoint and fresh:
        L.append("  if (a.flags & kPassFromZero) {")
        for m in range(R):
            x = lds_swizzle(m * T)
            if fresh == tile_mask:
                L.append(f"    lds[sw ^ {x}] = make_float2({'tid == 0 ? 1.f : 0.f' if m == 0 else '0.f'}, 0.f);")
            elif int(hi[m]) & fresh:
                L.append(f"    lds[sw ^ {x}] = make_float2(0.f, 0.f);")
            else:
                L.append(f"    {{ float2 v = make_float2(0.f, 0.f); if ((lo & {fresh}u) == 0u) v = bload(rp, lo * 8u, "
                         f"{int(hi[m]) * 8}u); lds[sw ^ {x}] = v; }}")
        L.append("  } else {")
    for m in range(0 if seeded or persistent else R):
        L.append(f"  lds[sw ^ {lds_swizzle(m * T)}] = bload(rp, lo * 8u, {int(hi[m]) * 8}u);")
        if adjoint:
            L.append(f"  ldl[sw ^ {lds_swizzle(m * T)}] = bload(rl, lo * 8u, {int(hi[m]) * 8}u);")
    if not adjoint and fresh:
        L.append("  }")
    L.append("  __syncthreads();")
    if persistent:
        # the next tile's loads are in flight during this tile's stages
        L.append("  if (TILE + gridDim.x < a.total_tiles)")
        L += prefetch("TILE + gridDim.x", "lo")
    regs = ("vp", "vl") if adjoint else ("v",)
    for s in range(nst):
        base = HDR_STAGES + STAGE_WORDS * s
        nops, op_off, nonreg = int(words[base]), int(words[base + 1]), int(words[base + 2])
        offs = [int(x) for x in words[base + 4:base + 4 + R]]
        L.append("  {")
        # the thread index is re-read through an empty asm in every stage: otherwise the compiler computes
        # the LDS offsets of all stages up front and keeps them live (up to ~100 more VGPRs in long passes)
        L.append('    int ts = threadIdx.x; asm volatile("" : "+v"(ts));')
        L.append(f"    const int sb = swz({_pdep_expr('ts', _bits(nonreg))});")
        for v in regs:
            L.append(f"    float2 {v}[{R}];")
        for m in range(R):
            for v, arr in zip(regs, ("lds", "ldl")):
                L.append(f"    {v}[{m}] = {arr}[sb ^ {offs[m]}];")
        for o in range(nops):
            w = words[op_off + o * OP_WORDS: op_off + (o + 1) * OP_WORDS]
            L.extend("    " + x for x in _op_lines(w, R, regs, adjoint, seg))
        for m in range(R):
            for v, arr in zip(regs, ("lds", "ldl")):
                L.append(f"    {arr}[sb ^ {offs[m]}] = {v}[{m}];")
        L.append("    __syncthreads();")
        L.append("  }")
    if adjoint:
        # kPassPsiReadOnly: psi is a checkpoint of the forward; the state at this pass's start is the
        # previous checkpoint, so psi is not written back
        L.append("  if (!(a.flags & kPassPsiReadOnly)) {")
        for m in range(R):
            L.append(f"    bstore(lds[sw ^ {lds_swizzle(m * T)}], rp, lo * 8u, {int(hi[m]) * 8}u);")
        L.append("  }")
        for m in range(R):
            L.append(f"  bstore(ldl[sw ^ {lds_swizzle(m * T)}], rl, lo * 8u, {int(hi[m]) * 8}u);")
        if persistent:
            # this tile's stores read LDS, which the next tile's first writes overwrite
            L += ["  __syncthreads();", "  }"]
    else:
        # kPassOutOfPlace: the result goes to the second state pointer (a checkpoint), psi keeps its input
        L.append("  const __amdgpu_buffer_rsrc_t ro = make_rsrc(((a.flags & kPassOutOfPlace) ? lam : psi) + gbase,"
                 " (uint32_t)(8u << a.n));")
        for m in range(R):
            L.append(f"  bstore(lds[sw ^ {lds_swizzle(m * T)}], ro, lo * 8u, {int(hi[m]) * 8}u);")
    L.append("}")
    return "\n".join(L)


def program_source(enc, n_fwd, n_adj, seed_masks=None, persistent=False, only=None):
    """Generated source of every pass of an encoded plan: kernels f0.. (forward) and a0.. (adjoint), and
    with ``seed_masks`` also a0s, the first adjoint pass forming lambda from psi (pass_kernel_source);
    ``persistent``: the adjoint kernels loop over tiles with a register prefetch (kPassPersistent);
    ``only``: the kernel names to emit (default all)."""
    parts = ['#include "sv_ops.h"', ""]
    offs = list(enc.fwd_offsets) + list(enc.adj_offsets) + [len(enc.prog)]
    starts = sorted(set(offs))
    ends = {o: starts[i + 1] for i, o in enumerate(starts[:-1])}
    seen = 0
    for i in range(n_fwd):
        o = enc.fwd_offsets[i]
        if only is None or f"f{i}" in only:
            parts.append(pass_kernel_source(f"f{i}", enc.prog[o:ends[o]], False, fresh=enc.tile_masks[i] & ~seen))
        seen |= enc.tile_masks[i]
    for i in range(n_adj):
        o = enc.adj_offsets[i]
        if only is None or f"a{i}" in only:
            parts.append(pass_kernel_source(f"a{i}", enc.prog[o:ends[o]], True, persistent=persistent))
    if seed_masks is not None and n_adj and (only is None or "a0s" in only):
        o = enc.adj_offsets[0]
        parts.append(pass_kernel_source("a0s", enc.prog[o:ends[o]], True, seed_masks, persistent=persistent))
    return "\n\n".join(parts) + "\n"


def scratch_bytes(fn):
    """Per-thread scratch bytes of a compiled kernel (nonzero: the compiler spilled registers)."""
    out = ctypes.c_int(0)
    hip.call("tedq_jit_function_scratch", fn, ctypes.byref(out))
    return out.value


class JitModule:
    """A compiled module of generated pass kernels on the current device.

    Modules are never unloaded: launches are asynchronous, so a module dropped
    by the host may still have kernels running.
    """

    def __init__(self, source, names):
        texts = []
        for h in _HEADERS:
            with open(os.path.join(_CSRC, h), "rb") as f:
                texts.append(f.read() + b"\0")
        srcs = (ctypes.c_char_p * len(texts))(*texts)
        hnames = (ctypes.c_char_p * len(_HEADERS))(*[h.encode() for h in _HEADERS])
        mod = ctypes.c_void_p()
        hip.call("tedq_jit_compile", source.encode(), len(_HEADERS), srcs, hnames, ctypes.byref(mod))
        self.module = mod
        self.functions = {}
        for name in names:
            fn = ctypes.c_void_p()
            hip.call("tedq_jit_function", mod, name.encode(), ctypes.byref(fn))
            self.functions[name] = fn
