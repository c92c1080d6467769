"""Serialise a fusion plan into the int32 programs read by the HIP kernels.

Formats are documented next to the kernels (csrc/sv_passes.hip header for
pass programs, csrc/sv_tables.hip for slot/component records); the constants
here must match them.
"""

from dataclasses import dataclass

import numpy as np

from .planner import (OP_U1, OP_D1, OP_CU1, OP_CX, OP_CZ, OP_GRAD1, OP_CGRAD1, OP_CD1, OP_R1, OP_PGRAD,
                      OP_GFLUSH, OP_PH1, OP_DAGGER, SLOT_1Q)

HDR_TILE_BITS = 4
HDR_HI_OFF = 36
HDR_STAGES = 100
STAGE_WORDS = 48
OP_WORDS = 8
NO_SRC = -(2 ** 31)


@dataclass
class EncodedPlan:
    """Host copies of everything the kernels need for one circuit structure."""
    prog: np.ndarray          # int32 words: forward pass programs, then adjoint ones
    fwd_offsets: list         # word offset of each forward pass
    adj_offsets: list         # word offset of each adjoint pass (reverse pass order)
    tile_masks: list          # physical tile mask per pass (forward order)
    slot_recs: np.ndarray     # int32 [nslots, 8]
    comp_recs: np.ndarray     # int32 [ncomps, 8]
    const_pool: np.ndarray    # float64 [npool, 2]
    pgslot: np.ndarray        # int32 [n_params]: gradient slot of each trainable parameter


def lds_swizzle(e):
    """LDS element swizzle of the pass kernels (csrc/tedq_common.h ``swz``); an XOR-linear involution."""
    x = (e >> 4) & 15
    lo = (x & 6) | ((x ^ (x >> 2)) & 1) | ((bin(x).count("1") & 1) << 3)
    return e ^ (lo ^ ((e >> 8) & 15))


def _pdep(v, positions):
    out = 0
    for i, p in enumerate(positions):
        if (v >> i) & 1:
            out |= 1 << p
    return out


def _phase_slot(sl, specialize):
    """Diagonal 1-qubit slot of a specialised plan: its table row also holds p = d1 conj(d0) in the
    unused entry 1 (U = d0 diag(1, p) for U = diag(d0, d1)), read by OP_PH1."""
    return specialize and sl.kind == SLOT_1Q and sl.diagonal


def _fwd_op_words(op, slots, specialize=False):
    """Forward op words.  In a specialised plan an uncontrolled diagonal U = d0 diag(1, p) is applied
    without its global phase d0 (OP_PH1: the |1> half times p); expectation values do not see a global
    phase, and the adjoint sweep, which applies (diag(1, p))^dagger, undoes exactly this map."""
    kind, b0, b1, cmask, si = op
    mat = slots[si].mat_off if si >= 0 else 0
    if kind == "u1" and cmask == 0 and _phase_slot(slots[si], specialize):
        return [OP_PH1, b0, 0, 0, mat, -1, 0, 0]
    if kind in ("u1", "cu1"):
        sl = slots[si]
        if sl.diagonal:
            t = OP_D1 if kind == "u1" else OP_CD1
        elif sl.real:
            t = OP_R1
        else:
            t = OP_U1 if kind == "u1" else OP_CU1
    elif kind == "cx":
        t = OP_CX
    elif kind == "cz":
        t = OP_CZ
    else:
        raise ValueError(kind)
    return [t, b0, b1, cmask, mat, -1, 0, 0]


def _adj_op_words(op, slots, specialize=False):
    """Reverse-sweep ops for one forward op: gradient bilinears first, then U^dagger.

    In a specialised plan an uncontrolled diagonal U^dagger = conj(d0) diag(1, conj(p)) is applied
    without its global phase conj(d0) (OP_PH1 with the dagger flag).  psi and lambda receive the same
    phase, so every bilinear conj(lambda_a) psi_b, hence every gradient, is unchanged."""
    kind, b0, b1, cmask, si = op
    words = []
    if si >= 0 and slots[si].gslot >= 0:
        sl = slots[si]
        if sl.scomp >= 0:
            pk, neg = sl.pauli
            words.append([OP_PGRAD, b0, 0, cmask, 0, sl.gslot, pk | (int(neg) << 4), sl.scomp])
        else:
            g = OP_GRAD1 if kind == "u1" else OP_CGRAD1
            words.append([g, b0, 0, cmask, 0, sl.gslot, 0, 0])
    w = _fwd_op_words(op, slots, specialize)
    w[0] |= OP_DAGGER
    words.append(w)
    return words


def _encode_pass(p, k, r, slots, adjoint, specialize=False):
    nst = len(p.stages)
    stages = list(reversed(p.stages)) if adjoint else p.stages
    header = np.zeros(HDR_STAGES + STAGE_WORDS * nst, dtype=np.int64)
    tT = k - r
    header[0], header[1], header[2], header[3] = k, r, nst, tT
    for i, b in enumerate(p.tile_bits):
        header[HDR_TILE_BITS + i] = b
    hi = np.zeros(32, dtype=np.int64)
    for m in range(1 << r):
        hi[m] = _pdep(m << tT, p.tile_bits)
    op_words = []
    op_start = len(header)
    for s_idx, st in enumerate(stages):
        base = HDR_STAGES + STAGE_WORDS * s_idx
        ops = []
        seq = reversed(st.ops) if adjoint else st.ops
        pending = -1          # gradient group collecting OP_PGRAD values
        for op in seq:
            if adjoint:
                for w in _adj_op_words(op, slots, specialize):
                    if w[0] == OP_PGRAD and w[5] != pending:
                        if pending >= 0:
                            ops.append([OP_GFLUSH, 0, 0, 0, 0, pending, 0, 0])
                        pending = w[5]
                    ops.append(w)
            else:
                ops.append(_fwd_op_words(op, slots, specialize))
        if pending >= 0:
            ops.append([OP_GFLUSH, 0, 0, 0, 0, pending, 0, 0])
        regmask = sum(1 << b for b in st.reg_bits)
        header[base + 0] = len(ops)
        header[base + 1] = op_start + len(op_words) * OP_WORDS
        header[base + 2] = ((1 << k) - 1) & ~regmask
        header[base + 3] = regmask
        for m in range(1 << r):
            header[base + 4 + m] = lds_swizzle(_pdep(m, st.reg_bits))
        op_words.extend(ops)
    words = header.astype(np.int32)
    words[HDR_HI_OFF:HDR_HI_OFF + 64] = hi.view(np.int32)
    body = np.asarray(op_words, dtype=np.int32).reshape(-1) if op_words else np.zeros(0, np.int32)
    out = np.concatenate([words, body])
    if len(out) % 2:
        out = np.concatenate([out, np.zeros(1, np.int32)])
    return out


def encode_plan(plan):
    """Serialise forward and adjoint pass programs plus the slot tables."""
    chunks, fwd_off, adj_off, masks = [], [], [], []
    pos = 0
    for p in plan.passes:
        w = _encode_pass(p, plan.k, plan.r, plan.slots, adjoint=False, specialize=plan.specialize)
        fwd_off.append(pos)
        chunks.append(w)
        pos += len(w)
        masks.append(sum(1 << b for b in p.tile_bits))
    for p in reversed(plan.passes):
        w = _encode_pass(p, plan.k, plan.r, plan.slots, adjoint=True, specialize=plan.specialize)
        adj_off.append(pos)
        chunks.append(w)
        pos += len(w)
    prog = np.concatenate(chunks) if chunks else np.zeros(2, np.int32)

    slot_recs = np.zeros((max(len(plan.slots), 1), 8), dtype=np.int32)
    comp_list = []
    pgslot = np.full(max(plan.n_params, 1), -1, dtype=np.int32)
    for si, s in enumerate(plan.slots):
        slot_recs[si] = [s.kind, len(s.comps), len(comp_list), s.mat_off, s.gslot, s.scomp + 1,
                         int(_phase_slot(s, plan.specialize)), 0]
        for c in s.comps:
            srcs = list(c.srcs) + [NO_SRC] * (3 - len(c.srcs))
            comp_list.append([c.code, int(c.dagger), srcs[0], srcs[1], srcs[2], c.const_mat, 0, 0])
            for src in c.srcs:
                if src >= 0:
                    pgslot[src] = s.gslot
        assert s.kind == SLOT_1Q or len(s.comps) == 1
    comp_recs = np.asarray(comp_list if comp_list else [[0] * 8], dtype=np.int32).reshape(-1, 8)
    pool = plan.const_pool if len(plan.const_pool) else np.zeros(1, np.complex128)
    const_pool = np.stack([pool.real, pool.imag], axis=1).astype(np.float64)
    return EncodedPlan(prog=prog.astype(np.int32), fwd_offsets=fwd_off, adj_offsets=adj_off,
                       tile_masks=masks, slot_recs=slot_recs, comp_recs=comp_recs,
                       const_pool=np.ascontiguousarray(const_pool), pgslot=pgslot)
