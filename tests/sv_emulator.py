"""CPU emulation of the engine's device programs (test helper).

Decodes the int32 pass programs and slot records produced by
``tedq_amd.engine.program`` and executes them with the kernels' semantics —
tile gather by physical bit set, register stages, op codes, dagger flag,
adjoint bilinears and the final ``2 Re sum A_ij C_ij`` — in complex128
numpy.  Gate matrices come from the oracle, so agreement with the oracle's
tensordot loop checks the planner, the encoder and the gradient algebra
without a GPU.
"""

import numpy as np
import torch

from oracle.sv_oracle import gate_matrix
from tedq_amd.engine import program as P
from tedq_amd.engine.planner import (OP_U1, OP_D1, OP_CU1, OP_CX, OP_CZ, OP_U2, OP_GRAD1, OP_CGRAD1, OP_CD1,
                                     OP_R1, OP_PGRAD, OP_GFLUSH, OP_PH1)

CODE_NAMES = {0: "I", 1: "Hadamard", 2: "PauliX", 3: "PauliY", 4: "PauliZ", 5: "S", 6: "T", 7: "SX",
              14: "RX", 15: "RY", 16: "RZ", 17: "Rot", 18: "PhaseShift"}


def _comp_matrix(rec, row, pool, which=-1):
    code, dagger = int(rec[0]), bool(rec[1])
    srcs = [int(s) for s in rec[2:5]]
    if code == 23:
        off = int(rec[5])
        m = (pool[off:off + 4, 0] + 1j * pool[off:off + 4, 1]).reshape(2, 2)
        return m.conj().T if dagger else m
    vals = []
    for s in srcs:
        if s == P.NO_SRC:
            continue
        vals.append(float(row[s]) if s >= 0 else float(pool[-1 - s, 0]))
    name = CODE_NAMES[code]
    if which < 0:
        m = gate_matrix(name, vals).detach().numpy()
    else:
        t = torch.tensor(vals, dtype=torch.float64, requires_grad=True)

        def f(x):
            return torch.view_as_real(gate_matrix(name, list(x)))
        jac = torch.autograd.functional.jacobian(f, t).numpy()   # [2,2,2(re/im), nparams]
        m = jac[..., 0, which] + 1j * jac[..., 1, which]
    return m.conj().T if dagger else m


def build_tables(enc, plan, row):
    """Slot matrices (complex128) and A matrices per trainable column for one parameter row."""
    pool = enc.const_pool
    mats = np.zeros(plan.mat_stride, dtype=np.complex128)
    amats = {}
    for si, s in enumerate(plan.slots):
        rec = enc.slot_recs[si]
        kind, ncomp, cstart, moff = int(rec[0]), int(rec[1]), int(rec[2]), int(rec[3])
        comps = enc.comp_recs[cstart:cstart + ncomp]
        if kind == 1:
            off = int(comps[0][5])
            m = (pool[off:off + 16, 0] + 1j * pool[off:off + 16, 1]).reshape(4, 4)
            if comps[0][1]:
                m = m.conj().T
            mats[moff:moff + 16] = m.reshape(-1)
            continue
        Us = [_comp_matrix(c, row, pool) for c in comps]
        S = np.eye(2, dtype=np.complex128)
        for U in Us:
            S = U @ S
        mats[moff:moff + 4] = S.reshape(-1)
        if int(rec[6]):
            mats[moff + 1] = S[1, 1] * np.conj(S[0, 0])
        scalar = int(rec[5])
        if scalar > 0:
            # single-value gradient in component c of the gslot: A picks it out of 2 Re sum A C
            c = scalar - 1
            A = np.zeros(4, dtype=np.complex128)
            A[c >> 1] = -0.5j if c & 1 else 0.5
            for comp in comps:
                for src in (int(x) for x in comp[2:5]):
                    if src != P.NO_SRC and src >= 0:
                        amats[src] = A.reshape(2, 2)
            continue
        prev = np.eye(2, dtype=np.complex128)
        for c, U in zip(comps, Us):
            cur = U @ prev
            srcs = [int(x) for x in c[2:5] if int(x) != P.NO_SRC]
            for w, src in enumerate(srcs):
                if src >= 0:
                    dU = _comp_matrix(c, row, pool, which=w)
                    amats[src] = S @ cur.conj().T @ dU @ prev @ S.conj().T
            prev = cur
    return mats, amats


def _pdep(v, positions):
    out = 0
    for i, p in enumerate(positions):
        if (v >> i) & 1:
            out |= 1 << p
    return out


def _apply_ops(regs_list, ops, mats, acc):
    """regs_list: list of arrays [threads, R] (psi, and lambda in the adjoint)."""
    R = regs_list[0].shape[1]
    for op in ops:
        typ, dag = int(op[0]) & 0xFF, (int(op[0]) >> 8) & 1
        b0, b1, cmask, moff, gslot = int(op[1]), int(op[2]), int(op[3]), int(op[4]), int(op[5])
        if typ == OP_GFLUSH:
            continue
        if typ == OP_PH1:
            ph = np.conj(mats[moff + 1]) if dag else mats[moff + 1]
            for v in regs_list:
                for j in range(R):
                    if (j >> b0) & 1 and (j & cmask) == cmask:
                        v[:, j] = ph * v[:, j]
            continue
        if typ == OP_PGRAD:
            p, l = regs_list
            kind, neg, comp = int(op[6]) & 15, int(op[6]) >> 4, int(op[7])
            val = 0.0
            for i in range(R):
                if i & (1 << b0) or (i & cmask) != cmask:
                    continue
                j = i | (1 << b0)
                c = lambda a, b: np.sum(np.conj(l[:, a]) * p[:, b])  # noqa: E731
                if kind == 0:
                    val += (c(i, j) + c(j, i)).imag
                elif kind == 1:
                    val += (c(j, i) - c(i, j)).real
                elif kind == 2:
                    val += (c(i, i) - c(j, j)).imag
                else:
                    val += -2.0 * c(j, j).imag
            if neg:
                val = -val
            cm = np.zeros(4, dtype=np.complex128)
            cm[comp >> 1] = val if comp % 2 == 0 else 1j * val
            acc[gslot] = acc.get(gslot, 0) + cm.reshape(2, 2)
            continue
        if typ in (OP_GRAD1, OP_CGRAD1):
            p, l = regs_list
            c = np.zeros((2, 2), dtype=np.complex128)
            for i in range(R):
                if i & (1 << b0) or (i & cmask) != cmask:
                    continue
                j = i | (1 << b0)
                idx = (i, j)
                for a in range(2):
                    for b in range(2):
                        c[a, b] += np.sum(np.conj(l[:, idx[a]]) * p[:, idx[b]])
            acc[gslot] = acc.get(gslot, 0) + c
            continue
        for v in regs_list:
            if typ in (OP_U1, OP_CU1, OP_D1, OP_CD1, OP_R1):
                m = mats[moff:moff + 4].reshape(2, 2)
                if typ in (OP_D1, OP_CD1):
                    m = np.diag(np.diag(m))
                if typ == OP_R1:
                    assert np.allclose(m.imag, 0)
                    m = m.real.astype(np.complex128)
                if dag:
                    m = m.conj().T
                for i in range(R):
                    if i & (1 << b0) or (i & cmask) != cmask:
                        continue
                    j = i | (1 << b0)
                    x, y = v[:, i].copy(), v[:, j].copy()
                    v[:, i] = m[0, 0] * x + m[0, 1] * y
                    v[:, j] = m[1, 0] * x + m[1, 1] * y
            elif typ == OP_CX:
                for i in range(R):
                    if i & (1 << b0) or (i & cmask) != cmask:
                        continue
                    j = i | (1 << b0)
                    v[:, [i, j]] = v[:, [j, i]]
            elif typ == OP_CZ:
                for i in range(R):
                    if (i & cmask) == cmask:
                        v[:, i] = -v[:, i]
            elif typ == OP_U2:
                m = mats[moff:moff + 16].reshape(4, 4)
                if dag:
                    m = m.conj().T
                for i in range(R):
                    if i & ((1 << b0) | (1 << b1)):
                        continue
                    idx = [i, i | (1 << b1), i | (1 << b0), i | (1 << b0) | (1 << b1)]
                    x = v[:, idx].copy()
                    v[:, idx] = x @ m.T
            else:
                raise ValueError(f"unknown op {typ}")


def run_pass(words, n, states, mats, acc, zero=0, fresh=0):
    """Execute one encoded pass over each state vector in ``states`` (in place); tiles whose non-tile bits
    meet ``zero`` are left untouched (the kernels' zero mask).  ``fresh`` (a forward sweep from |0..0>,
    kPassFromZero): amplitudes at local positions meeting these tile bits are taken as zero instead of
    read, and when every tile bit is fresh the tile is |0..0> (nothing read)."""
    k, r, nst = int(words[0]), int(words[1]), int(words[2])
    tile_bits = [int(b) for b in words[P.HDR_TILE_BITS:P.HDR_TILE_BITS + k]]
    rest = [b for b in range(n) if b not in tile_bits]
    R = 1 << r
    local_phys = np.array([_pdep(j, tile_bits) for j in range(1 << k)], dtype=np.int64)
    first = fresh != 0 and fresh == sum(1 << b for b in tile_bits)
    for t in range(1 << (n - k)):
        base = _pdep(t, rest)
        if base & zero:
            continue
        tiles = [s[base + local_phys].copy() for s in states]
        if first:
            for tile in tiles:
                tile[:] = 0
                tile[0] = 1
        elif fresh:
            for tile in tiles:
                tile[(local_phys & fresh) != 0] = 0
        for si in range(nst):
            st = words[P.HDR_STAGES + P.STAGE_WORDS * si:P.HDR_STAGES + P.STAGE_WORDS * (si + 1)]
            nops, ostart, nonreg = int(st[0]), int(st[1]), int(st[2])
            nonreg_bits = [b for b in range(k) if (nonreg >> b) & 1]
            reg_off = np.array([P.lds_swizzle(int(x)) for x in st[4:4 + R]], dtype=np.int64)
            thread_base = np.array([_pdep(tl, nonreg_bits) for tl in range(1 << (k - r))], dtype=np.int64)
            idx = thread_base[:, None] + reg_off[None, :]
            regs = [tile[idx] for tile in tiles]
            ops = words[ostart:ostart + nops * P.OP_WORDS].reshape(nops, P.OP_WORDS)
            _apply_ops(regs, ops, mats, acc)
            for tile, rg in zip(tiles, regs):
                tile[idx] = rg
        for s, tile in zip(states, tiles):
            s[base + local_phys] = tile


def run_pass_out(words, n, src, dst, mats, zero=0, fresh=0):
    """One forward pass read from ``src`` and written to ``dst`` (out of place), tiles meeting ``zero``
    skipped (neither read nor written); ``fresh`` as in run_pass."""
    k = int(words[0])
    tile_bits = [int(b) for b in words[P.HDR_TILE_BITS:P.HDR_TILE_BITS + k]]
    rest = [b for b in range(n) if b not in tile_bits]
    local_phys = np.array([_pdep(j, tile_bits) for j in range(1 << k)], dtype=np.int64)
    work = np.zeros_like(src)
    for t in range(1 << (n - k)):
        base = _pdep(t, rest)
        if base & zero:
            continue
        work[base + local_phys] = src[base + local_phys]
    run_pass(words, n, [work], mats, {}, zero, fresh)
    for t in range(1 << (n - k)):
        base = _pdep(t, rest)
        if base & zero:
            continue
        dst[base + local_phys] = work[base + local_phys]


def forward(enc, plan, row, init=None, zero_masks=None):
    n = plan.n
    mats, amats = build_tables(enc, plan, row)
    psi = np.zeros(2 ** n, dtype=np.complex128)
    if init is None:
        psi[0] = 1
    else:
        psi[:] = np.asarray(init).reshape(-1)
    for i, off in enumerate(enc.fwd_offsets):
        run_pass(enc.prog[off:], n, [psi], mats, {}, zero_masks[i] if zero_masks else 0)
    return psi, mats, amats


def adjoint_grad(enc, plan, psi, lam, mats, amats, n_params, zero_masks=None, ckpt=None):
    """Reverse sweep; ``ckpt`` (the forward's state after every pass): each pass reads psi from the
    checkpoint at its end instead of the state the previous adjoint pass left."""
    acc = {}
    psi, lam = psi.copy(), lam.copy()
    P = len(enc.adj_offsets)
    for i, off in enumerate(enc.adj_offsets):
        if ckpt is not None:
            psi = ckpt[P - 1 - i].copy()
        run_pass(enc.prog[off:], plan.n, [psi, lam], mats, acc, zero_masks[P - 1 - i] if zero_masks else 0)
    grad = np.zeros(n_params)
    for p in range(n_params):
        g = int(enc.pgslot[p])
        if g < 0 or p not in amats:
            continue
        grad[p] = 2 * np.real(np.sum(amats[p] * acc[g]))
    return grad, psi
