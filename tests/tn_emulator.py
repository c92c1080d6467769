"""CPU emulation of the tensor-network entry points, for host-logic tests.

Replaces the device calls of ``TNExecutor`` (gate tensors, one contraction
step, a level of batched small steps, a fused gate backward, parameter
gradients) with numpy code that reads the same
records and descriptors, so the executor's layouts, descriptors, light-cone
reduction, path and reverse-tree bookkeeping run unchanged on the CPU.  The
gate matrices come from the oracle; the contraction is a direct sum over the
descriptor's index roles.  Test infrastructure only.
"""

import numpy as np
import torch

from oracle import sv_oracle as O
from tedq_amd.tn import executor as X

_CODE_NAME = {14: "RX", 15: "RY", 16: "RZ", 17: "Rot", 18: "PhaseShift"}


def _scat(pos, nbits, x):
    off = np.zeros_like(x)
    for j in range(nbits):
        off |= ((x >> j) & 1) << int(pos[j])
    return off


def contract_np(desc, A, B, C):
    """C[b] (flat, in place) from A[b], B[b] per the descriptor; arrays are complex [batch, size]."""
    nH, nM, nN, nK = (int(v) for v in desc[:4])
    pos = desc[4:].reshape(9, 32)
    o = np.arange(1 << (nH + nM + nN), dtype=np.int64)
    n = o & ((1 << nN) - 1)
    m = (o >> nN) & ((1 << nM) - 1)
    h = o >> (nN + nM)
    k = np.arange(1 << nK, dtype=np.int64)
    oa = _scat(pos[0], nH, h) | _scat(pos[1], nM, m)
    ob = _scat(pos[3], nH, h) | _scat(pos[5], nN, n)
    oc = _scat(pos[6], nH, h) | _scat(pos[7], nM, m) | _scat(pos[8], nN, n)
    ka = _scat(pos[2], nK, k)
    kb = _scat(pos[4], nK, k)
    for b in range(C.shape[0]):
        a = A[b if A.shape[0] > 1 else 0]
        bb = B[b if B.shape[0] > 1 else 0]
        vals = (a[oa[:, None] | ka[None, :]] * bb[ob[:, None] | kb[None, :]]).sum(1)
        C[b, oc] = vals


def _view(t, stride, off, size, batch):
    arr = t.detach().numpy().reshape(-1) if t.dim() else t.numpy()
    if stride == 0:
        return arr[off:off + size][None, :]
    rows = arr.reshape(batch, stride) if t.numel() == batch * stride else arr.reshape(-1, stride)
    return rows[:, off:off + size]


def _pack(T, nq, dw):
    """Leaf entries of a gate matrix whose first dw wires are merged: the diagonal (dw = nq), the
    [merged][out][in] array (0 < dw < nq) or the whole matrix."""
    if dw == nq:
        return np.diag(T)
    f = nq - dw
    blocks = T.reshape(1 << dw, 1 << f, 1 << dw, 1 << f)
    return np.stack([blocks[c, :, c, :] for c in range(1 << dw)]).reshape(-1)


def install(ex):
    """Patch ``ex`` (a TNExecutor) to run its device calls on the CPU."""

    def leaves(rows):
        r = rows.detach().to(torch.float64).numpy()
        B = r.shape[0]
        lv = np.zeros((B, ex.leaf_stride), dtype=np.complex64)
        dm = np.zeros((B, max(ex.n_params, 1) * 4), dtype=np.complex64)
        pool = ex._pool  # pylint: disable=protected-access
        for rec in ex._recs:  # pylint: disable=protected-access
            code, s0, s1, s2, coff, nq, kind, bra, off, d0, d1, d2, dag, diag = (int(v) for v in rec[:14])
            D = 1 << nq
            for b in range(B):
                if kind == 1:
                    U = pool[coff:coff + D * D].reshape(D, D)
                    T = U.conj().T if bra else U
                    v = _pack(T, nq, diag)
                    lv[b, off:off + v.size] = v
                    continue

                def val(s):
                    return r[b, s] if s >= 0 else pool[-1 - s].real

                ps = [val(s) for s in (s0, s1, s2) if s != X._NO_SRC]  # pylint: disable=protected-access
                name = _CODE_NAME[code]
                V = O.gate_matrix(name, [torch.tensor(p, dtype=torch.float64) for p in ps],
                                  torch.complex128).numpy()
                if dag:
                    V = V.conj().T
                U = np.eye(D, dtype=np.complex128)
                U[D - 2:, D - 2:] = V
                T = U.conj().T if bra else U
                v = _pack(T, nq, diag)
                lv[b, off:off + v.size] = v
                for w, d in enumerate((d0, d1, d2)):
                    if d < 0:
                        continue
                    pp = [torch.tensor(p, dtype=torch.float64, requires_grad=True) for p in ps]
                    Vt = O.gate_matrix(name, pp, torch.complex128)
                    dV = np.zeros((2, 2), dtype=np.complex128)
                    for i in range(2):
                        for j in range(2):
                            gr = torch.autograd.grad(Vt[i, j].real, pp[w], retain_graph=True, allow_unused=True)[0]
                            gi = torch.autograd.grad(Vt[i, j].imag, pp[w], retain_graph=True, allow_unused=True)[0]
                            dV[i, j] = (0.0 if gr is None else float(gr)) + 1j * (0.0 if gi is None else float(gi))
                    if dag:
                        dV = dV.conj().T
                    dm[b, d * 4:d * 4 + 4] = dV.reshape(-1)
        return torch.from_numpy(lv), torch.from_numpy(dm)

    def contract(row, a, b, c, B, dev, buf):  # pylint: disable=unused-argument
        desc = ex._descs[row]  # pylint: disable=protected-access
        nH, nM, nN, nK = (int(v) for v in desc[:4])
        (ta, sa, oa), (tb, sb, ob), (tc, sc, oc) = a, b, c
        A = _view(ta, sa, oa, 1 << (nH + nM + nK), B)
        Bv = _view(tb, sb, ob, 1 << (nH + nK + nN), B)
        size_c = 1 << (nH + nM + nN)
        out = np.zeros((B, size_c), dtype=np.complex128)
        contract_np(desc, A.astype(np.complex128), Bv.astype(np.complex128), out)
        cv = tc.numpy().reshape(B, sc)
        cv[:, oc:oc + size_c] = out.astype(np.complex64)

    def small_levels(plan, arena, leaves, buf, B, dev):  # pylint: disable=unused-argument
        """tedq_tn_contract_small_batch: every record of every level, in level order."""
        consts = buf["consts"].numpy()
        bases = [leaves.detach().numpy().reshape(B, -1), consts.reshape(1, -1), arena.numpy()]
        for s, e in plan["bounds"]:
            for rec in plan["recs"][s:e]:
                row, ia, oa, ib, ob, oc, size, _ = (int(v) for v in rec)
                desc = ex._descs[row]  # pylint: disable=protected-access
                nH, nM, nN, nK = (int(v) for v in desc[:4])
                A = bases[ia][:, oa:oa + (1 << (nH + nM + nK))].astype(np.complex128)
                Bv = bases[ib][:, ob:ob + (1 << (nH + nN + nK))].astype(np.complex128)
                out = np.zeros((B, size), dtype=np.complex128)
                contract_np(desc, A, Bv, out)
                bases[2][:, oc:oc + size] = out.astype(np.complex64)

    def buffers(device):  # pylint: disable=unused-argument
        return {"consts": torch.from_numpy(ex._consts)}  # pylint: disable=protected-access

    def param_grad(env_arena, dmats, grad, B):
        env = env_arena.numpy()
        dm = dmats.numpy()
        g = grad.numpy()
        for p, (ket, bra, nq, ds) in enumerate(ex._param_recs):  # pylint: disable=protected-access
            diag = bool(int(nq) & 16)
            ctrl = bool(int(nq) & 32)
            D = 1 << (int(nq) & 15)
            for b in range(B):
                dV = dm[b, ds * 4:ds * 4 + 4].astype(np.complex128).reshape(2, 2)
                if diag:
                    Ek = np.diag(env[b, ket + D - 2:ket + D].astype(np.complex128))
                    Eb = np.diag(env[b, bra + D - 2:bra + D].astype(np.complex128))
                elif ctrl:
                    # [controls][out][in]: the block where every control is 1 is the last 2 x 2
                    Ek = env[b, ket:ket + 2 * D].reshape(-1, 2, 2)[-1].astype(np.complex128)
                    Eb = env[b, bra:bra + 2 * D].reshape(-1, 2, 2)[-1].astype(np.complex128)
                else:
                    Ek = env[b, ket:ket + D * D].reshape(D, D)[D - 2:, D - 2:]
                    Eb = env[b, bra:bra + D * D].reshape(D, D)[D - 2:, D - 2:]
                g[b, p] += float(np.real((Ek * dV).sum() + (Eb * dV.conj().T).sum()))

    ex._leaves = leaves  # pylint: disable=protected-access
    ex._contract_raw = contract  # pylint: disable=protected-access
    ex._small_levels = small_levels  # pylint: disable=protected-access

    def gate_backward(st, ecop, aop, bop, dst_a, dst_b, B, dev, buf):
        """tedq_tn_gate_backward as its two environment steps."""
        contract(st.bwd_a, ecop, bop, dst_a, B, dev, buf)
        contract(st.bwd_b, aop, ecop, dst_b, B, dev, buf)

    ex._gate_backward = gate_backward  # pylint: disable=protected-access
    ex._buffers = buffers  # pylint: disable=protected-access
    ex._param_grad = param_grad  # pylint: disable=protected-access
    return ex
