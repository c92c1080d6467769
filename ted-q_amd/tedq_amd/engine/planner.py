"""Fusion planner: gate list -> device passes of LDS-staged tiles.

State layout (the reference's, pytorch_backend.py:513-522 with
compiled_circuit.py:134-139): amplitudes of one circuit are a contiguous
C-order [2]*n complex64 tensor, so qubit q is flat bit p = n-1-q and a batch of
circuits is [B][2^n].

A *pass* is one launch of the block kernel.  It owns a set of k physical bits
(the *tile* bits, always including the lowest bits so that global loads are
contiguous runs); every workgroup loads one 2^k tile per circuit into LDS,
applies all gates of the pass that act inside the tile, and writes it back
once.  Inside a pass, gates are grouped into *stages*: each stage names r tile
bits that are held in registers (R = 2^r amplitudes per thread), so every
gate of the stage is pure VALU work on registers; stages are separated by one
LDS gather/scatter.

Pass selection is a greedy topological sweep: a gate may be scheduled once it
is the next pending gate on all its qubits, so gates on disjoint qubits are
freely reordered (they commute) and the CNOT chains of hardware-efficient
ansatze fold into about two passes per layer.
"""

from dataclasses import dataclass, field

import numpy as np

# Op codes shared with csrc/sv_ops.h (kOp* constants there).
OP_U1, OP_D1, OP_CU1, OP_CX, OP_CZ, OP_U2, OP_GRAD1, OP_CGRAD1, OP_CD1 = 1, 2, 3, 4, 5, 6, 7, 8, 9
OP_R1, OP_PGRAD, OP_GFLUSH, OP_PH1 = 10, 11, 12, 13
OP_DAGGER = 1 << 8

# kinds of OP_PGRAD (PauliGradKind in csrc/sv_ops.h) by gate code: RX, RY, RZ, PhaseShift blocks
PAULI_KIND = {14: 0, 15: 1, 16: 2, 18: 3}
GROUP_SIZE = 8                                  # gradient values per OP_PGRAD group (one partial row)

# Slot kinds shared with csrc/sv_tables.hip.
SLOT_1Q, SLOT_2Q = 0, 1

# Component sources: a parameter is either a trainable column of the params
# row (>= 0) or a constant in the fp64 constant pool (encoded -1 - index).

_DIAGONAL_CODES = {0, 4, 5, 6, 16, 18}          # I, Z, S, T, RZ, PhaseShift
_REAL_CODES = {0, 1, 2, 4, 15}                  # I, H, X, Z, RY
_ONEQ_CODES = {0, 1, 2, 3, 4, 5, 6, 7, 14, 15, 16, 17, 18, 23}
_PARAM_COUNT = {14: 1, 15: 1, 16: 1, 17: 3, 18: 1}


@dataclass
class Component:
    """One gate factor of a fused slot: its code, dagger flag and parameter sources."""
    code: int
    dagger: bool
    srcs: tuple            # parameter sources (see module note), up to 3
    const_mat: int = -1    # constant-pool offset of a user matrix (Unitary) or -1


@dataclass
class Slot:
    """A device matrix: product of components (1q, 2x2) or one 4x4 user unitary."""
    kind: int
    comps: list
    mat_off: int = 0       # complex64 offset inside one batch row of the table
    gslot: int = -1        # gradient slot index or -1
    scomp: int = -1        # component inside gslot of a single-value gradient (OP_PGRAD), else -1

    @property
    def dim(self):
        return 2 if self.kind == SLOT_1Q else 4

    @property
    def trainable(self):
        return [s for c in self.comps for s in c.srcs if s >= 0]

    @property
    def diagonal(self):
        return self.kind == SLOT_1Q and all(c.code in _DIAGONAL_CODES for c in self.comps)

    @property
    def real(self):
        return self.kind == SLOT_1Q and all(c.code in _REAL_CODES and c.const_mat < 0 for c in self.comps)

    @property
    def pauli(self):
        """(kind, negate) when the slot is one trainable rotation whose gradient is a single value
        (RX, RY, RZ, PhaseShift and their controlled blocks), else None."""
        if self.kind != SLOT_1Q or len(self.comps) != 1:
            return None
        c = self.comps[0]
        if c.code not in PAULI_KIND or c.const_mat >= 0 or len(c.srcs) != 1 or c.srcs[0] < 0:
            return None
        return PAULI_KIND[c.code], bool(c.dagger)


@dataclass
class Unit:
    """A gate as the kernels see it: physical bits it touches and how to apply it."""
    kind: str              # "u1", "cu1", "cx", "cz"
    bits: tuple            # for u1/cu1/cx: (target, *controls); cz: all bits
    comp: Component = None
    index: int = 0


@dataclass
class Stage:
    reg_bits: list         # local tile-bit indices held in registers (len r)
    ops: list = field(default_factory=list)   # tuples (type, b0, b1, cmask, slot_index)


@dataclass
class Pass:
    tile_bits: list        # physical bit positions, ascending (len k)
    stages: list = field(default_factory=list)


@dataclass
class Plan:
    n: int
    k: int
    r: int
    passes: list
    slots: list
    const_pool: np.ndarray
    n_params: int
    n_gslots: int
    mat_stride: int        # complex64 entries per batch row of the slot table
    specialize: bool = False   # adjoint plan of generated kernels (see build_plan)


def gate_to_units(n, gate, srcs, const_pool):
    """Lower one traced gate to kernel units.  ``srcs`` are its parameter sources."""
    code, qubits = gate.gate_code, list(gate.qubits)
    bits = [n - 1 - q for q in qubits]
    dagger = bool(getattr(gate, "is_inverse", False))
    name = gate.name[:-4] if dagger else gate.name

    def comp(c, s=(), mat=-1):
        return Component(c, dagger, tuple(s), mat)

    # a user matrix already reflects ``inverse()`` (GateBase.inverse daggers it)
    if code in _ONEQ_CODES and len(bits) == 1:
        if code == 23:
            return [Unit("u1", (bits[0],), _const_comp(gate.matrix, const_pool))]
        return [Unit("u1", (bits[0],), comp(code, srcs))]
    if name == "CNOT":
        return [Unit("cx", (bits[1], bits[0]))]
    if name == "Toffoli":
        return [Unit("cx", (bits[2], bits[0], bits[1]))]
    if name == "CZ":
        return [Unit("cz", (bits[0], bits[1]))]
    if name == "SWAP":
        a, b = bits
        return [Unit("cx", (b, a)), Unit("cx", (a, b)), Unit("cx", (b, a))]
    if name == "CSWAP":
        c, a, b = bits
        return [Unit("cx", (a, b)), Unit("cx", (b, c, a)), Unit("cx", (a, b))]
    if name == "CY":
        return [Unit("cu1", (bits[1], bits[0]), comp(3))]
    if name in ("CRX", "CRY", "CRZ", "ControlledPhaseShift"):
        block = {"CRX": 14, "CRY": 15, "CRZ": 16, "ControlledPhaseShift": 18}[name]
        return [Unit("cu1", (bits[1], bits[0]), comp(block, srcs))]
    if code == 23 and len(bits) == 2:
        return two_qubit_units(np.asarray(gate.matrix, dtype=np.complex128).reshape(4, 4), bits[0], bits[1],
                               const_pool)
    raise NotImplementedError(f"{gate.name} on {len(bits)} qubits is not supported by the SV engine")


def _const_comp(mat, const_pool):
    off = len(const_pool)
    const_pool.extend(np.asarray(mat, dtype=np.complex128).reshape(-1).tolist())
    return Component(23, False, (), off)


def _multiplexor(target, control, m0, m1, const_pool):
    """Apply m0 to ``target`` where ``control`` is 0 and m1 where it is 1."""
    return [Unit("cu1", (target, control), _const_comp(m1, const_pool)),
            Unit("cx", (control,)),
            Unit("cu1", (target, control), _const_comp(m0, const_pool)),
            Unit("cx", (control,))]


def two_qubit_units(m, b_hi, b_lo, const_pool):
    """Exact lowering of a 4x4 operator on physical bits (b_hi = first qubit, b_lo = second).

    Cosine-sine decomposition m = diag(U0, U1) . CS . diag(V0, V1), where the
    diagonal blocks act on the second qubit multiplexed by the first and CS is
    a rotation of the first qubit multiplexed by the second; each multiplexor
    is two controlled single-qubit ops.
    """
    from scipy.linalg import cossin  # pylint: disable=import-outside-toplevel
    u, cs, vdh = cossin(m, p=2, q=2)
    units = _multiplexor(b_lo, b_hi, vdh[:2, :2], vdh[2:, 2:], const_pool)
    r0 = np.array([[cs[0, 0], cs[0, 2]], [cs[2, 0], cs[2, 2]]])
    r1 = np.array([[cs[1, 1], cs[1, 3]], [cs[3, 1], cs[3, 3]]])
    units += _multiplexor(b_hi, b_lo, r0, r1, const_pool)
    units += _multiplexor(b_lo, b_hi, u[:2, :2], u[2:, 2:], const_pool)
    return units


def _is_diagonal_unit(u):
    if u.kind == "cz":
        return True
    return u.kind in ("u1", "cu1") and u.comp.const_mat < 0 and u.comp.code in _DIAGONAL_CODES


def fold_trailing(units, zmasks):
    """Fold trailing gates into Z-string observables (every measurement is some <psi| Z_m |psi>).

    A unit that no kept unit after it shares a bit with commutes to the end of the circuit, where
    <psi| G^dagger Z_m G |psi> takes its place.  Scanning backwards:
      - a unit on none of the observables' bits leaves them unchanged (G^dagger Z_m G = Z_m): dropped;
      - a diagonal unit (Z, S, T, RZ, PhaseShift, CZ, controlled diagonals) commutes with them: dropped;
      - a CNOT (target t, control c) maps Z_t to Z_c Z_t and Z_c to itself: every mask holding bit t
        toggles bit c, and the unit is dropped;
      - any other unit (and a user matrix, which need not be unitary) is kept, and its bits close.
    This is exact: values are unchanged, and the gradient of a parameter whose gates were all dropped is
    zero.  Returns (kept units in order, new masks, number of dropped units).
    """
    masks = list(zmasks)
    closed = set()
    keep = [True] * len(units)
    for i in range(len(units) - 1, -1, -1):
        u = units[i]
        bits = set(u.bits)
        if bits & closed:
            closed |= bits
            continue
        support = 0
        for m in masks:
            support |= m
        known = u.comp is None or u.comp.const_mat < 0
        if known and not any((support >> b) & 1 for b in bits):
            keep[i] = False
        elif _is_diagonal_unit(u):
            keep[i] = False
        elif u.kind == "cx" and len(u.bits) == 2:
            t, c = u.bits
            masks = [m ^ (1 << c) if (m >> t) & 1 else m for m in masks]
            keep[i] = False
        else:
            closed |= bits
    return [u for u, k in zip(units, keep) if k], masks, keep.count(False)


def compact_register(units, zmasks, n):
    """Drop qubits no unit touches from a circuit that starts in |0..0> and measures Z-strings only.

    Such a qubit stays |0> and contributes a factor <0|Z|0> = 1 to every Z-string, so its bit leaves the
    masks and the state needs only the used bits (renumbered in order, at least one).  Returns
    (units, masks, new n)."""
    used = sorted({b for u in units for b in u.bits}) or [0]
    if len(used) == n:
        return units, list(zmasks), n
    new = {b: i for i, b in enumerate(used)}
    out = [Unit(u.kind, tuple(new[b] for b in u.bits), u.comp, u.index) for u in units]
    masks = []
    for m in zmasks:
        masks.append(sum(1 << new[b] for b in used if (m >> b) & 1))
    return out, masks, len(used)


def _schedule(units, order, max_bits, seed_bits, allowed_bits=None):
    """Greedy topological sweep.

    ``order`` lists candidate unit indices in program order; returns the
    scheduled indices (a valid order) and the chosen bit set.  A unit is ready
    when it heads the pending queue of all its bits.
    """
    pending = {}
    for ui in order:
        for b in units[ui].bits:
            pending.setdefault(b, []).append(ui)
    heads = {b: 0 for b in pending}
    chosen = set(seed_bits)
    done = []
    remaining = list(order)
    progress = True
    while progress:
        progress = False
        nxt = []
        for ui in remaining:
            u = units[ui]
            if any(pending[b][heads[b]] != ui for b in u.bits):
                nxt.append(ui)
                continue
            ub = set(u.bits)
            if allowed_bits is not None and not ub <= allowed_bits:
                nxt.append(ui)
                continue
            if not ub <= chosen:
                if len(chosen | ub) > max_bits:
                    nxt.append(ui)
                    continue
                chosen |= ub
            for b in u.bits:
                heads[b] += 1
            done.append(ui)
            progress = True
        remaining = nxt
    return done, chosen


def build_plan(n, units, n_params, const_pool, k_max=12, r=4, coalesce_bits=3, specialize=False):
    """Partition ``units`` (program order) into passes and register stages.

    With ``specialize`` (the adjoint's plan) a trainable single-qubit gate is never fused with another
    gate, so that a rotation keeps its own slot: real gates then run as real 2x2 ops and the gradient
    of RX / RY / RZ / PhaseShift (and controlled forms) is one value per parameter (OP_PGRAD, 2-4
    multiply-adds per amplitude pair) instead of the slot's 8 bilinears (16 per pair), collected in
    groups of 8 per register stage.  Its uncontrolled diagonal gates run as phases on the |1> half
    (OP_PH1, one complex multiply per pair instead of two; see program.py)."""
    k = min(n, k_max)
    widest = max((len(u.bits) for u in units), default=1)
    if widest > k:
        raise ValueError(f"a gate acts on {widest} qubits but tiles hold only {k}")
    r = min(k, max(r, widest))
    slots = []
    passes = []
    todo = list(range(len(units)))
    while todo:
        # seed the tile with the lowest bits (contiguous global runs); shrink the
        # seed when it leaves no room for the next gate
        for c in range(min(n, coalesce_bits), -1, -1):
            done, chosen = _schedule(units, todo, k, list(range(c)))
            if done:
                break
        if not done:
            raise RuntimeError("planner made no progress")
        for b in range(n):
            if len(chosen) >= k:
                break
            chosen.add(b)
        tile_bits = sorted(chosen)
        done_set = set(done)
        todo = [u for u in todo if u not in done_set]
        passes.append(_build_pass(units, done, tile_bits, r, slots, specialize))
    mat_off = 0
    for s in slots:
        s.mat_off = mat_off
        mat_off += s.dim * s.dim
    gslot = 0
    if specialize:
        # single-value gradients share partial rows: groups of 8 per register stage, numbered in the
        # adjoint's order (stage ops reversed)
        for p in passes:
            for st in p.stages:
                group, used = -1, GROUP_SIZE
                for op in reversed(st.ops):
                    si = op[4]
                    if si < 0 or not slots[si].trainable:
                        continue
                    if slots[si].pauli is not None:
                        if used == GROUP_SIZE:
                            group, used = gslot, 0
                            gslot += 1
                        slots[si].gslot, slots[si].scomp = group, used
                        used += 1
                    else:
                        slots[si].gslot = gslot
                        gslot += 1
    else:
        for s in slots:
            if s.trainable:
                s.gslot = gslot
                gslot += 1
    return Plan(n=n, k=k, r=r, passes=passes, slots=slots,
                const_pool=np.asarray(const_pool, dtype=np.complex128),
                n_params=n_params, n_gslots=gslot, mat_stride=max(mat_off, 1), specialize=specialize)


def _build_pass(units, pass_units, tile_bits, r, slots, specialize=False):
    local = {b: i for i, b in enumerate(tile_bits)}
    allowed = set(tile_bits)
    p = Pass(tile_bits=tile_bits)
    todo = list(pass_units)
    while todo:
        done, chosen = _schedule(units, todo, r, [], allowed)
        if not done:
            raise RuntimeError("stage planner made no progress")
        for b in tile_bits:
            if len(chosen) >= r:
                break
            chosen.add(b)
        reg_phys = sorted(chosen)
        reg_of = {b: i for i, b in enumerate(reg_phys)}
        stage = Stage(reg_bits=[local[b] for b in reg_phys])
        last_on_bit = {}
        for ui in done:
            u = units[ui]
            if u.kind == "u1":
                b = u.bits[0]
                prev = last_on_bit.get(b)
                if prev is not None and stage.ops[prev][0] == "u1" and (
                        not specialize or (not slots[stage.ops[prev][4]].trainable
                                           and not any(x >= 0 for x in u.comp.srcs))):
                    slots[stage.ops[prev][4]].comps.append(u.comp)
                    continue
                slots.append(Slot(SLOT_1Q, [u.comp]))
                stage.ops.append(("u1", reg_of[b], 0, 0, len(slots) - 1))
            elif u.kind == "cu1":
                cmask = sum(1 << reg_of[c] for c in u.bits[1:])
                slots.append(Slot(SLOT_1Q, [u.comp]))
                stage.ops.append(("cu1", reg_of[u.bits[0]], 0, cmask, len(slots) - 1))
            elif u.kind == "cx":
                cmask = sum(1 << reg_of[c] for c in u.bits[1:])
                stage.ops.append(("cx", reg_of[u.bits[0]], 0, cmask, -1))
            elif u.kind == "cz":
                mask = sum(1 << reg_of[c] for c in u.bits)
                stage.ops.append(("cz", 0, 0, mask, -1))
            for b in u.bits:
                last_on_bit[b] = len(stage.ops) - 1
        p.stages.append(stage)
        done_set = set(done)
        todo = [u for u in todo if u not in done_set]
    return p
