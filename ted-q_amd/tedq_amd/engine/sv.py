"""State-vector executor: one compiled circuit structure, any batch of parameter rows.

Forward (replaces tedq/backends/pytorch_backend.py:358-391):
    tables  <- gate matrices of every fused slot for every parameter row (fp64 -> c64)
    psi     <- |0..0> (or the traced initial state) for every row
    psi     <- pass_P ... pass_1 psi              (one fused LDS-tiled kernel per pass)
    results <- FP64 compensated reductions (all Z-string observables in one pass),
               factored observables, probabilities, or the state itself.
Backward (replaces torch autograd through the tensordot/permute graph):
    lambda  <- O_eff psi with O_eff = sum_m dy_m O_m (+ probability/state seeds)
    for each pass in reverse: one adjoint kernel walks psi and lambda backwards,
               reducing the bilinears C_ij of every parametrised slot,
    grad    <- 2 Re sum_ij A_ij C_ij per parameter (fp64).
"""

from dataclasses import dataclass, field

import numpy as np
import torch

from . import hip
from .planner import build_plan, compact_register, fold_trailing
from .program import encode_plan

_Z = np.array([[1, 0], [0, -1]], dtype=np.complex128)
ZMAX = 32                      # Z-string masks per reduction launch (csrc/sv_measure.hip kMaxZ)
_I = np.eye(2, dtype=np.complex128)


@dataclass
class MeasSpec:
    """How one measurement is produced on the device."""
    kind: str                  # "zexp", "fexp", "probs", "state"
    shape: tuple               # per-circuit output shape
    zindex: int = -1           # column in the Z-string launch
    zmask: int = 0
    fbits: list = field(default_factory=list)     # factored observable: physical bits
    ffactors: np.ndarray = None                    # [q, 2, 2] complex
    pbits: list = field(default_factory=list)     # probability bits, MSB first (empty: all)
    # more than 4 non-Z factors, all Hermitian: F_i = V_i (a_i I + b_i Z) V_i^dagger, so the value is
    # sum_S c_S <phi| Z_{m_S} |phi> with phi = (x V_i^dagger) psi (basis: [(bit, V_i)], bmasks / bcoefs)
    basis: list = None
    bmasks: list = None
    bcoefs: list = None


def _observable_factors(n, meas):
    """Per-physical-bit 2x2 factor of an expval observable (reference order of application)."""
    obs = meas.obs if isinstance(meas.obs, list) else [meas.obs]
    factors = {}
    for ob in obs:
        if len(ob.qubits) != 1:
            raise NotImplementedError("multi-qubit observables are not supported by the SV engine")
        bit = n - 1 - ob.qubits[0]
        m = np.asarray(ob.matrix, dtype=np.complex128).reshape(2, 2)
        factors[bit] = m @ factors.get(bit, _I)
    return factors


def _basis_spec(zmask, rest):
    """Product observable with more than 4 non-Z single-qubit factors, each Hermitian: diagonalise every
    factor, F = V (a I + b Z) V^dagger (a, b from its eigenvalues), and expand the product of the
    (a I + b Z) into Z-strings with real coefficients (terms with a zero coefficient dropped; a Pauli
    factor has a = 0 and contributes one term)."""
    basis, terms = [], [(zmask, 1.0)]
    for bit in sorted(rest):
        f = rest[bit]
        if not np.allclose(f, f.conj().T, atol=1e-12):
            raise NotImplementedError("expval over more than 4 non-Z qubits needs Hermitian factors")
        w, v = np.linalg.eigh(f)
        a, b = 0.5 * (w[1] + w[0]), 0.5 * (w[1] - w[0])
        # eigh sorts ascending: put the larger eigenvalue on |0> so that the Z coefficient is b
        v = v[:, ::-1]
        basis.append((bit, v))
        nxt = []
        for m, c in terms:
            if abs(a) > 1e-15:
                nxt.append((m, c * a))
            if abs(b) > 1e-15:
                nxt.append((m | (1 << bit), c * b))
        terms = nxt
    if len(terms) > ZMAX:
        raise NotImplementedError(f"expval expands into {len(terms)} Z-strings (at most {ZMAX})")
    return MeasSpec("fexp", (), zmask=zmask, basis=basis, bmasks=[m for m, _ in terms],
                    bcoefs=[c for _, c in terms])


def compile_measurements(n, measurements):
    from ..measurement import Expectation, Probability, State  # pylint: disable=import-outside-toplevel
    specs, zmasks = [], []
    for meas in measurements:
        if meas.return_type is Expectation:
            factors = _observable_factors(n, meas)
            zmask, rest = 0, {}
            for bit, f in factors.items():
                if np.array_equal(f, _I):
                    continue
                if np.array_equal(f, _Z):
                    zmask |= 1 << bit
                else:
                    rest[bit] = f
            if not rest:
                specs.append(MeasSpec("zexp", (), zindex=len(zmasks), zmask=zmask))
                zmasks.append(zmask)
            elif len(rest) <= 4:
                bits = sorted(rest)
                specs.append(MeasSpec("fexp", (), zmask=zmask, fbits=bits,
                                      ffactors=np.stack([rest[b] for b in bits])))
            else:
                specs.append(_basis_spec(zmask, rest))
        elif meas.return_type is Probability:
            qs = sorted(set(meas.qubits)) if meas.qubits is not None else list(range(n))
            # every qubit in ascending order is the whole distribution (no marginal sum)
            pbits = [] if len(qs) == n else [n - 1 - q for q in qs]
            specs.append(MeasSpec("probs", tuple([2] * len(qs)), pbits=pbits))
        elif meas.return_type is State:
            specs.append(MeasSpec("state", tuple([2] * n)))
        else:
            raise NotImplementedError(f"measurement {meas.return_type} is not supported")
    return specs, zmasks


class _DeviceBuffers:
    def __init__(self, enc, device):
        self.prog = torch.from_numpy(enc.prog).to(device)
        self.slot_recs = torch.from_numpy(enc.slot_recs.reshape(-1)).to(device)
        self.comp_recs = torch.from_numpy(enc.comp_recs.reshape(-1)).to(device)
        self.const_pool = torch.from_numpy(enc.const_pool.reshape(-1)).to(device)
        self.pgslot = torch.from_numpy(enc.pgslot).to(device)
        self.prog_base = self.prog.data_ptr()
        # unknown-gate-code flag of the table builder (set once; never cleared inside a graph), read back
        # after the first table build that is not being captured into a graph
        self.err = torch.zeros(1, dtype=torch.int32, device=device)
        self.err_checked = False

    def pass_ptr(self, word_offset):
        import ctypes  # pylint: disable=import-outside-toplevel
        return ctypes.c_void_p(self.prog_base + 4 * word_offset)


class _Program:
    """One fusion plan (tile bits k, register bits r) with its encoded device program."""

    def __init__(self, n, units, n_params, const_pool, k, r, specialize=False):
        self.n = n
        self.n_params = n_params
        self.plan = build_plan(n, units, n_params, const_pool, k_max=k, r=r, specialize=specialize)
        self.enc = encode_plan(self.plan)
        self._dev = {}
        self._jit = {}          # device key -> JitModule of generated pass kernels
        # From |0..0>, before pass p the state vanishes wherever a bit that no earlier pass held in its tile
        # is 1, and pass p keeps the non-tile bits: its tiles with such a bit set hold zeros before and after
        # it (forward), and in the adjoint sweep psi vanishes there and no earlier pass reads lambda there.
        # zero_masks[p] (forward order) are those non-tile bits; the first pass runs one tile per circuit.
        full = (1 << n) - 1
        seen = 0
        self.zero_masks = []
        for mask in self.enc.tile_masks:
            self.zero_masks.append(full & ~mask & ~seen if mask != full else 0)
            seen |= mask
        # gradient slots each pass writes (a contiguous range, assigned in pass order)
        self.pass_gslots = []
        for p in self.plan.passes:
            gs = {self.plan.slots[op[4]].gslot for st in p.stages for op in st.ops
                  if op[4] >= 0 and self.plan.slots[op[4]].gslot >= 0}
            if gs and max(gs) - min(gs) + 1 != len(gs):
                raise RuntimeError("gradient slots of a pass are not contiguous")
            self.pass_gslots.append((min(gs), max(gs) + 1) if gs else (0, 0))

    def enable_jit(self, device, forward=True, adjoint=True, seed_masks=None):
        """Compile per-pass kernels for this program on ``device`` (hiprtc); returns True if any are in use.

        Forward and adjoint kernels are compiled separately, once each; modules stay loaded for the
        life of the program (kernels may still be running when a program is dropped).  ``seed_masks``:
        the Z-string masks of an executor whose outputs are all Z-string expectations; the adjoint
        module then also holds a first pass that forms lambda from psi (jit.pass_kernel_source).
        """
        from . import jit  # pylint: disable=import-outside-toplevel
        key = (device.type, device.index)
        k, r = self.plan.k, self.plan.r
        have = self._jit.setdefault(key, {"f": None, "a": None, "a_seed": None, "mods": []})
        for kind, want, adj in (("f", forward, False), ("a", adjoint, True)):
            if not want or have[kind] is not None or not jit.supported(self.n, k, r, adj):
                continue
            offs = self.enc.adj_offsets if adj else self.enc.fwd_offsets
            names = [f"{kind}{i}" for i in range(len(offs))]
            masks = seed_masks if adj and offs else None
            all_names = names + (["a0s"] if masks is not None else [])
            n_f, n_a = (0, len(offs)) if adj else (len(offs), 0)
            persistent = jit.persistent_pays(k, adj)
            with torch.cuda.device(device):
                mod = jit.JitModule(jit.program_source(self.enc, n_f, n_a, masks, persistent=persistent), all_names)
                fns = dict(mod.functions)
                plain = set()
                if persistent:
                    # a kernel that spills with the prefetch registers runs in the one-tile-per-workgroup form
                    plain = {nm for nm in all_names if jit.scratch_bytes(fns[nm]) > 0}
                    if plain:
                        mod2 = jit.JitModule(jit.program_source(self.enc, n_f, n_a, masks, only=plain), sorted(plain))
                        have["mods"].append(mod2)
                        fns.update(mod2.functions)
            have["mods"].append(mod)
            have[kind] = [fns[nm] for nm in names]
            # launch flag per generated kernel: kPassPersistent for those that loop over tiles
            have.setdefault("flags", {}).update({fns[nm].value: (8 if persistent and nm not in plain else 0)
                                                 for nm in all_names})
            if masks is not None:
                have["a_seed"] = (tuple(masks), fns["a0s"])
        return have["f"] is not None or have["a"] is not None

    def seed_fn(self, device, masks):
        """The generated first adjoint pass that forms lambda from psi for these Z-string masks, or None."""
        got = self._jit.get((device.type, device.index))
        if got is None or got["a_seed"] is None or got["a_seed"][0] != tuple(masks):
            return None
        return got["a_seed"][1]

    def _jit_fns(self, device, adjoint):
        got = self._jit.get((device.type, device.index))
        if got is None:
            return None
        return got["a"] if adjoint else got["f"]

    def buffers(self, device):
        key = (device.type, device.index)
        if key not in self._dev:
            self._dev[key] = _DeviceBuffers(self.enc, device)
        return self._dev[key]

    def threads_log2(self):
        k, r = self.plan.k, self.plan.r
        tT = k - r
        if k < self.n:
            return tT
        return max(tT, min(8, max(6, tT)))

    def partial_rows(self):
        k, r = self.plan.k, self.plan.r
        tpt = 1 << (k - r)
        w = max(1, tpt // 64)
        return (1 << (self.n - k)) * w

    def tables(self, params, need_grad):
        """Slot matrices (and A matrices) for [B, P] rows."""
        dev = params.device
        buf = self.buffers(dev)
        B = params.shape[0]
        mats = torch.empty((B, self.plan.mat_stride, 2), dtype=torch.float32, device=dev)
        # every trainable column's A matrix is written by the table kernel; other
        # columns are never read (their gradient slot is -1)
        amats = torch.empty((B, max(self.n_params, 1), 4, 2), dtype=torch.float64, device=dev) if need_grad else None
        params = params.contiguous()
        hip.call("tedq_sv_build_tables", hip.ptr(params) if self.n_params else None,
                 int(params.dtype == torch.float64), B, self.n_params, hip.ptr(buf.slot_recs),
                 len(self.plan.slots), hip.ptr(buf.comp_recs), hip.ptr(buf.const_pool), hip.ptr(mats),
                 self.plan.mat_stride, hip.ptr(amats), hip.ptr(buf.err), hip.stream_handle(dev))
        if not buf.err_checked and not torch.cuda.is_current_stream_capturing():
            check_device_flag(buf.err, "tedq_sv_build_tables: a slot record holds an unknown gate code")
            buf.err_checked = True
        return mats, amats

    def forward_passes(self, psi, mats, probe=None, from_zero=False, outs=None):
        """Apply the passes to psi in place.  ``from_zero``: the sweep starts from |0..0>, so each pass skips
        the tiles that hold zeros (zero_masks); generated kernels also take the amplitudes at tile bits no
        earlier pass held as zero and the first pass builds |0..0> itself (kPassFromZero), so psi then
        needs no initialisation (the interpreted kernels read it: psi must hold |0..0>).  ``outs``
        (generated kernels only): one state buffer per pass after the first; pass i >= 1 reads the
        previous output and writes outs[i] (out of place), so psi and outs[1:] end up holding the state
        after every pass (checkpoints)."""
        dev = psi.device
        buf = self.buffers(dev)
        B = psi.shape[0]
        st = hip.stream_handle(dev)
        tl = self.threads_log2()
        fns = self._jit_fns(dev, False)
        if outs is not None and not fns:
            raise RuntimeError("checkpointed forward passes need the generated forward kernels")
        for i, (off, mask) in enumerate(zip(self.enc.fwd_offsets, self.enc.tile_masks)):
            zero = self.zero_masks[i] if from_zero else 0
            if probe:
                probe.begin("sv_pass_fwd")
            if fns:
                src, dst, flags = psi, None, 4 if from_zero else 0
                if outs is not None and i > 0:
                    src, dst, flags = outs[i - 1], outs[i], flags | 1
                hip.call("tedq_jit_pass", fns[i], 0, hip.ptr(src), hip.ptr(dst), B, self.n, self.plan.k, self.plan.r,
                         mask, zero, hip.ptr(mats), self.plan.mat_stride, None, 0, 0, None, flags, st)
            else:
                hip.call("tedq_sv_apply_pass", hip.ptr(psi), B, self.n, self.plan.k, self.plan.r, tl,
                         buf.pass_ptr(off), mask, zero, hip.ptr(mats), self.plan.mat_stride, st)
            if probe:
                probe.end("sv_pass_fwd")

    def adjoint_passes(self, psi, lam, mats, partials, probe=None, from_zero=False, seed=None, ckpt=None):
        """Reverse sweep over psi and lam (both in place), gradient partial rows into ``partials``.
        ``from_zero``: the forward started from |0..0>, so each pass skips the tiles where psi vanishes;
        the partial rows of skipped tiles are set to zero here.  ``seed`` = (kernel, dy [B, M] float64):
        the first pass forms lambda from psi with that generated kernel (lam is only written).
        ``ckpt``: the forward state after every pass of this same partition (forward_passes' outputs);
        each pass then reads psi from the checkpoint at its end and leaves it unwritten."""
        dev = psi.device
        buf = self.buffers(dev)
        B = psi.shape[0]
        st = hip.stream_handle(dev)
        tl = self.threads_log2()
        rows = self.partial_rows()
        G = self.plan.n_gslots
        fns = self._jit_fns(dev, True)
        if (seed is not None or ckpt is not None) and not fns:
            raise RuntimeError("a fused adjoint seed or checkpoints need the generated adjoint kernels")
        # generated adjoint kernels that loop over tiles are launched with the resident grid only
        kflags = self._jit[(dev.type, dev.index)].get("flags", {}) if fns else {}
        P = len(self.plan.passes)
        if from_zero:
            per_tile = rows >> (self.n - self.plan.k)
            for zero, (g0, g1) in zip(self.zero_masks, self.pass_gslots):
                if zero and g1 > g0:
                    live = per_tile << (self.n - self.plan.k - bin(zero).count("1"))
                    partials[:, g0:g1, live:, :].zero_()
        for i, (off, mask) in enumerate(zip(self.enc.adj_offsets, reversed(self.enc.tile_masks))):
            zero = self.zero_masks[P - 1 - i] if from_zero else 0
            if probe:
                probe.begin("sv_pass_adj")
            if fns:
                fn, dy = (seed[0], hip.ptr(seed[1])) if seed is not None and i == 0 else (fns[i], None)
                src, flags = (ckpt[P - 1 - i], 2) if ckpt is not None else (psi, 0)
                flags |= kflags.get(fn.value, 0)
                hip.call("tedq_jit_pass", fn, 1, hip.ptr(src), hip.ptr(lam), B, self.n, self.plan.k, self.plan.r,
                         mask, zero, hip.ptr(mats), self.plan.mat_stride, hip.ptr(partials), rows, G, dy, flags,
                         st)
            else:
                hip.call("tedq_sv_adjoint_pass", hip.ptr(psi), hip.ptr(lam), B, self.n, self.plan.k, self.plan.r,
                         tl, buf.pass_ptr(off), mask, zero, hip.ptr(mats), self.plan.mat_stride,
                         hip.ptr(partials), rows, G, st)
            if probe:
                probe.end("sv_pass_adj")


class SVExecutor:
    """Compiled SV program for one circuit structure.

    The forward and the adjoint sweep may use different tile / register
    shapes (``k``, ``r`` and ``k_bwd``, ``r_bwd``), each with its own fusion plan.
    """

    def __init__(self, n, units, n_params, const_pool, measurements, init_state=None, k_max=12, r=4,
                 k_bwd=None, r_bwd=None, fold=True):
        self.n = n
        self.n_params = n_params
        self._pool = list(const_pool)
        self.specs, self.zmasks = compile_measurements(n, measurements)
        # when every measurement is a Z-string expectation, trailing gates that map Z-strings to Z-strings
        # are folded into the observables (planner.fold_trailing) and, from |0..0>, qubits that no gate
        # is left on are dropped (planner.compact_register): psi is then the state of the n kept qubits
        # before the folded gates; n_circuit stays the circuit's width
        self.n_circuit = n
        self.folded = 0
        if fold and self.specs and all(s.kind == "zexp" for s in self.specs):
            units, self.zmasks, self.folded = fold_trailing(units, self.zmasks)
            if init_state is None:
                units, self.zmasks, self.n = compact_register(units, self.zmasks, n)
            for s in self.specs:
                s.zmask = self.zmasks[s.zindex]
        n = self.n
        self._units = units
        # expectation values and probabilities only (no state returned, none taken after a measurement):
        # the forward may drop global phases (see enable_jit)
        self.phase_free = (bool(self.specs) and all(s.kind in ("zexp", "fexp", "probs") for s in self.specs)
                           and not any(getattr(m, "after_state", False) for m in measurements))
        # every forward measurement also reduces <psi|psi> in fp64 (the Z-string of mask 0, compensated
        # sums), appended after the circuit's own masks; the last batch's values stay in last_norm
        self._zall = list(self.zmasks) + [0]
        self.last_norm = None
        # when set, every measurement checks max |<psi|psi> - 1| <= norm_tol (one read-back per call)
        self.norm_tol = None
        self.init_state = None
        if init_state is not None:
            flat = np.asarray(init_state, dtype=np.complex64).reshape(-1)
            if flat.size != 2 ** n:
                raise ValueError(f"initial state has {flat.size} amplitudes, expected {2 ** n}")
            self.init_state = flat
        self._dev = {}
        # optional timing probe with begin(kind)/end(kind) around every pass launch
        self.probe = None
        # device bytes the forward may spend on per-pass checkpoints for the backward (None: 40% of the
        # device's memory; 0: never)
        self.checkpoint_bytes = None
        # compile per-pass kernels on first use on a device (see enable_jit)
        self.jit = False
        self._jitted = set()
        self.configure(k_max, r, k_bwd, r_bwd)

    def enable_jit(self, device=None, forward=True, adjoint=True):
        """Compile the forward and adjoint programs into per-pass kernels (hiprtc) for ``device``.

        The adjoint switches to its specialised plan (unfused rotations with single-value gradients,
        real 2x2 ops, diagonal gates without their global phase) when generated kernels can run it:
        that plan has about twice the ops, which only generated code executes without per-op overhead.
        The forward does the same when every measurement is an expectation value or a probability
        (which no global phase changes; a returned state or states after a measurement keep the exact
        forward).  Returns
        (forward_jitted, adjoint_jitted).  Compiled programs are kept per (k, r) shape.
        """
        from . import jit  # pylint: disable=import-outside-toplevel
        device = device or torch.device("cuda", torch.cuda.current_device())
        if forward:
            k, r = self.fwd.plan.k, self.fwd.plan.r
            if self.phase_free and jit.supported(self.n, k, r, False):
                self.fwd = self._program(k, r, specialize=True)
            self.fwd.enable_jit(device, True, False)
        if adjoint:
            k, r = self.bwd.plan.k, self.bwd.plan.r
            if jit.supported(self.n, k, r, True):
                self.bwd = self._program(k, r, specialize=True)
            self.bwd.enable_jit(device, False, True, self._seed_masks())
        # pylint: disable=protected-access
        return self.fwd._jit_fns(device, False) is not None, self.bwd._jit_fns(device, True) is not None

    def _seed_masks(self):
        """Z-string masks whose adjoint seed a generated first pass may form (every output a Z-string
        expectation, at most ZMAX of them), else None."""
        if self.specs and all(s.kind == "zexp" for s in self.specs) and 0 < len(self.zmasks) <= ZMAX:
            return tuple(self.zmasks)
        return None

    def _ensure_jit(self, device):
        key = (device.type, device.index)
        if self.jit and key not in self._jitted:
            self.enable_jit(device)
            self._jitted.add(key)

    def _program(self, k, r, specialize=False):
        """Fusion plan for (k, r), built once per executor (with its compiled kernels, if any).  The
        specialised plan keeps trainable rotations unfused (single-value gradients, real 2x2 ops) and
        applies diagonal gates without their global phase, for generated kernels."""
        if not hasattr(self, "_programs"):
            self._programs = {}
        key = (k, r, specialize)
        if key not in self._programs:
            self._programs[key] = _Program(self.n, self._units, self.n_params, list(self._pool), k, r,
                                           specialize=specialize)
        return self._programs[key]

    def configure(self, k, r, k_bwd=None, r_bwd=None):
        """Select the forward and adjoint programs for the given tile / register shapes."""
        self._jitted = set()
        k_bwd = k if k_bwd is None else k_bwd
        # the adjoint holds psi and lambda registers: more than 16 amplitudes each would spill
        r_bwd = min(4, r if r_bwd is None else r_bwd)
        self.fwd = self._program(k, r)
        self.bwd = self._program(k_bwd, r_bwd)
        if self.jit:
            self._ensure_jit(torch.device("cuda", torch.cuda.current_device()))

    # ------------------------------------------------------------------
    @property
    def plan(self):
        return self.fwd.plan

    @property
    def enc(self):
        return self.fwd.enc

    @property
    def n_passes(self):
        return len(self.fwd.plan.passes)

    def partial_rows(self):
        return self.bwd.partial_rows()

    def _zmask_buffer(self, device):
        key = (device.type, device.index)
        if key not in self._dev:
            z = torch.tensor(np.asarray(self._zall, dtype=np.uint64).view(np.int64), device=device)
            init = None
            if self.init_state is not None:
                init = torch.from_numpy(self.init_state.view(np.float32).copy()).to(device)
            self._dev[key] = (z, init)
        return self._dev[key]

    # ------------------------------------------------------------------
    def _checkpoints_fit(self, dev, B, from_zero):
        """Whether a forward that a backward follows keeps the state after every pass: generated kernels
        for both directions, one pass partition for both (the same tile bits k), every tile written by the
        last pass, and the extra states within ``checkpoint_bytes`` (None: 40% of the device's memory)."""
        if self.fwd._jit_fns(dev, False) is None or self.bwd._jit_fns(dev, True) is None:  # pylint: disable=protected-access
            return False
        if not self.checkpoints_compatible():
            return False
        if from_zero and self.fwd.zero_masks[-1]:
            return False
        budget = self.checkpoint_bytes
        if budget is None:
            budget = int(0.4 * torch.cuda.get_device_properties(dev).total_memory)
        return (len(self.fwd.plan.passes) - 1) * B * (8 << self.n) <= budget

    def checkpoints_compatible(self):
        """Whether the forward's state after each pass is the state the adjoint walks back to: at least two
        passes, one pass partition, and one phase convention.  A specialised plan applies diagonal gates
        without their global phase; the in-place sweep does not care (psi and lambda carry the same
        phases), but an exact forward's checkpoints would meet a lambda whose phases differ."""
        f, b = self.fwd, self.bwd
        return (len(f.plan.passes) >= 2 and f.enc.tile_masks == b.enc.tile_masks
                and f.plan.specialize == b.plan.specialize)

    def run_forward(self, params, need_grad=False):
        """Apply the circuit to every parameter row.  Returns psi [B, 2^n] complex64.

        With ``need_grad`` (a backward follows) and when they fit (_checkpoints_fit), the passes write
        out of place and the states after every pass stay attached to the returned psi as checkpoints:
        the adjoint sweep then reads psi from them instead of writing it back (a quarter less traffic)."""
        dev = params.device
        B = params.shape[0]
        self._ensure_jit(dev)
        mats, _ = self.fwd.tables(params, False)
        psi = torch.empty((B, 2 ** self.n), dtype=torch.complex64, device=dev)
        st = hip.stream_handle(dev)
        _, init = self._zmask_buffer(dev)
        if init is not None:
            hip.call("tedq_sv_init_copy", hip.ptr(psi), hip.ptr(init), B, self.n, st)
        elif self.fwd._jit_fns(dev, False) is None or self.fwd.zero_masks[-1]:  # pylint: disable=protected-access
            # interpreted kernels read |0..0>; with generated ones the last pass writes every tile unless it
            # skips some, and what it skips is read later (measurements) as zeros
            hip.call("tedq_sv_init_zero", hip.ptr(psi), B, self.n, st)
        outs = None
        if need_grad and self._checkpoints_fit(dev, B, init is None):
            outs = [psi] + [torch.empty_like(psi) for _ in range(len(self.fwd.plan.passes) - 1)]
        self.fwd.forward_passes(psi, mats, self.probe, from_zero=init is None, outs=outs)
        if outs is None:
            return psi
        final = outs[-1]
        # the adjoint program they serve and the states before the last pass's output (which is psi itself:
        # holding it here would make a reference cycle that keeps every state alive until a collection)
        final.tedq_checkpoints = (self.bwd, outs[:-1])
        return final

    def measure(self, psi):
        """List of per-measurement outputs, each [B, *shape] (expvals fp64, probs fp32, state c64)."""
        dev = psi.device
        zbuf, _ = self._zmask_buffer(dev)
        B = psi.shape[0]
        st = hip.stream_handle(dev)
        nchunk = hip.load().tedq_sv_reduce_chunks(self.n)
        # all Z-string observables and the norm, at most ZMAX masks per reduction pass
        M = len(self._zall)
        zout = torch.empty((B, M), dtype=torch.float64, device=dev)
        partial = torch.empty((B, nchunk, min(M, ZMAX)), dtype=torch.float64, device=dev)
        for c0 in range(0, M, ZMAX):
            m = min(ZMAX, M - c0)
            part = zout if m == M else torch.empty((B, m), dtype=torch.float64, device=dev)
            hip.call("tedq_sv_expval_z", hip.ptr(psi), B, self.n, _offset_ptr(zbuf, c0), m, hip.ptr(partial),
                     hip.ptr(part), st)
            if m != M:
                zout[:, c0:c0 + m] = part
        self.last_norm = zout[:, M - 1]
        if self.norm_tol is not None and not torch.cuda.is_current_stream_capturing():
            dev_max = norm_deviation(self)
            if not dev_max <= self.norm_tol:
                raise hip.EngineError(f"state norm check failed: max |<psi|psi> - 1| = {dev_max:.3e} over "
                                      f"{B} rows exceeds {self.norm_tol:.1e}")
        outs = []
        for spec in self.specs:
            if spec.kind == "zexp":
                outs.append(zout[:, spec.zindex])
            elif spec.kind == "fexp" and spec.basis is not None:
                phi = self._basis_apply(spec, psi, inverse=True)
                M = len(spec.bmasks)
                masks = self._basis_masks(spec, dev)
                part = torch.empty((B, nchunk, M), dtype=torch.float64, device=dev)
                vals = torch.empty((B, M), dtype=torch.float64, device=dev)
                hip.call("tedq_sv_expval_z", hip.ptr(phi), B, self.n, hip.ptr(masks), M, hip.ptr(part),
                         hip.ptr(vals), st)
                outs.append(vals @ torch.tensor(spec.bcoefs, dtype=torch.float64, device=dev))
            elif spec.kind == "fexp":
                bits, fac = _factored_args(spec, dev)
                partial = torch.empty((B, nchunk), dtype=torch.float64, device=dev)
                out = torch.empty((B,), dtype=torch.float64, device=dev)
                hip.call("tedq_sv_expval_factored", hip.ptr(psi), B, self.n, len(spec.fbits), _hptr(bits),
                         _hptr(fac), spec.zmask, hip.ptr(partial), hip.ptr(out), st)
                outs.append(out)
            elif spec.kind == "probs":
                k = len(spec.pbits) if spec.pbits else self.n
                out = torch.empty((B, 2 ** k), dtype=torch.float32, device=dev)
                if spec.pbits:
                    bits = np.ascontiguousarray(spec.pbits, dtype=np.int32)
                    rows = hip.load().tedq_sv_probs_chunks(self.n, k)
                    partial = torch.empty((B, rows, 2 ** k), dtype=torch.float64, device=dev)
                    hip.call("tedq_sv_probs", hip.ptr(psi), B, self.n, k, _hptr(bits), hip.ptr(partial),
                             hip.ptr(out), st)
                else:
                    hip.call("tedq_sv_probs", hip.ptr(psi), B, self.n, k, None, None, hip.ptr(out), st)
                outs.append(out.view((B,) + spec.shape))
            else:
                outs.append(psi.view((B,) + spec.shape))
        return outs

    def _basis_program(self, spec, inverse):
        """Pass program applying V_i^dagger (inverse) or V_i to every factor bit of a basis spec."""
        from .planner import Unit, _const_comp  # pylint: disable=import-outside-toplevel
        if not hasattr(self, "_basis_progs"):
            self._basis_progs = {}
        key = (id(spec), inverse)
        if key not in self._basis_progs:
            pool = []
            units = [Unit("u1", (bit,), _const_comp(v.conj().T if inverse else v, pool)) for bit, v in spec.basis]
            self._basis_progs[key] = _Program(self.n, units, 0, pool, self.fwd.plan.k, self.fwd.plan.r)
        return self._basis_progs[key]

    def _basis_masks(self, spec, device):
        key = ("bmask", id(spec), device.type, device.index)
        if key not in self._dev:
            self._dev[key] = torch.tensor(np.asarray(spec.bmasks, dtype=np.uint64).view(np.int64), device=device)
        return self._dev[key]

    def _basis_apply(self, spec, psi, inverse, in_place=False):
        """(x V^dagger) psi (inverse) or (x V) psi over the factor bits, on a copy unless in_place."""
        prog = self._basis_program(spec, inverse)
        out = psi if in_place else psi.clone()
        B = psi.shape[0]
        mats, _ = prog.tables(torch.empty((B, 0), dtype=torch.float32, device=psi.device), False)
        prog.forward_passes(out, mats)
        return out

    def post_measurement_states(self, psi_row, qubits):
        """Normalised post-measurement states of one circuit for every outcome of ``qubits`` (caller's
        order): [2^k, 2^(n-k)] complex64, outcome y in the bit order of ``qubits``, the remaining qubits
        in ascending order (the reference's after_state, pytorch_backend.py:474-493)."""
        dev = psi_row.device
        k = len(qubits)
        bits = np.ascontiguousarray([self.n - 1 - q for q in qubits], dtype=np.int32)
        st = hip.stream_handle(dev)
        rows = hip.load().tedq_sv_probs_chunks(self.n, k)
        partial = torch.empty((1, rows, 2 ** k), dtype=torch.float64, device=dev)
        probs = torch.empty((1, 2 ** k), dtype=torch.float32, device=dev)
        hip.call("tedq_sv_probs", hip.ptr(psi_row), 1, self.n, k, _hptr(bits), hip.ptr(partial), hip.ptr(probs), st)
        out = torch.empty((2 ** k, 2 ** (self.n - k)), dtype=torch.complex64, device=dev)
        hip.call("tedq_sv_post_measure", hip.ptr(psi_row), self.n, k, _hptr(bits), hip.ptr(probs), hip.ptr(out), st)
        return out

    # ------------------------------------------------------------------
    def _zexp_columns(self, grads, B, dev):
        """Upstream gradient of every Z-string mask, [B] float64 each (summed over the outputs sharing a
        mask, zeros where none)."""
        cols = [None] * len(self.zmasks)
        for spec, g in zip(self.specs, grads):
            if spec.kind == "zexp" and g is not None:
                g = g.to(torch.float64).reshape(B)
                cols[spec.zindex] = g if cols[spec.zindex] is None else cols[spec.zindex] + g
        zero = None
        for i, c in enumerate(cols):
            if c is None:
                if zero is None:
                    zero = torch.full((B,), 0.0, dtype=torch.float64, device=dev)
                cols[i] = zero
        return cols

    def adjoint(self, params, psi, grads, grad_dtype=torch.float32):
        """Seed lambda from the upstream gradients (aligned with self.specs) and run the adjoint sweep;
        psi is consumed.  Returns grad [B, P].  When every output is a Z-string expectation and generated
        kernels run the sweep, its first pass forms lambda from psi itself (no separate seed kernel)."""
        dev = psi.device
        if not self.n_params:
            return torch.empty((psi.shape[0], 0), dtype=grad_dtype, device=dev)
        self._ensure_jit(dev)
        masks = self._seed_masks()
        fn = self.bwd.seed_fn(dev, masks) if masks is not None else None
        if fn is not None:
            dy = torch.stack(self._zexp_columns(grads, psi.shape[0], dev), 1).contiguous()
            return self.run_backward(params, psi, torch.empty_like(psi), grad_dtype, seed=(fn, dy))
        return self.run_backward(params, psi, self.seed_adjoint(psi, grads), grad_dtype)

    def seed_adjoint(self, psi, grads):
        """lambda = sum_m (d out_m)^* seed_m applied to psi; grads aligned with self.specs (None = zero)."""
        dev = psi.device
        zbuf, _ = self._zmask_buffer(dev)
        B = psi.shape[0]
        st = hip.stream_handle(dev)
        lam = torch.empty_like(psi)
        M = len(self.zmasks)
        dy = None
        cols = []
        if M:
            cols = self._zexp_columns(grads, B, dev)
            dy = torch.stack(cols[:ZMAX], 1).contiguous()
        m0 = min(M, ZMAX)
        hip.call("tedq_sv_lambda_z", hip.ptr(psi), hip.ptr(lam), B, self.n, hip.ptr(zbuf) if m0 else None, m0,
                 hip.ptr(dy), st)
        # observables past the first ZMAX: lam += dy_m Z-string_m psi, one accumulating pass each
        for m in range(ZMAX, M):
            coef = cols[m].contiguous()
            hip.call("tedq_sv_lambda_factored", hip.ptr(psi), hip.ptr(lam), B, self.n, 0, None, None,
                     self.zmasks[m], hip.ptr(coef), st)
        for spec, g in zip(self.specs, grads):
            if g is None or spec.kind == "zexp":
                continue
            if spec.kind == "fexp" and spec.basis is not None:
                # lam += g O psi with O = (x V) [sum_S c_S Z_S] (x V^dagger)
                phi = self._basis_apply(spec, psi, inverse=True)
                coef = g.to(torch.float64).reshape(B, 1) * torch.tensor(spec.bcoefs, dtype=torch.float64,
                                                                        device=dev)
                mu = torch.empty_like(psi)
                hip.call("tedq_sv_lambda_z", hip.ptr(phi), hip.ptr(mu), B, self.n, hip.ptr(self._basis_masks(spec, dev)),
                         len(spec.bmasks), hip.ptr(coef.contiguous()), st)
                mu = self._basis_apply(spec, mu, inverse=False, in_place=True)
                hip.call("tedq_sv_axpy", hip.ptr(lam), hip.ptr(mu), B, self.n, 1.0, st)
            elif spec.kind == "fexp":
                bits, fac = _factored_args(spec, dev)
                coef = g.to(torch.float64).reshape(B).contiguous()
                hip.call("tedq_sv_lambda_factored", hip.ptr(psi), hip.ptr(lam), B, self.n, len(spec.fbits),
                         _hptr(bits), _hptr(fac), spec.zmask, hip.ptr(coef), st)
            elif spec.kind == "probs":
                k = len(spec.pbits) if spec.pbits else self.n
                gg = g.to(torch.float32).reshape(B, 2 ** k).contiguous()
                bits = np.ascontiguousarray(spec.pbits, dtype=np.int32) if spec.pbits else None
                hip.call("tedq_sv_lambda_probs", hip.ptr(psi), hip.ptr(lam), B, self.n, k, _hptr(bits),
                         hip.ptr(gg), st)
            elif spec.kind == "state":
                gs = g.to(torch.complex64).reshape(B, 2 ** self.n).contiguous()
                hip.call("tedq_sv_axpy", hip.ptr(lam), hip.ptr(gs), B, self.n, 0.5, st)
        return lam

    def run_backward(self, params, psi, lam, grad_dtype=torch.float32, seed=None):
        """Adjoint sweep for [B, P] rows; consumes psi and lam (both are overwritten).  Returns grad [B, P].
        ``seed``: (kernel, dy) for a first pass that forms lambda from psi (see ``adjoint``)."""
        dev = psi.device
        B = psi.shape[0]
        grad = torch.empty((B, max(self.n_params, 1)), dtype=grad_dtype, device=dev)
        if self.n_params == 0:
            return grad[:, :0]
        self._ensure_jit(dev)
        prog = self.bwd
        mats, amats = prog.tables(params, True)
        G = prog.plan.n_gslots
        rows = prog.partial_rows()
        partials = torch.empty((B, max(G, 1), rows, 8), dtype=torch.float32, device=dev)
        ck = getattr(psi, "tedq_checkpoints", None)
        ckpt = list(ck[1]) + [psi] if ck is not None and ck[0] is prog else None
        prog.adjoint_passes(psi, lam, mats, partials, self.probe, from_zero=self.init_state is None, seed=seed,
                            ckpt=ckpt)
        C = torch.empty((B, max(G, 1), 8), dtype=torch.float64, device=dev)
        buf = prog.buffers(dev)
        hip.call("tedq_sv_grad_reduce", hip.ptr(partials), B, G, rows, hip.ptr(amats), hip.ptr(buf.pgslot),
                 self.n_params, hip.ptr(C), hip.ptr(grad), int(grad_dtype == torch.float64),
                 hip.stream_handle(dev))
        return grad

    def value_and_grad(self, rows, grads_out):
        """Measurements and parameter gradient for [B, P] rows in one launch sequence, without autograd.

        ``grads_out`` holds one upstream gradient per measurement (None = 0).
        Returns (list of measurement outputs, grad [B, P]).
        """
        psi = self.run_forward(rows, need_grad=True)
        outs = self.measure(psi)
        grad = self.adjoint(rows, psi, grads_out, grad_dtype=rows.dtype)
        return outs, grad

    def memory_bytes(self, batch, with_grad=True):
        """Device bytes one launch sequence of ``batch`` rows needs (states, tables, partials)."""
        amps = batch * (2 ** self.n) * 8
        tables = batch * (self.fwd.plan.mat_stride + self.bwd.plan.mat_stride) * 8 + batch * max(self.n_params, 1) * 64
        if not with_grad:
            return amps + tables
        partials = batch * max(self.bwd.plan.n_gslots, 1) * self.bwd.partial_rows() * 32
        return 2 * amps + tables + partials


def norm_deviation(ex):
    """Largest |<psi|psi> - 1| over the rows of an executor's last forward measurement (fp64, reduced on
    the device with compensated sums); None before the first measurement.  Reads the values back."""
    if ex.last_norm is None:
        return None
    return float((ex.last_norm - 1.0).abs().max())


def check_device_flag(flag, what):
    """Raise EngineError if a device error flag (int32, written by a table kernel) is nonzero.

    Reads one int back, so it synchronises the stream once; callers do it on the first uncaptured
    build of a program only."""
    code = int(flag.item())
    if code:
        raise hip.EngineError(f"{what} (device error flag {code})")


def _factored_args(spec, device):  # pylint: disable=unused-argument
    """Host arrays (the library copies them into the kernel arguments)."""
    bits = np.ascontiguousarray(spec.fbits, dtype=np.int32)
    fac = np.ascontiguousarray(spec.ffactors.astype(np.complex64)).view(np.float32).reshape(-1)
    return bits, fac


def _offset_ptr(t, elems):
    """Device pointer ``elems`` elements into a contiguous tensor."""
    import ctypes  # pylint: disable=import-outside-toplevel
    return ctypes.c_void_p(t.data_ptr() + elems * t.element_size())


def _hptr(a):
    return None if a is None else a.ctypes.data_as(hip.c_vp)
