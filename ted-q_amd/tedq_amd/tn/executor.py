"""Tensor-network execution of compiled circuits on the GPU.

Mirrors the reference's contraction mode (tedq/backends/pytorch_backend.py:
242-355): one bra/ket network per measurement (``gen_tensor_networks``), the
zero vectors, gate operands, observable and conjugated gate operands as
leaves, contracted pairwise along a path, then ``squeeze(result.real)`` for
expectation values and probabilities.

What is this module's own:

* leaves are built on the device for every parameter row at once (one
  launch: tedq_tn_gate_tensors), with the derivative blocks of trainable
  parameters;
* every network is first reduced by cancelling ket/bra gate pairs outside
  the observable's light cone (``cancel_inverse_pairs``), then given a
  deterministic path (``search_path``); path and index layouts are frozen at
  construction;
* each pairwise step is one tedq_tn_contract launch over the whole parameter
  batch, reading operands and writing the result through index-bit layouts
  (no permuted copies);
* gradients come from the reverse contraction tree: the environment of a
  step's result gives the environments of its two operands (two more
  contraction launches), and the environments of the gate leaves give
  d value / d theta (tedq_tn_param_grad).  Intermediates are recomputed per
  network in the backward pass, and when they do not fit the memory cap the
  path is cut into segments whose boundary tensors are checkpointed and whose
  insides are recomputed.
"""

import ctypes
import math
from dataclasses import dataclass

import numpy as np
import torch

from ..engine import hip
from ..engine.sv import MeasSpec, check_device_flag
from ..measurement import Expectation, Probability, State
from .network import cancel_inverse_pairs, diagonal_reduce
from .path import external_path, path_sets, path_stats, search_path
from .slicing import choose_slices, slice_network

_BLOCK_CODE = {"RX": 14, "RY": 15, "RZ": 16, "Rot": 17, "PhaseShift": 18,
               "CRX": 14, "CRY": 15, "CRZ": 16, "ControlledPhaseShift": 18}
# gates that are diagonal for every parameter value (their tensors become hyper-index diagonals)
_DIAGONAL_GATES = {"RZ", "PhaseShift", "PauliZ", "S", "T", "I", "CZ", "CRZ", "ControlledPhaseShift"}


# controlled gates: their leading (control) wires are diagonal for every parameter value
_CONTROL_WIRES = {"CNOT": 1, "CY": 1, "CRX": 1, "CRY": 1, "Toffoli": 2, "CSWAP": 1}


def _is_diagonal_matrix(m):
    m = np.asarray(m)
    return m.ndim == 2 and m.shape[0] == m.shape[1] and not np.any(m - np.diag(np.diag(m)))


def _leading_diagonal_wires(m, nq):
    """Leading wires (qubit order, wire 0 = most significant bit) in which the matrix is diagonal: its
    entries vanish wherever row and column differ in that wire's bit."""
    m = np.asarray(m).reshape(1 << nq, 1 << nq)
    idx = np.arange(1 << nq)
    for w in range(nq):
        bit = 1 << (nq - 1 - w)
        differ = ((idx[:, None] ^ idx[None, :]) & bit) != 0
        if np.any(m[differ]):
            return w
    return nq


def diagonal_wires(op):
    """Number of leading wires in which the operator's matrix is diagonal whatever its parameters: all of
    them for diagonal gates, the controls of controlled gates (by type), else from a constant matrix."""
    name = op.name[:-4] if getattr(op, "is_inverse", False) else op.name
    nq = len(op.qubits)
    if name in _DIAGONAL_GATES:
        return nq
    if name in _CONTROL_WIRES:
        return _CONTROL_WIRES[name]
    if op.num_params == 0:
        return _leading_diagonal_wires(op.matrix, nq)
    return 0
_NO_SRC = -2 ** 31
_DESC_WORDS = 4 + 9 * 32
_ROLE = {"AH": 0, "AM": 1, "AK": 2, "BH": 3, "BK": 4, "BN": 5, "CH": 6, "CM": 7, "CN": 8}


def _bitpos(lst):
    r = len(lst)
    return {s: r - 1 - i for i, s in enumerate(lst)}


def _gate_step(desc):
    """Steps whose backward runs as one pass (tedq_tn_gate_backward_workspace's rule, tn_contract.hip):
    no kept-shared index, at most two summed indices, one operand gate-sized (at most two new indices)
    and the other at least 2^10 rows."""
    nh, nm, nn, nk = (int(v) for v in desc[:4])
    return nh == 0 and nk <= 2 and ((nn <= 2 and nm >= 10) or (nm <= 2 and nn >= 10))


def gemm_shaped(nm, nn, nk):
    """Steps the library runs on the matrix-core GEMM kernel (tn_contract.hip, plan_launch)."""
    return nm >= 6 and nn >= 6 and nk >= 4


def step_roles(la, lb, lc):
    """Index roles of C = contract(A, B): kept-shared h, A-only m, B-only n, summed k (kernel bit order)."""
    sa, sb, sc = set(la), set(lb), set(lc)
    pa, pc = _bitpos(la), _bitpos(lc)
    return {"h": sorted(sa & sb & sc, key=pc.get), "m": sorted((sa - sb) & sc, key=pc.get),
            "n": sorted((sb - sa) & sc, key=pc.get), "k": sorted((sa & sb) - sc, key=pa.get)}


def gemm_operand_layouts(la, lb, lc):
    """For a GEMM-shaped step, operand index lists with the summed indices as the lowest address bits,
    in the same order in both operands (A's order), the other indices keeping their order, so that the
    kernel's tile loads read contiguous runs along k; None for other steps.  The result does not depend
    on C's layout, so a producer can write an operand in it once and for all."""
    r = step_roles(la, lb, lc)
    if not gemm_shaped(len(r["m"]), len(r["n"]), len(r["k"])):
        return None
    ks = list(reversed(r["k"]))  # role bit 0 is the least significant address bit
    kset = set(ks)
    return [x for x in la if x not in kset] + ks, [x for x in lb if x not in kset] + ks


# GEMM forms of the descriptor's last word (tn_contract.hip, kDescForm): 4 real products on 128 x 128 tiles
# where the free sizes allow (default), the 3-product Gauss form, or 4 products on 64 x 64 tiles only
GEMM_FORMS = {"4m": 0, "3m": 3, "4m64": 64}


def make_desc(la, lb, lc, gemm_form=0):
    """Descriptor of C = contract(A, B) for index lists (most significant first) of A, B and C.

    ``gemm_form`` (a value of GEMM_FORMS) goes to the last word and selects the kernel form of a
    GEMM-shaped step."""
    sa, sb, sc = set(la), set(lb), set(lc)
    h = sa & sb & sc
    m = (sa - sb) & sc
    n = (sb - sa) & sc
    k = (sa & sb) - sc
    if (sa | sb) - sc - k or sc - (sa | sb):
        raise ValueError("contraction step: an index is neither kept nor summed over both operands")
    pa, pb, pc = _bitpos(la), _bitpos(lb), _bitpos(lc)
    roles = {
        "h": sorted(h, key=pc.get), "m": sorted(m, key=pc.get),
        "n": sorted(n, key=pc.get), "k": sorted(k, key=pa.get),
    }
    d = np.zeros(_DESC_WORDS, dtype=np.int32)
    d[0:4] = [len(roles["h"]), len(roles["m"]), len(roles["n"]), len(roles["k"])]

    def put(role_name, syms, pos):
        base = 4 + 32 * _ROLE[role_name]
        for j, s in enumerate(syms):
            d[base + j] = pos[s]

    put("AH", roles["h"], pa)
    put("AM", roles["m"], pa)
    put("AK", roles["k"], pa)
    put("BH", roles["h"], pb)
    put("BK", roles["k"], pb)
    put("BN", roles["n"], pb)
    put("CH", roles["h"], pc)
    put("CM", roles["m"], pc)
    put("CN", roles["n"], pc)
    if gemm_form not in GEMM_FORMS.values():
        raise ValueError(f"unknown GEMM form {gemm_form}")
    d[_DESC_WORDS - 1] = gemm_form
    return d


def _step_layout(la, lb, res):
    """Index order of an intermediate.  When a small tensor (a gate) meets a large one, the result keeps
    the large tensor's layout with each summed index replaced by the matching new index of the small one
    (wire i of a gate's [new..., old...] pairs new[i] with old[i]), so the step reads and writes the large
    tensor at the same addresses; otherwise A's kept indices, then B's."""
    big, small = (la, lb) if len(la) >= len(lb) else (lb, la)
    if len(small) <= 8 and set(res) == set(big):
        return list(big)          # a diagonal (hyper-index) factor: elementwise, in the big tensor's layout
    if len(small) <= 8:
        sbig = set(big)
        summed = [x for x in small if x in sbig and x not in res]
        new = [x for x in small if x in res and x not in sbig]
        if summed and len(summed) == len(new):
            mp = dict(zip(summed, new))
            return [mp.get(x, x) for x in big]
    return [x for x in la if x in res] + [x for x in lb if x in res and x not in la]


@dataclass
class _Step:
    a: int
    b: int
    c: int
    fwd: int            # row of the descriptor table
    bwd_a: int          # descriptor of E_A = contract(E_C, B), -1 if A needs no environment
    bwd_b: int


class _Network:
    """One reduced network with its frozen path and index layouts."""

    def __init__(self, tn, leaf_src, desc_rows, trials, external=None, slicing=None, sliced=None, is_diagonal=None,
                 searched=None):
        # an externally found order applies to the network as built (no light-cone reduction)
        red = tn if external is not None else cancel_inverse_pairs(tn)
        diag = [0] * red.num_tensors
        if is_diagonal is not None and external is None:
            red, diag = diagonal_reduce(red, is_diagonal)
        self.n_diagonal = sum(1 for d in diag if d)
        inputs = [list(t) for t in red.input_indices]
        sources = list(red.sources)
        if len(inputs) < 2:                       # keep at least one step
            inputs.append([])
            sources.append(("one",))
            diag.append(0)
            if external is not None:
                external = [(0, 1)]
        self.measurement = tn.measurement
        self.output = list(tn.output_indices)
        if external is not None:
            path = [tuple(p) for p in external]
            sets = path_sets(inputs, self.output, path)
            largest, _ = path_stats(path, sets, len(inputs))
            info = {"strategy": "external", "trial": None,
                    "macs": sum(1 << len(sets[i] | sets[j]) for i, j in path), "largest_rank": largest}
        elif searched is not None:
            path, sets, info = searched
        else:
            path, sets, info = search_path(inputs, self.output, trials)
        # sliced indices: given with an external tree, or chosen on the frozen path from slicing options
        sliced = list(sliced) if sliced else []
        if not sliced and slicing:
            sliced = choose_slices(path, sets, len(inputs), self.output, slicing.get("target_size"),
                                   slicing.get("target_slices"))
        self.sliced = sliced
        self.sel_off = -1          # first constant slot of the selector vectors (set by the executor)
        if sliced:
            inputs, sources, path = slice_network(inputs, sources, self.output, path, sliced)
            diag += [0] * (len(inputs) - len(diag))
            sets = path_sets(inputs, self.output, path)
            largest, _ = path_stats(path, sets, len(inputs))
            info = dict(info, sliced=list(sliced), slices=1 << len(sliced), largest_rank=largest,
                        macs=sum(1 << len(sets[i] | sets[j]) for i, j in path) << len(sliced))
        self.n_inputs = len(inputs)
        self.sources = sources
        self.leaves = [leaf_src(s, d) for s, d in zip(sources, diag)]
        self.path_info = info
        self.path = path
        lay = [list(t) for t in inputs]
        # which leaves and intermediates need an environment (gate leaves and everything above them)
        needs = [s[0] in ("gate", "adj") for s in sources]
        self.steps = []
        for s, (i, j) in enumerate(path):
            res = sets[self.n_inputs + s]
            if s == len(path) - 1:
                out = list(self.output)
                if set(out) != set(res):
                    raise ValueError("tensor network: final tensor does not carry the output indices")
            else:
                out = _step_layout(lay[i], lay[j], res)
            lay.append(out)
            needs.append(needs[i] or needs[j])
            ops = gemm_operand_layouts(lay[i], lay[j], out)
            if ops is not None:
                # an operand produced by a GEMM-shaped step is written there in the layout this step
                # reads (summed indices lowest) instead of being copied into it before the step
                for tid, want in ((i, ops[0]), (j, ops[1])):
                    p = tid - self.n_inputs
                    if p < 0 or lay[tid] == want or tid == self.n_inputs + len(path) - 1:
                        continue
                    ps = self.steps[p]
                    if gemm_operand_layouts(lay[ps.a], lay[ps.b], lay[tid]) is None:
                        continue
                    lay[tid] = list(want)
                    self.steps[p] = _Step(ps.a, ps.b, tid, desc_rows(lay[ps.a], lay[ps.b], lay[tid]),
                                          desc_rows(lay[tid], lay[ps.b], lay[ps.a]) if needs[ps.a] else -1,
                                          desc_rows(lay[ps.a], lay[tid], lay[ps.b]) if needs[ps.b] else -1)
            fwd = desc_rows(lay[i], lay[j], out)
            ba = desc_rows(out, lay[j], lay[i]) if needs[i] else -1
            bb = desc_rows(lay[i], out, lay[j]) if needs[j] else -1
            self.steps.append(_Step(i, j, self.n_inputs + s, fwd, ba, bb))
        self.lay = lay
        self.ranks = [len(x) for x in lay]
        self.needs = needs
        # step index of the consumer of every tensor id
        self.consumer = {}
        for s, st in enumerate(self.steps):
            self.consumer[st.a] = s
            self.consumer[st.b] = s
        self.root = self.n_inputs + len(self.steps) - 1

    def intermediate_elems(self):
        return sum(1 << self.ranks[st.c] for st in self.steps)


def _external_paths(cc, paths):
    """Per network: (SSA path, sliced indices) from ``paths`` (SSA or linear, or path objects, one per
    measurement), or from the optimiser object the circuit was compiled with (the reference's
    ``use_cotengra`` / ``use_jdopttn`` calls, tedq/backends/compiled_circuit.py:340-393), or
    (None, None) for the own search."""
    # pylint: disable=protected-access
    nets = cc._tensor_networks
    if paths is not None:
        if len(paths) != len(nets):
            raise ValueError(f"{len(paths)} contraction paths given for {len(nets)} networks")
        return [(external_path(p, tn.num_tensors), sliced_indices(p)) for p, tn in zip(paths, nets)]
    hyper = cc._hyper_opt if isinstance(cc._hyper_opt, dict) else {}
    ctg, jd = cc._use_cotengra, cc._use_jdopttn
    out = []
    for tn in nets:
        tree = None
        if ctg is not None and hasattr(ctg, "HyperOptimizer"):
            opt = ctg.HyperOptimizer(methods=hyper.get("methods", ["kahypar"]),
                                     max_repeats=hyper.get("max_repeats", 128),
                                     progbar=hyper.get("progbar", False), minimize=hyper.get("minimize", "flops"),
                                     score_compression=hyper.get("score_compression", 0.5),
                                     slicing_opts=hyper.get("slicing_opts", None))
            tree = opt.search(tn.input_indices, tn.output_indices, tn.size_dict)
        elif callable(jd) and not isinstance(jd, bool):
            tree = jd(tn.input_indices, tn.size_dict, output=tn.output_indices, imbalance=0.2,
                      max_repeats=hyper.get("max_repeats", 128), search_parallel=hyper.get("search_parallel", True),
                      slicing_opts=hyper.get("slicing_opts", None))
        out.append((None, None) if tree is None else (external_path(tree, tn.num_tensors), sliced_indices(tree)))
    return out


def sliced_indices(tree):
    """Index symbols an external contraction tree slices (cotengra's ``sliced_inds``, a dict or a
    sequence), or None."""
    got = getattr(tree, "sliced_inds", None)
    if not got:
        return None
    return sorted(got.keys()) if isinstance(got, dict) else sorted(got)


class TNExecutor:
    """Contraction-mode executor of one compiled circuit (all its measurements).

    Same duck interface as the state-vector executor: ``run_forward``,
    ``measure``, ``seed_adjoint``, ``run_backward``, ``value_and_grad``,
    ``specs`` and ``n_params``.
    """

    def __init__(self, cc, trials=None, max_bytes=None, paths=None, gemm_form="4m", slicing=None,
                 diagonal_reduction=None):
        from .path import SEARCH_TRIALS  # pylint: disable=import-outside-toplevel
        trials = SEARCH_TRIALS if trials is None else trials
        self.n = cc._num_qubits  # pylint: disable=protected-access
        # bytes of intermediates the backward may keep before it checkpoints segments and recomputes
        # them; None: 40% of the device's memory (115 GB on a 288 GB MI355X), 64 GiB off the GPU
        self.max_bytes = max_bytes
        # forward: run the small steps of each tree level in one launch (tedq_tn_contract_small_batch)
        self.small_batch = True
        # backward: both environments of a gate step from one pass (tedq_tn_gate_backward)
        self.fused_backward = True
        self._small = {}
        ext = _external_paths(cc, paths)
        # diagonal reduction (the reference's tn_simplify, by gate type): the wires along which gates are
        # diagonal (all of a diagonal gate's, the controls of a controlled gate) are merged into
        # hyper-indices and the tensors stored without the repeated entries.  Off for externally given
        # orders, which refer to the network as built; otherwise used when the modelled device time of the
        # searched paths (summed over the networks) is lower with it than without it
        self._diag_ops = {idx: diagonal_wires(op) for idx, op in cc._operators.items()}  # pylint: disable=protected-access
        self._searched = None
        self.simplify = False
        self.simplify_report = None
        allowed = bool(getattr(cc, "_tn_simplify", True)) and all(p is None for p, _ in ext)
        if diagonal_reduction is not None:          # forced on or off (tests, experiments)
            self.simplify = allowed and bool(diagonal_reduction)
        elif allowed and any(self._diag_ops.values()):
            self.simplify, self._searched, self.simplify_report = self._choose_reduction(cc, trials)
        pool = []          # complex128 constants: matrices and constant parameter values
        recs = []
        self._gate_off = {}
        off = 0
        flat = 0
        param_recs = []
        for idx, op in cc._operators.items():  # pylint: disable=protected-access
            trainable = cc._trainableparams[idx]  # pylint: disable=protected-access
            nq = len(op.qubits)
            D = 1 << nq
            srcs = [_NO_SRC] * 3
            dsl = [-1] * 3
            for i, pos in enumerate(trainable):
                srcs[pos] = flat + i
                dsl[pos] = flat + i
            if trainable:
                name = op.name[:-4] if getattr(op, "is_inverse", False) else op.name
                if name not in _BLOCK_CODE:
                    raise NotImplementedError(f"trainable {op.name} is not supported in contraction mode")
                for pos in range(op.num_params):
                    if srcs[pos] == _NO_SRC:
                        from ..operators import as_float  # pylint: disable=import-outside-toplevel
                        pool.append(complex(as_float(cc._operatorparams[idx][pos]), 0.0))  # pylint: disable=protected-access
                        srcs[pos] = -1 - (len(pool) - 1)
                code, kind, coff = _BLOCK_CODE[name], 0, 0
            else:
                mat = np.asarray(op.matrix, dtype=np.complex128).reshape(D, D)
                coff = len(pool)
                pool.extend(mat.reshape(-1).tolist())
                code, kind = 0, 1
            dag = int(bool(getattr(op, "is_inverse", False)))
            # merged leading wires: a leaf keeps 2^dw x 4^(nq - dw) entries (D when fully diagonal)
            dw = self._diag_ops[idx] if self.simplify else 0
            if kind == 0 and 0 < dw < nq - 1:
                raise RuntimeError(f"{op.name}: a parametrised gate is diagonal in all wires or its controls")
            size = (1 << dw) << (2 * (nq - dw))
            ket, bra = off, off + size
            off += 2 * size
            self._gate_off[idx] = (ket, bra)
            for is_bra, o in ((0, ket), (1, bra)):
                d = dsl if not is_bra else [-1, -1, -1]
                recs.append([code, srcs[0], srcs[1], srcs[2], coff, nq, kind, is_bra, o, d[0], d[1], d[2],
                             dag, dw, 0, 0])
            flag = 16 if dw == nq else 32 if dw else 0      # leaf layout: diagonal / controls merged / full
            for i, _ in enumerate(trainable):
                param_recs.append([ket, bra, nq | flag, flat + i])
            flat += len(trainable)
        self.n_params = flat
        self.leaf_stride = max(off, 1)
        self._recs = np.asarray(recs, dtype=np.int32).reshape(-1, 16)
        self._param_recs = np.asarray(param_recs, dtype=np.int32).reshape(-1, 4)
        self._pool = np.asarray(pool if pool else [0j], dtype=np.complex128)

        # constant leaves (batch stride 0): |0>, the scalar 1, observables
        consts = [np.array([1.0, 0.0], dtype=np.complex64), np.array([1.0], dtype=np.complex64)]
        const_off = [0, 2]
        self._const_len = 3
        descs = []

        lists = []

        form = GEMM_FORMS[gemm_form]

        def desc_rows(la, lb, lc):
            descs.append(make_desc(la, lb, lc, form))
            lists.append((list(la), list(lb), list(lc)))
            return len(descs) - 1

        self.networks = []
        self.specs = []
        # own slicing: explicit options, else the reference's hyper_opt["slicing_opts"] (cotengra keys
        # target_size = largest intermediate in elements, target_slices = least number of slices)
        if slicing is None and isinstance(cc._hyper_opt, dict):  # pylint: disable=protected-access
            slicing = cc._hyper_opt.get("slicing_opts")  # pylint: disable=protected-access
        for tn in cc._tensor_networks:  # pylint: disable=protected-access
            meas = tn.measurement

            selectors = {}

            def leaf_src(s, diag=False, meas=meas, selectors=selectors):
                kind = s[0]
                if kind == "select":
                    # selector vector of a sliced index (both caps of index j share one slot; the
                    # slots of a network are contiguous and rewritten before each slice)
                    if s[1] not in selectors:
                        consts.append(np.array([1.0, 0.0], dtype=np.complex64))
                        const_off.append(self._const_len)
                        selectors[s[1]] = self._const_len
                        self._const_len += 2
                    return ("const", selectors[s[1]])
                if kind == "gate":
                    return ("arena", self._gate_off[s[1]][0])
                if kind == "adj":
                    return ("arena", self._gate_off[s[1]][1])
                if kind == "zero":
                    return ("const", 0)
                if kind == "one":
                    return ("const", 2)
                mat = meas.obs[s[1]].matrix if kind == "obs" else meas.obs.matrix
                m = np.asarray(mat, dtype=np.complex64)
                m = np.diag(m.reshape(int(np.sqrt(m.size)), -1)).copy() if diag else m.reshape(-1)
                o = self._const_len
                consts.append(m)
                const_off.append(o)
                self._const_len += m.size
                return ("const", o)

            ext_path, ext_sliced = ext[len(self.networks)]

            searched = self._searched[len(self.networks)] if self._searched else None
            net = _Network(tn, leaf_src, desc_rows, trials, ext_path, slicing, ext_sliced,
                           self._is_diagonal(meas) if self.simplify else None, searched)
            if net.sliced:
                offs = [selectors[j] for j in range(len(net.sliced))]
                if offs != list(range(offs[0], offs[0] + 2 * len(offs), 2)):
                    raise RuntimeError("tensor network: selector slots are not contiguous")
                net.sel_off = offs[0]
            self.networks.append(net)
            if meas.return_type is Expectation:
                self.specs.append(MeasSpec("fexp", ()))
            elif meas.return_type is Probability:
                k = len(net.output)
                self.specs.append(MeasSpec("probs", tuple([2] * k)))
            elif meas.return_type is State:
                self.specs.append(MeasSpec("state", tuple([2] * self.n)))
            else:
                raise NotImplementedError(f"measurement {meas.return_type} in contraction mode")
        # GEMM-shaped steps: copy operands whose summed indices are not already the lowest address bits
        # into that layout first (a contraction with the scalar 1), then contract the copies
        self._gemm = {}
        for row, (la, lb, lc) in enumerate(list(lists)):
            lay = gemm_operand_layouts(la, lb, lc)
            if lay is None:
                continue
            ca, cb = lay
            pa = desc_rows(la, [], ca) if ca != la else -1
            pb = desc_rows(lb, [], cb) if cb != lb else -1
            if pa < 0 and pb < 0:
                continue
            main = desc_rows(ca if pa >= 0 else la, cb if pb >= 0 else lb, lc)
            self._gemm[row] = (pa, len(ca), pb, len(cb), main)
        self._consts = np.concatenate(consts).astype(np.complex64)
        self._sel = {}
        self._descs = np.stack(descs).astype(np.int32) if descs else np.zeros((1, _DESC_WORDS), np.int32)
        self._dev = {}

    def _is_diagonal(self, meas):
        """Merged leading wires of network sources (gates by type, observables when diagonal)."""
        def pred(s):
            if s[0] in ("gate", "adj"):
                return self._diag_ops[s[1]]
            if s[0] == "obs":
                m = np.asarray(meas.obs[s[1]].matrix)
            elif s[0] == "obsmat":
                m = np.asarray(meas.obs.matrix)
            else:
                return 0
            m = m.reshape(int(np.sqrt(m.size)), -1)
            return int(np.log2(m.shape[0])) if _is_diagonal_matrix(m) else 0
        return pred

    def _choose_reduction(self, cc, trials):
        """Search every (light-cone reduced) network with and without diagonal reduction; keep the variant
        whose paths have the lower modelled device time in total.  Returns (use_reduction, the kept
        variant's searched paths per network, report)."""
        variants = {False: [], True: []}
        for tn in cc._tensor_networks:  # pylint: disable=protected-access
            red = cancel_inverse_pairs(tn)
            for use in (False, True):
                net = diagonal_reduce(red, self._is_diagonal(tn.measurement))[0] if use else red
                inputs = [list(t) for t in net.input_indices]
                if len(inputs) < 2:
                    variants[use].append(None)
                    continue
                variants[use].append(search_path(inputs, list(tn.output_indices), trials))
        model = {use: sum(v[2]["model_s"] for v in vs if v is not None) for use, vs in variants.items()}
        use = model[True] < model[False]
        report = {"diagonal_reduction": use, "model_s_with": model[True], "model_s_without": model[False]}
        return use, variants[use], report

    # ------------------------------------------------------------------
    def _buffers(self, device):
        key = str(device)
        if key not in self._dev:
            self._dev[key] = {
                "recs": torch.from_numpy(self._recs.reshape(-1).copy()).to(device),
                "precs": torch.from_numpy(self._param_recs.reshape(-1).copy()).to(device),
                "pool": torch.from_numpy(self._pool.view(np.float64).copy()).to(device),
                "consts": torch.from_numpy(self._consts.copy()).to(device),
                "descs": torch.from_numpy(self._descs.reshape(-1).copy()).to(device),
                "err": torch.zeros(1, dtype=torch.int32, device=device),
                "err_checked": False,
            }
        return self._dev[key]

    def _leaves(self, rows):
        dev = rows.device
        B = rows.shape[0]
        buf = self._buffers(dev)
        leaves = torch.empty((B, self.leaf_stride), dtype=torch.complex64, device=dev)
        dmats = torch.empty((B, max(self.n_params, 1) * 4), dtype=torch.complex64, device=dev)
        hip.call("tedq_tn_gate_tensors", hip.ptr(rows), int(rows.dtype == torch.float64), B, self.n_params,
                 hip.ptr(buf["recs"]), self._recs.shape[0], hip.ptr(buf["pool"]), hip.ptr(leaves),
                 self.leaf_stride, hip.ptr(dmats), dmats.shape[1], hip.ptr(buf["err"]), hip.stream_handle(dev))
        if not buf["err_checked"] and not torch.cuda.is_current_stream_capturing():
            check_device_flag(buf["err"], "tedq_tn_gate_tensors: a gate record holds an unknown gate code")
            buf["err_checked"] = True
        return leaves, dmats

    def _operand(self, net, vals, tid, leaves, buf):
        """(tensor, batch stride, element offset) of tensor id ``tid``."""
        if tid < net.n_inputs:
            where, off = net.leaves[tid]
            if where == "arena":
                return leaves, self.leaf_stride, off
            return buf["consts"], 0, off
        t = vals[tid]
        if isinstance(t, tuple):            # a small step's result inside the level arena
            return t
        return t, t.shape[1], 0

    # small steps, batched per tree level ---------------------------------
    SMALL_RANK = 12     # operands and result
    SMALL_SUM_BITS = 6  # summed indices

    def _small_plan(self, net):
        """The forward steps whose operands are leaves or results of such steps, with ranks <= 12 and at
        most 2^6 summed terms (the gate merges of a path), grouped by tree level: records for
        tedq_tn_contract_small_batch, level bounds, result offsets in one arena and its length."""
        key = id(net)
        if key in self._small:
            return self._small[key]
        offs, level, entries, arena = {}, {}, [], 0
        for st in net.steps:
            nh, nm, nn, nk = (int(v) for v in self._descs[st.fwd][:4])
            if (self._gemm.get(st.fwd) is not None or nk > self.SMALL_SUM_BITS
                    or max(nh + nm + nk, nh + nn + nk, nh + nm + nn) > self.SMALL_RANK
                    or not all(t < net.n_inputs or t in offs for t in (st.a, st.b))):
                continue
            ops = []
            for t in (st.a, st.b):
                if t < net.n_inputs:
                    where, o = net.leaves[t]
                    ops += [0 if where == "arena" else 1, o]
                else:
                    ops += [2, offs[t]]
            lv = 1 + max(level.get(t, 0) for t in (st.a, st.b))
            level[st.c] = lv
            offs[st.c] = arena
            size = 1 << net.ranks[st.c]
            entries.append((lv, [st.fwd] + ops + [arena, size, 0]))
            arena += size
        entries.sort(key=lambda e: e[0])
        recs = np.array([r for _, r in entries], dtype=np.int64).reshape(-1, 8)
        bounds, start = [], 0
        for i in range(1, len(entries) + 1):
            if i == len(entries) or entries[i][0] != entries[start][0]:
                bounds.append((start, i))
                start = i
        plan = {"recs": recs, "bounds": bounds, "offs": offs, "len": max(arena, 1), "dev": {}}
        self._small[key] = plan
        return plan

    def _small_levels(self, plan, arena, leaves, buf, B, dev):
        """Run a plan's levels in order, one launch each, results into ``arena`` [B, plan len]."""
        key = str(dev)
        if key not in plan["dev"]:
            plan["dev"][key] = torch.from_numpy(plan["recs"].reshape(-1).copy()).to(dev)
        drecs = plan["dev"][key]
        consts = buf["consts"]
        bases = (ctypes.c_void_p * 3)(leaves.data_ptr(), consts.data_ptr(), arena.data_ptr())
        strides = (ctypes.c_int64 * 3)(self.leaf_stride, 0, plan["len"])
        lens = (ctypes.c_int64 * 3)(self.leaf_stride, consts.numel(), plan["len"])
        hrecs = plan["recs"]
        for s, e in plan["bounds"]:
            hip.call("tedq_tn_contract_small_batch", bases, strides, lens,
                     hrecs[s:e].ctypes.data_as(hip.c_vp), ctypes.c_void_p(drecs.data_ptr() + 64 * s), e - s, B,
                     self._descs.ctypes.data_as(hip.c_vp), hip.ptr(buf["descs"]), self._descs.shape[0],
                     hip.stream_handle(dev))

    def _contract(self, row, a, b, c, B, dev, buf):
        g = self._gemm.get(row)
        if g is not None:
            pa, ra, pb, rb, main = g
            one = (buf["consts"], 0, 2)
            if pa >= 0:
                t = torch.empty((B, 1 << ra), dtype=torch.complex64, device=dev)
                self._contract_raw(pa, a, one, (t, t.shape[1], 0), B, dev, buf)
                a = (t, t.shape[1], 0)
            if pb >= 0:
                t = torch.empty((B, 1 << rb), dtype=torch.complex64, device=dev)
                self._contract_raw(pb, b, one, (t, t.shape[1], 0), B, dev, buf)
                b = (t, t.shape[1], 0)
            return self._contract_raw(main, a, b, c, B, dev, buf)
        return self._contract_raw(row, a, b, c, B, dev, buf)

    def _contract_raw(self, row, a, b, c, B, dev, buf):
        (ta, sa, oa), (tb, sb, ob), (tc, sc, oc) = a, b, c
        lib = hip.load()
        desc = self._descs[row]
        hdesc = desc.ctypes.data_as(hip.c_vp)
        ws_bytes = lib.tedq_tn_workspace_bytes(hdesc, B)
        if ws_bytes < 0:
            raise hip.EngineError("tensor network: invalid step descriptor")
        ws = torch.empty(max(ws_bytes, 1), dtype=torch.uint8, device=dev) if ws_bytes else None
        ddesc = ctypes.c_void_p(buf["descs"].data_ptr() + row * _DESC_WORDS * 4)
        hip.call("tedq_tn_contract", ctypes.c_void_p(ta.data_ptr() + 8 * oa), sa,
                 ctypes.c_void_p(tb.data_ptr() + 8 * ob), sb, ctypes.c_void_p(tc.data_ptr() + 8 * oc), sc, B,
                 hdesc, ddesc, hip.ptr(ws), ws_bytes, hip.stream_handle(dev))
        return ws

    def _set_selectors(self, net, a, dev):
        """Write the selector vectors of slice ``a`` (bit j of ``a`` = value of sliced index j) into the
        network's constant slots; stream-ordered after the previous slice's launches."""
        key = (str(dev), id(net))
        sel = self._sel.get(key)
        if sel is None:
            s = len(net.sliced)
            vals = np.zeros((1 << s, s, 2), dtype=np.complex64)
            for i in range(1 << s):
                for j in range(s):
                    vals[i, j, (i >> j) & 1] = 1.0
            sel = self._sel[key] = torch.from_numpy(vals.reshape(1 << s, 2 * s)).to(dev)
        consts = self._buffers(dev)["consts"]
        consts[net.sel_off:net.sel_off + sel.shape[1]].copy_(sel[a])

    def _forward_network(self, net, leaves, B, dev, keep=None):
        """Contract one network; ``keep(tid)`` says which intermediates to retain.  Returns (root, kept).
        A sliced network (forward only) is contracted once per slice and the roots are added in slice
        order."""
        if net.sliced and keep is None:
            total = None
            for a in range(1 << len(net.sliced)):
                self._set_selectors(net, a, dev)
                root, _ = self._forward_once(net, leaves, B, dev)
                total = root.clone() if total is None else total.add_(root)
            return total, {}
        return self._forward_once(net, leaves, B, dev, keep)

    def _forward_once(self, net, leaves, B, dev, keep=None):
        buf = self._buffers(dev)
        vals = {}
        kept = {}
        small = self._small_plan(net) if self.small_batch else None
        if small is not None and small["bounds"]:
            arena = torch.empty((B, small["len"]), dtype=torch.complex64, device=dev)
            self._small_levels(small, arena, leaves, buf, B, dev)
            for tid, o in small["offs"].items():
                vals[tid] = (arena, small["len"], o)
                if keep is not None and tid != net.root and keep(tid):
                    kept[tid] = vals[tid]
        for st in net.steps:
            if small is not None and st.c in small["offs"]:
                continue
            out = torch.empty((B, 1 << net.ranks[st.c]), dtype=torch.complex64, device=dev)
            vals[st.c] = out
            self._contract(st.fwd, self._operand(net, vals, st.a, leaves, buf),
                           self._operand(net, vals, st.b, leaves, buf), (out, out.shape[1], 0), B, dev, buf)
            for t in (st.a, st.b):
                if t >= net.n_inputs:
                    if keep is not None and keep(t):
                        kept[t] = vals[t]
                    del vals[t]
        root = vals[net.root]
        if isinstance(root, tuple):
            t, _, o = root
            root = t[:, o:o + (1 << net.ranks[net.root])].contiguous()
        return root, kept

    # ------------------------------------------------------------------
    def run_forward(self, rows, need_grad=False):  # pylint: disable=unused-argument
        """Contract every network for [B, P] rows.  Returns a run record for ``measure``/``run_backward``."""
        B = rows.shape[0]
        dev = rows.device
        leaves, dmats = self._leaves(rows)
        roots = [self._forward_network(net, leaves, B, dev)[0] for net in self.networks]
        return {"rows": rows, "leaves": leaves, "dmats": dmats, "roots": roots}

    def measure(self, run):
        outs = []
        B = run["rows"].shape[0]
        for spec, root in zip(self.specs, run["roots"]):
            if spec.kind == "fexp":
                outs.append(root.reshape(B).real)
            elif spec.kind == "probs":
                outs.append(root.real.reshape((B,) + spec.shape))
            else:
                outs.append(root.reshape((B,) + spec.shape))
        return outs

    def seed_adjoint(self, run, grads):
        """Upstream gradients per measurement (kept with the run; None = no contribution)."""
        return list(grads)

    # ------------------------------------------------------------------
    def _keep_budget(self, dev=None):
        if self.max_bytes is not None:
            return self.max_bytes
        if dev is not None and dev.type == "cuda":
            return int(0.4 * torch.cuda.get_device_properties(dev).total_memory)
        return 64 << 30

    def _segments(self, net, B, dev=None):
        total = net.intermediate_elems() * B * 8
        if total <= self._keep_budget(dev):
            return [list(range(len(net.steps)))]
        seg = max(1, int(math.ceil(math.sqrt(len(net.steps)))))
        return [list(range(s, min(s + seg, len(net.steps)))) for s in range(0, len(net.steps), seg)]

    def _backward_network(self, net, leaves, dmats, root_env, grad, B, dev):
        """Reverse contraction tree of one network; gradients are added into ``grad``.  Returns the
        network's root.  A sliced network runs forward and backward once per slice: the parameter
        gradients of the slices add up in ``grad`` and their roots are added in slice order."""
        if not net.sliced:
            return self._backward_once(net, leaves, dmats, root_env, grad, B, dev)
        total = None
        for a in range(1 << len(net.sliced)):
            self._set_selectors(net, a, dev)
            root = self._backward_once(net, leaves, dmats, root_env, grad, B, dev)
            total = root.clone() if total is None else total.add_(root)
        return total

    def _backward_once(self, net, leaves, dmats, root_env, grad, B, dev):
        buf = self._buffers(dev)
        segs = self._segments(net, B, dev)
        seg_of = {}
        for si, seg in enumerate(segs):
            for s in seg:
                seg_of[net.steps[s].c] = si
        # forward, checkpointing the tensors that cross a segment boundary
        if len(segs) == 1:
            def keep(_t):
                return True
        else:
            def keep(t):
                return seg_of.get(t, -1) != seg_of.get(net.steps[net.consumer[t]].c, -1)
        root, stored = self._forward_once(net, leaves, B, dev, keep)
        env_arena = torch.zeros((B, self.leaf_stride), dtype=torch.complex64, device=dev)
        envs = {net.root: root_env}
        for si in range(len(segs) - 1, -1, -1):
            seg = segs[si]
            vals = dict(stored)
            if len(segs) > 1:
                # recompute the inside of this segment from its checkpointed inputs
                for s in seg:
                    st = net.steps[s]
                    if st.c in vals:
                        continue
                    out = torch.empty((B, 1 << net.ranks[st.c]), dtype=torch.complex64, device=dev)
                    vals[st.c] = out
                    self._contract(st.fwd, self._operand(net, vals, st.a, leaves, buf),
                                   self._operand(net, vals, st.b, leaves, buf), (out, out.shape[1], 0), B, dev,
                                   buf)
            for s in reversed(seg):
                st = net.steps[s]
                ec = envs.pop(st.c, None)
                if ec is None:
                    continue
                ecop = (ec, ec.shape[1], 0)
                dsts = {}
                for which, desc_row, tid in (("a", st.bwd_a, st.a), ("b", st.bwd_b, st.b)):
                    if desc_row < 0:
                        continue
                    if tid < net.n_inputs:
                        where, off = net.leaves[tid]
                        if where != "arena":
                            continue
                        dsts[which] = (env_arena, self.leaf_stride, off)
                    else:
                        e = torch.empty((B, 1 << net.ranks[tid]), dtype=torch.complex64, device=dev)
                        envs[tid] = e
                        dsts[which] = (e, e.shape[1], 0)
                aop = self._operand(net, vals, st.a, leaves, buf)
                bop = self._operand(net, vals, st.b, leaves, buf)
                if len(dsts) == 2 and self.fused_backward and _gate_step(self._descs[st.fwd]):
                    self._gate_backward(st, ecop, aop, bop, dsts["a"], dsts["b"], B, dev, buf)
                    dsts = {}
                if "a" in dsts:
                    self._contract(st.bwd_a, ecop, bop, dsts["a"], B, dev, buf)
                if "b" in dsts:
                    self._contract(st.bwd_b, aop, ecop, dsts["b"], B, dev, buf)
                if st.c in vals and st.c != net.root:
                    del vals[st.c]
            for s in seg:
                stored.pop(net.steps[s].c, None)
        if self.n_params:
            self._param_grad(env_arena, dmats, grad, B)
        return root

    def _gate_backward(self, st, ecop, aop, bop, dst_a, dst_b, B, dev, buf):
        """Both environments of a gate step from one read of E_C (tedq_tn_gate_backward)."""
        desc = self._descs[st.fwd]
        a_big = int(desc[1]) >= 10
        x, g, dx, dg = (aop, bop, dst_a, dst_b) if a_big else (bop, aop, dst_b, dst_a)
        lib = hip.load()
        hdesc = desc.ctypes.data_as(hip.c_vp)
        ws_bytes = lib.tedq_tn_gate_backward_workspace(hdesc, B)
        if ws_bytes < 0:
            raise hip.EngineError("tensor network: step is not a gate step")
        ws = torch.empty(max(ws_bytes, 1), dtype=torch.uint8, device=dev)
        ddesc = ctypes.c_void_p(buf["descs"].data_ptr() + st.fwd * _DESC_WORDS * 4)

        def p(op):
            return ctypes.c_void_p(op[0].data_ptr() + 8 * op[2]), op[1]
        hip.call("tedq_tn_gate_backward", *p(ecop), *p(x), *p(g), *p(dx), *p(dg), B, hdesc, ddesc, hip.ptr(ws),
                 ws_bytes, hip.stream_handle(dev))

    def _param_grad(self, env_arena, dmats, grad, B):
        dev = env_arena.device
        buf = self._buffers(dev)
        ones = torch.ones(B, dtype=torch.float32, device=dev)
        hip.call("tedq_tn_param_grad", hip.ptr(env_arena), self.leaf_stride, hip.ptr(dmats), dmats.shape[1],
                 hip.ptr(buf["precs"]), self.n_params, hip.ptr(ones), hip.ptr(grad), B, hip.stream_handle(dev))

    @staticmethod
    def _root_env(spec, g, B, net):
        """Environment of the root: conj of the upstream gradient (real outputs: the gradient itself)."""
        if spec.kind == "fexp":
            return g.to(torch.float32).reshape(B, 1).to(torch.complex64)
        if spec.kind == "probs":
            return g.to(torch.float32).reshape(B, 1 << len(net.output)).to(torch.complex64)
        return g.to(torch.complex64).reshape(B, -1).conj().contiguous()

    def run_backward(self, rows, run, grads, grad_dtype=torch.float32):
        """Gradient [B, P] of sum_m <grads_m, out_m> by the reverse contraction tree of every network."""
        B = rows.shape[0]
        dev = rows.device
        grad = torch.zeros((B, max(self.n_params, 1)), dtype=torch.float64, device=dev)
        for spec, net, g in zip(self.specs, self.networks, grads):
            if g is None:
                continue
            self._backward_network(net, run["leaves"], run["dmats"], self._root_env(spec, g, B, net), grad, B, dev)
        return grad[:, :self.n_params].to(grad_dtype)

    def value_and_grad(self, rows, grads_out):
        """Outputs and gradient, contracting each network once with its intermediates kept for the backward."""
        B = rows.shape[0]
        dev = rows.device
        leaves, dmats = self._leaves(rows)
        grad = torch.zeros((B, max(self.n_params, 1)), dtype=torch.float64, device=dev)
        roots = []
        for spec, net, g in zip(self.specs, self.networks, grads_out):
            if g is None:
                roots.append(self._forward_network(net, leaves, B, dev)[0])
            else:
                roots.append(self._backward_network(net, leaves, dmats, self._root_env(spec, g, B, net), grad, B,
                                                    dev))
        outs = self.measure({"rows": rows, "roots": roots})
        return outs, grad[:, :self.n_params].to(rows.dtype)

    def memory_bytes(self, batch, with_grad=True):  # pylint: disable=unused-argument
        peak = max((net.intermediate_elems() for net in self.networks), default=0)
        return batch * 8 * (self.leaf_stride * 2 + min(peak, self._keep_budget() // max(batch * 8, 1)))

    def path_report(self):
        return [dict(net.path_info, steps=len(net.steps), tensors=net.n_inputs) for net in self.networks]
