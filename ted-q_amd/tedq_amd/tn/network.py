"""Circuit -> bra/ket tensor network, one network per measurement.

Restates the index bookkeeping of the reference ``gen_tensor_networks``
(tedq/tensor_network/tensor_network.py:850-1099) and ``get_symbol``
(:1109-1127):

* n initial |0> vectors with indices [a_q];
* every gate in program order gets indices [new..., old...] (output wires
  first, then the wires it consumes), matching a gate tensor stored as
  matrix.reshape([2]*2k) with (out..., in...) axes;
* per measurement:
    - ``state``: output indices are the current wires, nothing else is added;
    - ``expval``: the observable tensor(s) are appended like gates, then the
      adjoint gates in reverse program order, then n closing <0| vectors;
    - ``probs(qubits)``: the current wires of ``qubits`` become output
      indices (shared by ket and bra, i.e. a diagonal), then adjoint gates
      and closing vectors as for ``expval``.

Each tensor also records where its data comes from at execution time, so
the device executor can assemble the operands for any parameter batch.
"""

_EINSUM_SYMBOLS_BASE = "abcdefghijklmnopqrstuvwxyzABCDEFGHIJKLMNOPQRSTUVWXYZ"


def get_symbol(i):
    """Index name of integer ``i``: the 52 ASCII letters, then ``chr(i + 140)``."""
    if i < 52:
        return _EINSUM_SYMBOLS_BASE[i]
    return chr(i + 140)


class TensorNetwork:
    """Index structure of one network plus the data source of every tensor.

    ``sources[t]`` is one of ``("zero",)``, ``("gate", instance_id)``,
    ``("adj", instance_id)``, ``("obs", k)`` (k-th observable factor of the
    measurement) or ``("obsmat",)`` (the whole observable matrix).
    """

    def __init__(self, input_indices, output_indices, size_dict, sources, measurement):
        if len(input_indices) != len(sources):
            raise ValueError(
                f"Error! number of input arrays {len(sources)} and input indices {len(input_indices)} do not match!!")
        used = set()
        for ix in input_indices:
            used |= set(ix)
        used |= set(output_indices)
        if used ^ set(size_dict):
            raise ValueError("Error! indices of input do not match with size_dict!!")
        self.input_indices = input_indices
        self.output_indices = output_indices
        self.size_dict = size_dict
        self.sources = sources
        self.measurement = measurement

    @property
    def num_tensors(self):
        return len(self.input_indices)


def _append_op(indices, sources, layer, cur, qubits, source):
    k = len(qubits)
    cur += k
    new = [get_symbol(cur - (k - 1 - i)) for i in range(k)]
    old = [get_symbol(layer[q]) for q in qubits]
    indices.append(new + old)
    sources.append(source)
    for i, q in enumerate(qubits):
        layer[q] = cur - (k - 1 - i)
    return cur


def gen_tensor_networks(num_qubits, operators, appliedqubits, measurements):
    """One ``TensorNetwork`` per measurement (see module docstring)."""
    from ..measurement import Expectation, Probability, State  # pylint: disable=import-outside-toplevel

    layer = list(range(num_qubits))
    cur = num_qubits - 1
    base_ix = [[get_symbol(i)] for i in range(num_qubits)]
    base_src = [("zero",)] * num_qubits
    for idx, qbts in appliedqubits.items():
        cur = _append_op(base_ix, base_src, layer, cur, list(qbts), ("gate", idx))

    networks = []
    for meas in measurements:
        ix = [list(x) for x in base_ix]
        src = list(base_src)
        lay = list(layer)
        c = cur
        out = []
        if meas.return_type is State:
            out = [get_symbol(lay[i]) for i in range(num_qubits)]
        else:
            if meas.return_type is Expectation:
                if isinstance(meas.obs, list):
                    for k, ob in enumerate(meas.obs):
                        c = _append_op(ix, src, lay, c, [ob.qubits[0]], ("obs", k))
                else:
                    c = _append_op(ix, src, lay, c, list(meas.obs.qubits), ("obsmat",))
            elif meas.return_type is Probability and meas.qubits is not None:
                out = [get_symbol(lay[q]) for q in meas.qubits]
            for idx, qbts in reversed(appliedqubits.items()):
                if len(qbts) > 3:
                    raise ValueError("Error!! unknown operator with len of applied qubits larger than 3!")
                c = _append_op(ix, src, lay, c, list(qbts), ("adj", idx))
            for i in lay:
                ix.append([get_symbol(i)])
                src.append(("zero",))
        size = {}
        for t in ix:
            for s in t:
                size[s] = 2
        for s in out:
            size[s] = 2
        networks.append(TensorNetwork(ix, out, size, src, meas))
    return networks


def cancel_inverse_pairs(tn):
    """Drop ket/bra gate pairs that multiply to the identity (the light-cone rule).

    A ket gate ``U`` with indices [new..., old...] and its bra copy
    ``U^dagger`` with indices [new'..., old'...] contract to the identity
    when old' == new (the bra copy consumes exactly the ket gate's outputs,
    in order): both are removed and new' is renamed to old.  Processed in the
    bra's order (outermost pair first), so cancellations cascade inward from
    the closing <0| vectors.  A <0| vector and a |0> vector left on the same
    index contract to 1 and are removed too.  Pairs that touch an output
    index (probability outputs) are kept.  The reference contracts these
    pairs numerically (its TensorNetwork.simplify is off in every benchmark,
    tedq/tensor_network/tensor_network.py:90-221); the value is the same up to
    single-precision rounding.
    """
    ix = [list(t) for t in tn.input_indices]
    src = list(tn.sources)
    out = set(tn.output_indices)
    alive = [True] * len(ix)
    ket = {s[1]: t for t, s in enumerate(src) if s[0] == "gate"}
    ren = {}

    def root(s):
        while s in ren:
            s = ren[s]
        return s

    for t, s in enumerate(src):
        if s[0] != "adj" or s[1] not in ket:
            continue
        g = ket[s[1]]
        k = len(ix[g]) // 2
        knew = [root(x) for x in ix[g][:k]]
        bold = [root(x) for x in ix[t][k:]]
        if knew == bold and not out.intersection(knew):
            for a, b in zip(ix[t][:k], ix[g][k:]):
                ren[a] = root(b)
            alive[g] = alive[t] = False
    kept = [t for t in range(len(ix)) if alive[t]]
    new_ix = [[root(x) for x in ix[t]] for t in kept]
    new_src = [src[t] for t in kept]
    zeros = {}
    for j, (t, s) in enumerate(zip(new_ix, new_src)):
        if s[0] == "zero" and len(t) == 1 and t[0] not in out:
            zeros.setdefault(t[0], []).append(j)
    drop = {j for js in zeros.values() if len(js) == 2 for j in js}
    fin_ix = [t for j, t in enumerate(new_ix) if j not in drop]
    fin_src = [s for j, s in enumerate(new_src) if j not in drop]
    size = {x: 2 for t in fin_ix for x in t}
    for x in tn.output_indices:
        size[x] = 2
    return TensorNetwork(fin_ix, list(tn.output_indices), size, fin_src, tn.measurement)


def diagonal_reduce(tn, diagonal_wires):
    """Collapse the wires along which tensors are diagonal into hyper-indices.

    A gate tensor [new..., old...] that is diagonal in wire i (its matrix
    entries vanish wherever the row and column differ in that wire's bit)
    carries the same value on new_i and old_i: the two indices merge into one.
    All wires of a diagonal gate (RZ, PhaseShift, Z, S, T, CZ, CRZ, controlled
    phase, diagonal observables and user matrices) merge, and the tensor keeps
    its diagonal, [w...]; the control wires of a controlled gate (CNOT, CY,
    CRX, CRY, Toffoli, CSWAP) merge and the tensor becomes
    [w_controls..., new_targets..., old_targets...].  Chains of such wires
    become one index shared by all tensors on them.  The reference's
    TensorNetwork.diagonal_reduce (tensor_network.py:316-392) finds diagonal
    axis pairs from the trace-time array values, which changes meaning when a
    parameter later takes another value; here ``diagonal_wires(source)`` gives
    the number of leading wires (in the gate's qubit order) that are diagonal
    for every parameter value, from the gate type or a constant matrix.
    Output indices keep their names.  Returns (network, diag) with diag[t] the
    number of merged leading wires of tensor t (0: unchanged).
    """
    parent = {}

    def root(s):
        while parent.get(s, s) != s:
            s = parent[s]
        return s

    out = set(tn.output_indices)
    diag = []
    for t, (ix, s) in enumerate(zip(tn.input_indices, tn.sources)):
        d = 0
        if s[0] in ("gate", "adj", "obs", "obsmat") and len(ix) % 2 == 0 and len(ix) > 0:
            d = min(int(diagonal_wires(s)), len(ix) // 2)
        diag.append(d)
        if not d:
            continue
        k = len(ix) // 2
        for a, b in zip(ix[:d], ix[k:k + d]):
            ra, rb = root(a), root(b)
            if ra == rb:
                continue
            # the merged wire is named by an output index when it holds one, else by its earliest symbol
            if ra in out and rb in out:
                raise ValueError("diagonal reduction would merge two output indices")
            keep, drop = (ra, rb) if ra in out else (rb, ra)
            parent[drop] = keep
    new_ix = []
    for ix, d in zip(tn.input_indices, diag):
        if d:
            k = len(ix) // 2
            new_ix.append([root(x) for x in ix[k:k + d]] + [root(x) for x in ix[d:k]] + [root(x) for x in ix[k + d:]])
        else:
            new_ix.append([root(x) for x in ix])
    size = {x: 2 for t in new_ix for x in t}
    for x in tn.output_indices:
        size[x] = 2
    return TensorNetwork(new_ix, list(tn.output_indices), size, list(tn.sources), tn.measurement), diag
