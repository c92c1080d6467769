"""Backend-neutral compiled circuit: parameter bookkeeping and public properties.

Mirrors the reference ``CompiledCircuit`` (tedq/backends/compiled_circuit.py:34-647):
per-gate ordered dicts keyed by ``instance_id``, the tensordot axes /
permutation lists of the state-vector method (:126-202, kept because they are
part of the observable surface), the tensor-network operands of the
contraction method (:207-336), ``_update_parameters`` (:492-547) which stores
the call arguments by reference in gate order, and the properties.
"""

from collections import OrderedDict

from ..tn.network import gen_tensor_networks


class CompiledCircuit:
    """Structure of a traced circuit as seen by an execution backend."""

    def __init__(self, backend, circuit, use_cotengra=False, use_jdopttn=False, tn_mode=False, hyper_opt=None,
                 tn_simplify=False, print_output=False):
        self._calculation_mode = False
        self._use_cotengra = use_cotengra
        self._use_jdopttn = use_jdopttn
        if use_cotengra and use_jdopttn:
            raise ValueError("Error!!!! can not use contengra, opt_einsum and cyc at the same time!")
        outer = bool(use_cotengra or use_jdopttn)
        self._tn_mode = tn_mode
        if outer and tn_mode:
            raise ValueError("Error!!!! can not use contengra, opt_einsum and cyc at the same time!")
        self._tn_simplify = tn_simplify
        self._hyper_opt = hyper_opt
        self._num_qubits = circuit.num_qubits
        self._backend = backend
        self._operatorname = OrderedDict()
        self._appliedqubits = OrderedDict()
        self._operatorparams = OrderedDict()
        self._trainableparams = OrderedDict()
        self._operators = OrderedDict()
        self._matrixs = OrderedDict()
        for op in circuit.operators:
            self._operators[op.instance_id] = op
            self._operatorname[op.instance_id] = op.name
            self._appliedqubits[op.instance_id] = op.qubits
            self._operatorparams[op.instance_id] = op.parameters
            self._trainableparams[op.instance_id] = op.trainable_params
            self._matrixs[op.instance_id] = op.matrix
        self._measurements = circuit.measurements
        self._init_state = circuit.init_state

        n = self._num_qubits
        self._axeslist = []
        self._permutationlist = []
        flops = 0.0
        for qbts in self._appliedqubits.values():
            k = len(qbts)
            if k == 1:
                gate_pos = [1]
                perms = list(range(1, qbts[0] + 1)) + [0] + list(range(qbts[0] + 1, n))
            else:
                gate_pos = list(range(k, 2 * k))
                order = list(qbts) + [x for x in range(n) if x not in qbts]
                perms = [order.index(i) for i in range(n)]
            flops += (2 ** n) * (2 * (2 ** k) - 1)
            self._axeslist.append((gate_pos, qbts))
            self._permutationlist.append(perms)
        self._axeslist.reverse()
        self._permutationlist.reverse()
        if print_output:
            import math  # pylint: disable=import-outside-toplevel
            print("log10(flops) of this quantum circuit:  ", math.log10(flops + 1.0e-10))

        self._tensor_networks = None
        self._cotengra_operands = []
        if self._use_cotengra or self._use_jdopttn or self._tn_mode:
            if self._init_state:
                raise ValueError(
                    "Error!!!! tensor network contraction mode do not support user-defined initial quantum state!")
            self._tensor_networks = gen_tensor_networks(n, self._operators, self._appliedqubits, self._measurements)
            for tn in self._tensor_networks:
                self._cotengra_operands.append((tn.input_indices, tn.output_indices, tn.size_dict))

    # ------------------------------------------------------------------
    def execute(self, *params):
        if params:
            self._update_parameters(*params)

    def _update_parameters(self, *params):
        """Write flattened call arguments into the trainable slots, in gate order."""
        count = 0
        for idx, trainable in self._trainableparams.items():
            if trainable:
                for i, pos in enumerate(trainable):
                    if count + i < len(params):
                        self._operatorparams[idx][pos] = params[count + i]
                count += len(trainable)
        if len(params) != count:
            raise ValueError(
                f"Error!!!! number of parameters are not matched!! required {count} but {len(params)} are given")

    def trainable_count(self):
        return sum(len(t) for t in self._trainableparams.values())

    # ------------------------------------------------------------------
    @property
    def operators(self):
        return list(self._operators.values())

    @property
    def gates_names(self):
        return list(self._operatorname.values())

    @property
    def gate_parameters(self):
        return list(self._operatorparams.values())

    @property
    def parameters(self):
        out = []
        for pars in self._operatorparams.values():
            out.extend(pars)
        return out

    @property
    def trainable_parameters(self):
        out = []
        for idx, pars in self._trainableparams.items():
            for i in pars:
                out.append(self._operatorparams[idx][i])
        return out

    @property
    def qubits(self):
        return list(self._appliedqubits.values())

    @property
    def interface(self):
        return self._interface  # pylint: disable=no-member

    @property
    def diff_method(self):
        return self._diff_method  # pylint: disable=no-member

    @property
    def backend(self):
        return self._backend

    @property
    def measurements(self):
        return self._measurements

    @property
    def calculation_mode(self):
        return self._calculation_mode
