"""Host-side front end and compiled-circuit bookkeeping (no GPU).

Restates the reference's host tests: test/test_circuit.py (storage, tracing,
errors), test/test_operators.py (construction checks, adjoint/inverse),
test/test_compiled_circuit.py (axes/permutation lists, TN index structure,
trainable-parameter bookkeeping) and the construction/kwargs parts of
test/test_pytorch_backend.py:38-136.
"""

import math
import random

import numpy as np
import pytest
import torch

import tedq_amd as qai
from tedq_amd.storage import StorageBase, CircuitStorage
from tedq_amd.operators import GateBase
from tedq_amd.measurement import QuantumMeasurement, Expectation, Probability, State
from tedq_amd.quantum_error import QuantumStorageError, QuantumCircuitError
from tedq_amd.backends import CompiledCircuit, PyTorchBackend


def test_storage_recording_append_remove():
    st = CircuitStorage()
    assert isinstance(st, StorageBase)
    with st as s:
        assert CircuitStorage.recording()
        assert s is CircuitStorage.active_context()
        objs = [1, ["a"], ("b",), {"c": 1}]
        for o in objs:
            CircuitStorage.append(o)
        assert all(x is y for x, y in zip(st.storage_context, objs))
        CircuitStorage.remove(objs[0])
        assert objs[0] not in st.storage_context
    assert CircuitStorage.recording() is False


def _two(*p):
    qai.RY(p[0], qubits=[0])
    qai.RZ(p[1], qubits=[1])
    return qai.expval(qai.PauliZ(qubits=[0]))


def test_circuit_construction_and_errors():
    params = (random.random(), random.random())
    assert isinstance(qai.Circuit(_two, 2, *params), qai.Circuit)
    with pytest.raises(ValueError):
        qai.Circuit(_two, 1, *params)

    def bad(*p):
        qai.RY(p[0], qubits=[0])
        CircuitStorage.append("wrong")
        return qai.expval(qai.PauliZ(qubits=[0]))
    with pytest.raises(QuantumStorageError):
        qai.Circuit(bad, 1, *params)

    def no_meas(*p):
        qai.RY(p[0], qubits=[0])
    with pytest.raises(QuantumCircuitError):
        qai.Circuit(no_meas, 1, 0.1)
    with pytest.raises(ValueError):
        qai.RX(0.1, qubits=[0])          # no active storage
    with pytest.raises(ValueError):
        qai.RX(0.1, 0.2, qubits=[0], do_queue=False)
    with pytest.raises(ValueError):
        qai.CNOT(qubits=[0], do_queue=False)
    with pytest.raises(TypeError):
        qai.RX(0.1)                       # qubits is keyword-only and required


def test_gates_inside_circuit_and_instance_ids():
    def cd(*p):
        qai.Hadamard(qubits=[0]); qai.PauliX(qubits=[1]); qai.PauliY(qubits=[2]); qai.PauliZ(qubits=[3])
        qai.I(qubits=[4]); qai.S(qubits=[5]); qai.T(qubits=[6]); qai.SX(qubits=[7])
        qai.CNOT(qubits=[0, 1]); qai.CY(qubits=[1, 2]); qai.CZ(qubits=[2, 3]); qai.SWAP(qubits=[3, 4])
        qai.CSWAP(qubits=[1, 2, 0]); qai.Toffoli(qubits=[1, 2, 0])
        qai.RX(p[0], qubits=[0]); qai.RY(p[1], qubits=[1]); qai.RZ(p[2], qubits=[2])
        qai.Rot(p[3], p[4], p[5], qubits=[0]); qai.PhaseShift(p[6], qubits=[0])
        qai.ControlledPhaseShift(p[7], qubits=[1, 0]); qai.CRX(p[8], qubits=[0, 3])
        qai.CRY(p[9], qubits=[0, 3]); qai.CRZ(p[10], qubits=[0, 3])
        return qai.expval(qai.PauliZ(qubits=[0]))
    c = qai.Circuit(cd, 8, *[0.1 * i for i in range(11)])
    gates = c.operators
    assert len(gates) == 23 and gates[0].name == "Hadamard" and gates[22].name == "CRZ"
    ids = [g.instance_id for g in gates]
    assert ids == sorted(ids) and len(set(ids)) == 23
    assert all(isinstance(g, GateBase) for g in gates)


def test_measurements_replace_observables():
    def cd():
        qai.Hadamard(qubits=[0])
        return [qai.expval(qai.PauliZ(qubits=[0])), qai.probs(qubits=[0]), qai.state(),
                qai.expval([qai.PauliX(qubits=[0]), qai.PauliZ(qubits=[1])])]
    c = qai.Circuit(cd, 2)
    assert len(c.operators) == 1
    kinds = [m.return_type for m in c.measurements]
    assert kinds == [Expectation, Probability, State, Expectation]
    assert isinstance(c.measurements[3].obs, list)
    with pytest.raises(ValueError):
        QuantumMeasurement(Expectation, obs=qai.PauliZ(qubits=[0], do_queue=False), qubits=[0], do_queue=False)
    with pytest.raises(qai.QuantumValueError):
        def cd2():
            qai.expval(qai.RX(0.1, qubits=[0]))
        qai.Circuit(cd2, 1)


def test_measurement_only_and_init_state_circuits():
    c = qai.Circuit(lambda: qai.state(), 3)
    assert c.init_state is None and c.operators == []
    init = np.zeros(4)
    init[3] = 1

    def cd():
        qai.IintStateVector(init)
        qai.PauliX(qubits=[0])
        return qai.state()
    c2 = qai.Circuit(cd, 2)
    assert c2.init_state is not None and len(c2.operators) == 1


def test_gate_matrices_inverse_adjoint():
    rx = qai.RX(0.5, qubits=[0], do_queue=False)
    assert np.allclose(rx.matrix, [[math.cos(0.25), -1j * math.sin(0.25)], [-1j * math.sin(0.25), math.cos(0.25)]])
    adj = rx.adjoint()
    assert np.allclose(adj.matrix, rx.matrix.conj().T)
    s = qai.S(qubits=[0], do_queue=False)
    s.inverse()
    assert s.name == "S.inv" and np.allclose(s.matrix, np.diag([1, -1j]))
    s.inverse()
    assert s.name == "S"
    rot = qai.Rot(0.3, 0.4, 0.5, qubits=[0], do_queue=False)
    assert np.allclose(rot.adjoint().matrix, rot.matrix.conj().T)
    assert qai.CRX(0.2, qubits=[0, 1], do_queue=False).get_parameter_shift()[2][2] == pytest.approx(3 * math.pi / 2)
    assert qai.RY(0.2, qubits=[0], do_queue=False).get_parameter_shift() == [[0.5, 1.0, math.pi / 2],
                                                                            [-0.5, 1.0, -math.pi / 2]]
    with pytest.raises(ValueError):
        qai.CNOT(qubits=[0, 1], do_queue=False).get_parameter_shift()
    # tensor-valued parameters, including complex ones, are accepted at trace time
    assert np.allclose(qai.RY(torch.tensor([0.5], dtype=torch.complex64), qubits=[0], do_queue=False).matrix,
                       qai.RY(0.5, qubits=[0], do_queue=False).matrix, atol=1e-7)
    u = qai.Unitary(np.array([[0, 1], [1, 0]]), qubits=[0], do_queue=False)
    assert sorted(np.real(u.eigvals).tolist()) == [-1.0, 1.0]


def test_backend_construction_kwargs():
    params = (random.random(), random.random())
    circuit = qai.Circuit(_two, 2, *params)
    cc = circuit.compilecircuit(backend="pytorch")
    assert isinstance(cc, PyTorchBackend) and isinstance(cc, CompiledCircuit)
    assert cc.diff_method == "back_prop" and cc.backend == "pytorch" and cc.interface == "pytorch"
    cc = circuit.compilecircuit(backend="pytorch", diff_method="param_shift")
    assert cc.diff_method == "param_shift"
    with pytest.raises(ValueError):
        circuit.compilecircuit(backend="pytorch", interface="jax")
    with pytest.raises(ValueError, match="Error!!!! can not use contengra, opt_einsum and cyc at the same time!"):
        circuit.compilecircuit(backend="pytorch", use_cotengra=True, use_jdopttn=True)
    with pytest.raises(ValueError):
        circuit.compilecircuit(backend="nonsense")
    assert cc.operators == circuit.operators and cc.measurements == circuit.measurements
    # unknown differential method: bare Exception at call time, as in the reference
    bad = circuit.compilecircuit(backend="pytorch", diff_method="arbitrary")
    with pytest.raises(Exception):
        bad(torch.tensor([0.1]), torch.tensor([0.2]))
    with pytest.raises(NotImplementedError):
        circuit.compilecircuit(backend="pytorch", diff_method="finite_diff")(torch.tensor([0.1]), torch.tensor([0.2]))
    with pytest.raises(ValueError, match="must be type of pytorch tensor"):
        cc(0.1, 0.2)
    ps = circuit.compilecircuit(backend="pytorch", diff_method="param_shift")
    with pytest.raises(ValueError, match="same data type"):
        ps(torch.tensor([0.1], dtype=torch.float32), torch.tensor([0.2], dtype=torch.complex64))


def test_axes_and_permutation_lists():
    # test/test_compiled_circuit.py:88-105
    cc = qai.Circuit(_two, 2, 0.1, 0.2).compilecircuit(backend="pytorch")
    assert cc._axeslist == [([1], [1]), ([1], [0])]
    assert cc._permutationlist == [[1, 0], [0, 1]]


def test_tensor_network_index_structure():
    # index lists of gen_tensor_networks (tensor_network.py:850-1099): [new, old] per gate
    def cd(*p):
        qai.RY(p[0], qubits=[0])
        qai.RZ(p[1], qubits=[1])
        return [qai.expval(qai.PauliZ(qubits=[0])), qai.state()]
    cc = qai.Circuit(cd, 2, 0.1, 0.2).compilecircuit(backend="pytorch", tn_mode=True)
    ix0, out0, size0 = cc._cotengra_operands[0]
    assert ix0 == [['a'], ['b'], ['c', 'a'], ['d', 'b'], ['e', 'c'], ['f', 'd'], ['g', 'e'], ['g'], ['f']]
    assert out0 == [] and size0 == {s: 2 for s in "abcdefg"}
    ix1, out1, size1 = cc._cotengra_operands[1]
    assert ix1 == [['a'], ['b'], ['c', 'a'], ['d', 'b']]
    assert out1 == ['c', 'd'] and size1 == {s: 2 for s in "abcd"}
    with pytest.raises(ValueError):
        def cdi():
            qai.IintStateVector(np.array([1, 0, 0, 0]))
            return qai.state()
        qai.Circuit(cdi, 2).compilecircuit(backend="pytorch", tn_mode=True)


def test_trainable_parameter_bookkeeping():
    # test/test_compiled_circuit.py:143-211
    params = tuple(torch.tensor([random.random()], requires_grad=True) for _ in range(4))

    def cd(*p):
        qai.Rot(p[0], p[1], p[2], qubits=[0], trainable_params=[0, 2])
        qai.RZ(p[3], qubits=[1])
        return qai.expval(qai.PauliZ(qubits=[0]))
    cc = qai.Circuit(cd, 2, *params).compilecircuit(backend="pytorch")
    assert cc.gate_parameters == [[params[0], params[1], params[2]], [params[3]]]
    assert cc.parameters == list(params)
    assert cc.trainable_parameters == [params[0], params[2], params[3]]
    new = tuple(torch.tensor([random.random()]) for _ in range(3))
    cc._flatten_and_update(new)
    got = [[x.item() for x in g] for g in cc.gate_parameters]
    assert got == [[new[0].item(), params[1].item(), new[1].item()], [new[2].item()]]
    assert [x.item() for x in cc.trainable_parameters] == [x.item() for x in new]
    with pytest.raises(ValueError, match="number of parameters are not matched"):
        cc._flatten_and_update(new[:2])
