"""tedq_amd: MI355X-native execution engine for TeD-Q compiled circuits.

Drop-in for the reference front end and PyTorch backend::

    import tedq_amd as qai
    def circuitDef(*params):
        qai.RX(params[0], qubits=[0])
        qai.RY(params[1], qubits=[0])
        return qai.expval(qai.PauliZ(qubits=[0]))
    circuit = qai.Circuit(circuitDef, 1, 0.54, 0.12)
    compiled = circuit.compilecircuit(backend="pytorch")
    compiled(torch.tensor([0.54], requires_grad=True), torch.tensor([0.12], requires_grad=True))

State-vector and tensor-network execution run in hand-written gfx950 HIP
kernels (``_lib/libtedq_hip.so``); there is no CPU execution path.
"""

from .global_variables import GlobalVariables
from .quantum_error import QuantumCircuitError, QuantumValueError, QuantumStorageError
from .operators import (Hadamard, PauliX, PauliY, PauliZ, I, S, T, SX, CNOT, CZ, CY, SWAP, CSWAP, Toffoli, RX, RY,
                        RZ, Rot, PhaseShift, ControlledPhaseShift, CRX, CRY, CRZ, Unitary, OperatorBase, GateBase,
                        ObservableBase)
from . import measurement
from .measurement import expval, probs, state, var, sample
from .prepared_state import IintStateVector
from .storage import StorageBase, CircuitStorage
from .circuit import Circuit
from .tools import dec_to_bin, rescale_state
from .backends import CompiledCircuit, PyTorchBackend

__version__ = "0.1.0"
