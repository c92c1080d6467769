"""``Circuit``: trace a user function into operators and measurements, then compile it.

Follows the reference front end (tedq/QInterpreter/circuits/circuit.py:38-98,
268-285): ``Circuit(circuitDef, num_qubits, *params)`` runs ``circuitDef``
inside a ``CircuitStorage`` and splits what was recorded into the optional
initial-state preparation, the gates and the measurements;
``compilecircuit(backend="pytorch", **kwargs)`` returns the compiled callable.
"""

import numpy as np

from .operators import GateBase
from .measurement import QuantumMeasurement
from .quantum_error import QuantumStorageError, QuantumCircuitError
from .storage import CircuitStorage


def generate_parameters(parameter_shapes):
    """Random parameters for each shape in ``parameter_shapes`` (circuit.py:407-417)."""
    return tuple((np.random.rand(*shape) + 0.01) * np.e / 1.77 for shape in parameter_shapes)


class Circuit:
    """A traced quantum circuit."""

    def __init__(self, func, num_qubits, *params, **kwargs):
        if isinstance(func, dict):
            self._func = None
            self._circuit = None
            self._operators = func["operators"]
            self._measurements = func["measurements"]
            self._init_state = func["init_state"]
            self._num_qubits = num_qubits
        else:
            self._func = func
            if not num_qubits:
                raise ValueError("Error in Circuit class, num_qubits cannot be None!")
            self._num_qubits = num_qubits
            parameter_shapes = kwargs.get("parameter_shapes", None)
            if parameter_shapes:
                params = generate_parameters(parameter_shapes)
            with CircuitStorage() as storage:
                self._func(*params)
            self._circuit = storage
            content = storage.storage_context
            first = content[0] if content else None
            self._init_state = first if getattr(first, "_is_preparation", False) else None
            self._operators = [op for op in content if isinstance(op, GateBase)]
            self._measurements = [op for op in content if isinstance(op, QuantumMeasurement)]
            if len(self._operators) + len(self._measurements) + bool(self._init_state) != len(content):
                raise QuantumStorageError("Error in Circuit class, unknown content in the storage_context!")

        if len(self._measurements) == 0:
            raise QuantumCircuitError("No measurement! please specify a quantum measurement!")
        max_qubit = self.maximum_qubit_num()
        if max_qubit + 1 > self._num_qubits:
            raise ValueError(
                f"Input number of qubits is not large enough! "
                f"Maximum qubit number of operators is {max_qubit + 1}")

    def maximum_qubit_num(self):
        """Largest qubit index used by any gate, measured qubit or observable."""
        max_qubit = 0
        for op in self._operators:
            max_qubit = max(max_qubit, max(op.qubits))
        for meas in self._measurements:
            if meas.qubits is not None:
                max_qubit = max(max_qubit, max(meas.qubits))
            if meas.obs is not None:
                observables = meas.obs if isinstance(meas.obs, list) else [meas.obs]
                for ob in observables:
                    max_qubit = max(max_qubit, max(ob.qubits))
        return max_qubit

    @property
    def operators(self):
        return self._operators

    @property
    def num_qubits(self):
        return self._num_qubits

    @property
    def measurements(self):
        return self._measurements

    @property
    def init_state(self):
        return self._init_state

    def compilecircuit(self, backend=None, **kwargs):
        """Compile for ``backend``; only ``"pytorch"`` (the MI355X engine) is provided here."""
        if backend == "pytorch":
            from .backends.pytorch_backend import PyTorchBackend  # pylint: disable=import-outside-toplevel
            return PyTorchBackend(backend, self, **kwargs)
        if backend in ("jax", "pytorch_QUDIO", "IBMQ_hardware", "Quafu_hardware"):
            raise ValueError(f"{backend}: backend is not part of the MI355X engine")
        raise ValueError(f"{backend}: unknown backend input")

    def __str__(self):
        lines = ["# operators"]
        lines += [f"qai.{gate.name}(qubits={gate.qubits})" for gate in self.operators]
        lines += ["", "", "# measurements"]
        for meas in self.measurements:
            if meas.obs is None:
                lines.append(f"qai.measurement.{meas.return_type.value}(qubits={meas.qubits})")
            else:
                observables = meas.obs if isinstance(meas.obs, list) else [meas.obs]
                names = ", ".join(f"qai.{ob.name}(qubits={ob.qubits})" for ob in observables)
                lines.append(f"qai.measurement.expval({names})")
        return "\n".join(lines) + "\n"
