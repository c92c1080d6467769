"""Measurement records placed at the end of a traced circuit.

Same contract as the reference (tedq/QInterpreter/operators/measurement.py:29-266):
``expval(obs)`` takes the observable that was just traced (or a list of
them, for a product observable) out of the active storage and puts a
``QuantumMeasurement`` in its place; ``probs(qubits=None)`` and ``state()``
append a measurement directly.
"""

from enum import Enum

from .global_variables import GlobalVariables
from .quantum_error import QuantumCircuitError, QuantumValueError
from .operators import ObservableBase


class MeasurementReturnTypes(Enum):
    """Kind of result a measurement produces."""

    Sample = "sample"
    Variance = "var"
    Expectation = "expval"
    Probability = "probs"
    State = "state"

    def __repr__(self):
        return str(self.value)


Sample = MeasurementReturnTypes.Sample
Variance = MeasurementReturnTypes.Variance
Expectation = MeasurementReturnTypes.Expectation
Probability = MeasurementReturnTypes.Probability
State = MeasurementReturnTypes.State


class QuantumMeasurement:
    """One measurement: expectation value, probabilities, or the final state."""

    _is_preparation = False

    @property
    def is_preparation(self):
        return False

    def __init__(self, return_type, obs=None, qubits=None, do_queue=True, after_state=False):
        self.return_type = return_type
        self.obs = obs
        self.do_queue = do_queue
        self.after_state = after_state
        if qubits is not None and obs is not None:
            raise ValueError("If an observable is provied, the qubit(s) can not be specified!")
        self.qubits = qubits
        if self.do_queue:
            self.circuit_queue()

    def circuit_queue(self):
        """Swap the traced observable(s) for this measurement in the active storage."""
        warehouse = GlobalVariables.get_value("global_deque")
        if not warehouse:
            raise QuantumCircuitError(
                "No active warehouse, tedq software is problematic, need to check 'global_variables.py'")
        storage = warehouse[-1]
        if self.obs is None:
            storage._append(self)  # pylint: disable=protected-access
            return
        observables = self.obs if isinstance(self.obs, list) else [self.obs]
        for ob in reversed(observables):
            last = storage.storage_context[-1]
            if last.instance_id != ob.instance_id:
                raise ValueError("Last content operator should be the same as 'obs'!")
            storage.remove(ob)
        storage.append(self)


def expval(observable, do_queue=True):
    """Expectation value of an observable (or of the product of a list of observables)."""
    observables = observable if isinstance(observable, list) else [observable]
    for ob in observables:
        if not isinstance(ob, ObservableBase):
            raise QuantumValueError(
                f"{getattr(ob, 'name', ob)} is not a subclass of ObservableBase: cannot be used with expval")
    return QuantumMeasurement(Expectation, obs=observable, do_queue=do_queue)


def var(observable, do_queue=True):
    """Variance measurement (not provided by the reference either)."""
    raise NotImplementedError


def sample(observable, num_shots, do_queue=True):
    """Sampling measurement (not provided by the reference either)."""
    raise NotImplementedError


def probs(qubits=None, do_queue=True, after_state=False):
    """Computational-basis probabilities of all qubits, or marginals over ``qubits``."""
    return QuantumMeasurement(Probability, qubits=qubits, do_queue=do_queue, after_state=after_state)


def state(do_queue=True):
    """The final state vector as a complex64 [2]*n tensor."""
    return QuantumMeasurement(State, do_queue=do_queue)
