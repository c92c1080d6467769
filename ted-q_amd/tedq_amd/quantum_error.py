"""Error types raised by circuit construction (tedq/quantum_error.py:1-7)."""


class QuantumCircuitError(Exception):
    """The quantum circuit is malformed (e.g. it has no measurement)."""


class QuantumValueError(Exception):
    """A value given to a circuit element is not usable (e.g. a non-observable in expval)."""


class QuantumStorageError(Exception):
    """The tracing storage holds something that is neither a gate, a measurement nor a preparation."""
