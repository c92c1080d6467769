"""User-supplied initial state (tedq/QInterpreter/operators/prepared_state.py:29-57).

Traced first in ``circuitDef``; ``Circuit`` recognises it through
``is_preparation`` and the backend starts propagation from this vector
instead of |0...0>.  The class name keeps the reference spelling.
"""

import numpy as np

from .operators import OperatorBase


class IintStateVector(OperatorBase):
    """Initial state vector with 2^n amplitudes (any shape that flattens to 2^n)."""

    _num_params = 0
    _num_qubits = 0

    def __init__(self, matrix, qubits=(), do_queue=True, **kwargs):
        self._user_defined_matrix = np.asarray(matrix)
        super().__init__(qubits=list(qubits), do_queue=do_queue, is_preparation=True, **kwargs)

    def get_matrix(self):
        return self._user_defined_matrix
