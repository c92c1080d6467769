"""Lower a compiled circuit's gate list to engine units with parameter sources.

Trainable parameters get flat column indices in the order the reference's
``_update_parameters`` assigns call arguments (tedq/backends/compiled_circuit.py:522-547):
gate by gate, and inside a gate in the order of its ``trainable_params``
list.  Non-trainable parameters keep their trace-time value and become fp64
constants of the table program.
"""

from ..operators import as_float
from .planner import gate_to_units


def lower_circuit(cc):
    """Return (n_qubits, units, n_trainable, const_pool) for a ``CompiledCircuit``."""
    n = cc._num_qubits  # pylint: disable=protected-access
    units, pool = [], []
    flat = 0
    for idx, op in cc._operators.items():  # pylint: disable=protected-access
        trainable = cc._trainableparams[idx]  # pylint: disable=protected-access
        srcs = [None] * op.num_params
        for i, pos in enumerate(trainable):
            srcs[pos] = flat + i
        flat += len(trainable)
        for pos in range(op.num_params):
            if srcs[pos] is None:
                value = as_float(cc._operatorparams[idx][pos])  # pylint: disable=protected-access
                pool.append(complex(value, 0.0))
                srcs[pos] = -1 - (len(pool) - 1)
        units.extend(gate_to_units(n, op, tuple(srcs), pool))
    return n, units, flat, pool
