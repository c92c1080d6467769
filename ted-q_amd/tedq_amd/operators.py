"""Gate and observable records captured while tracing a circuit.

The public surface follows the reference's operator classes
(tedq/QInterpreter/operators/ops_abc.py:28-380 and qubit.py:44-1786):
constructor ``Gate(*params, qubits=[...], do_queue=True, trainable_params=...)``,
the ``matrix`` / ``parameters`` / ``qubits`` / ``trainable_params`` /
``instance_id`` properties, parameter-shift recipes, ``adjoint`` and
``inverse``.  Matrices are the textbook definitions in the computational
basis with the first listed qubit as the most significant index (control
first for controlled gates, qubit.py:644).

Every gate class additionally carries ``gate_code``: the integer the device
table builder uses to rebuild the gate matrix from its parameters on the GPU.
"""

# pylint: disable=invalid-name

import abc
import cmath
import itertools
import math

import numpy as np

from .global_variables import GlobalVariables

INV_SQRT2 = 1.0 / math.sqrt(2.0)
PI = math.pi

# Four-term shift rule for controlled rotations (qubit.py:32-42): the
# derivative of f at phi is sum_k c_k f(phi + s_k) with these (c, a, s) rows.
_M1 = INV_SQRT2 * (math.sqrt(2.0) + 1.0) / 4.0
_M2 = INV_SQRT2 * (math.sqrt(2.0) - 1.0) / 4.0
CONTROL_GRAD_RECIPE = [
    [_M1, 1.0, PI / 2.0],
    [-_M1, 1.0, -PI / 2.0],
    [-_M2, 1.0, 3.0 * PI / 2.0],
    [_M2, 1.0, -3.0 * PI / 2.0],
]

# Integer codes shared with csrc/sv_tables.hip (kGate* constants there).
GATE_CODES = {
    "I": 0, "Hadamard": 1, "PauliX": 2, "PauliY": 3, "PauliZ": 4, "S": 5, "T": 6,
    "SX": 7, "CNOT": 8, "CZ": 9, "CY": 10, "SWAP": 11, "CSWAP": 12, "Toffoli": 13,
    "RX": 14, "RY": 15, "RZ": 16, "Rot": 17, "PhaseShift": 18,
    "ControlledPhaseShift": 19, "CRX": 20, "CRY": 21, "CRZ": 22, "Unitary": 23,
}


def as_float(value):
    """Real value of a gate parameter given as a python/numpy number or a one-element tensor."""
    if hasattr(value, "detach"):
        value = value.detach()
        if value.is_complex():
            value = value.real
        return float(value.reshape(-1)[0].item())
    if isinstance(value, complex):
        return float(value.real)
    arr = np.asarray(value)
    if np.iscomplexobj(arr):
        arr = arr.real
    return float(arr.reshape(-1)[0])


def complex_conjugate(ts):
    """Conjugate transpose of a square operator given as a matrix or a [2]*2k tensor."""
    ts = np.asarray(ts)
    shape = ts.shape
    size = int(round(math.sqrt(int(np.prod(shape)))))
    return ts.reshape(size, size).conj().T.reshape(shape)


class OperatorBase(abc.ABC):
    """Base class of every traced operator (gates, observables, state preparations)."""

    _count = itertools.count()
    _num_params = 0
    _num_qubits = 1
    gate_code = -1

    def __init__(self, *params, qubits, do_queue=True, **kwargs):
        self._name = self.__class__.__name__
        if len(params) != self.num_params:
            raise ValueError(
                f"{self._name}: # of parameters is not matched! expected {self.num_params}"
                f"parameters, but got {len(params)}.")
        if len(qubits) != self.num_qubits:
            raise ValueError(
                f"{self._name}: # of qubits this operator applied on is not matched! expected "
                f"{self.num_qubits} qubits, but got {len(qubits)}")
        self._instance_id = next(self._count)
        if do_queue:
            warehouse = GlobalVariables.get_value("global_deque")
            if not warehouse:
                raise ValueError("There's no global_deque for storing information!")
            if not warehouse[-1]:
                raise ValueError("There's no StorageBase for storing information!")
            warehouse[-1]._append(self)  # pylint: disable=protected-access
        self._do_queue = do_queue
        self._qubits = list(qubits)
        self._params = list(params)
        self._trainable_params = kwargs.pop("trainable_params", list(range(self.num_params)))
        self._is_preparation = kwargs.pop("is_preparation", False)
        self._expected_index = kwargs.pop("expected_index", 0)
        self._matrix = self.get_matrix()

    def get_matrix(self):
        raise NotImplementedError

    @property
    def matrix(self):
        return self._matrix

    @property
    def eigvals(self):
        return np.linalg.eigvals(self.matrix)

    @property
    def num_params(self):
        return self._num_params

    @property
    def num_qubits(self):
        return self._num_qubits

    @property
    def name(self):
        return self._name

    @property
    def qubits(self):
        return self._qubits

    @property
    def parameters(self):
        return self._params

    @property
    def instance_id(self):
        return self._instance_id

    @property
    def trainable_params(self):
        return self._trainable_params

    @property
    def is_preparation(self):
        return self._is_preparation

    def _fval(self, i):
        return as_float(self._params[i])


class GateBase(OperatorBase):
    """A unitary gate; parametrised gates support the parameter-shift rule."""

    def __init__(self, *params, qubits, do_queue=True, **kwargs):
        self._inverse = False
        super().__init__(*params, qubits=qubits, do_queue=do_queue, **kwargs)

    @property
    def grad_style(self):
        return "param-shift" if self.num_params > 0 else None

    def get_parameter_shift(self):
        """Two-term rule (ops_abc.py:261-280): df = 0.5 [f(p + pi/2) - f(p - pi/2)]."""
        shift = PI / 2.0
        m = 0.5 / math.sin(shift)
        if self.grad_style == "param-shift":
            return [[m, 1.0, shift], [-m, 1.0, -shift]]
        raise ValueError(f"{self.name} : This quantum gate do not support parameter-shift method")

    @property
    def is_inverse(self):
        return self._inverse

    def inverse(self):
        """Toggle this instance into its inverse: name gains/loses ``.inv``, matrix is daggered."""
        if self._inverse:
            self._inverse = False
            self._name = self._name[:-4]
        else:
            self._inverse = True
            self._name = self._name + ".inv"
        self._matrix = self._matrix.conj().T
        return self

    def adjoint(self, do_queue=False):
        raise NotImplementedError


class ObservableBase(OperatorBase):
    """An operator that can be measured with ``expval``."""

    return_type = None

    def diagonalizing_gates(self, do_queue=False):
        raise NotImplementedError


# ---------------------------------------------------------------------------
# Fixed single-qubit gates / observables
# ---------------------------------------------------------------------------

class Hadamard(ObservableBase, GateBase):
    """Hadamard gate, (X + Z)/sqrt(2)."""
    _num_params = 0
    _num_qubits = 1
    gate_code = GATE_CODES["Hadamard"]

    def get_matrix(self):
        return np.array([[INV_SQRT2, INV_SQRT2], [INV_SQRT2, -INV_SQRT2]])

    def diagonalizing_gates(self, do_queue=False):
        return [RY(-PI / 4, qubits=self.qubits, do_queue=do_queue)]

    def adjoint(self, do_queue=False):
        return Hadamard(qubits=self.qubits, do_queue=do_queue)


class PauliX(ObservableBase, GateBase):
    """Pauli X (bit flip)."""
    _num_params = 0
    _num_qubits = 1
    gate_code = GATE_CODES["PauliX"]

    def get_matrix(self):
        return np.array([[0, 1], [1, 0]])

    def diagonalizing_gates(self, do_queue=False):
        return [Hadamard(qubits=self.qubits, do_queue=do_queue)]

    def adjoint(self, do_queue=False):
        return PauliX(qubits=self.qubits, do_queue=do_queue)


class PauliY(ObservableBase, GateBase):
    """Pauli Y."""
    _num_params = 0
    _num_qubits = 1
    gate_code = GATE_CODES["PauliY"]

    def get_matrix(self):
        return np.array([[0, -1j], [1j, 0]])

    def diagonalizing_gates(self, do_queue=False):
        return [PauliZ(qubits=self.qubits, do_queue=do_queue),
                S(qubits=self.qubits, do_queue=do_queue),
                Hadamard(qubits=self.qubits, do_queue=do_queue)]

    def adjoint(self, do_queue=False):
        return PauliY(qubits=self.qubits, do_queue=do_queue)


class PauliZ(ObservableBase, GateBase):
    """Pauli Z (phase flip)."""
    _num_params = 0
    _num_qubits = 1
    gate_code = GATE_CODES["PauliZ"]

    def get_matrix(self):
        return np.array([[1, 0], [0, -1]])

    def diagonalizing_gates(self, do_queue=False):
        return [I(qubits=self.qubits, do_queue=do_queue)]

    def adjoint(self, do_queue=False):
        return PauliZ(qubits=self.qubits, do_queue=do_queue)


class I(ObservableBase, GateBase):
    """Identity."""
    _num_params = 0
    _num_qubits = 1
    gate_code = GATE_CODES["I"]

    def get_matrix(self):
        return np.array([[1, 0], [0, 1]])

    def diagonalizing_gates(self, do_queue=False):
        return [I(qubits=self.qubits, do_queue=do_queue)]

    def adjoint(self, do_queue=False):
        return I(qubits=self.qubits, do_queue=do_queue)


class S(GateBase):
    """Phase gate diag(1, i)."""
    _num_params = 0
    _num_qubits = 1
    gate_code = GATE_CODES["S"]

    def get_matrix(self):
        return np.array([[1, 0], [0, 1j]])

    def adjoint(self, do_queue=False):
        return S(qubits=self.qubits, do_queue=do_queue).inverse()


class T(GateBase):
    """T gate diag(1, e^{i pi/4})."""
    _num_params = 0
    _num_qubits = 1
    gate_code = GATE_CODES["T"]

    def get_matrix(self):
        return np.array([[1, 0], [0, cmath.exp(1j * PI / 4)]])

    def adjoint(self, do_queue=False):
        return T(qubits=self.qubits, do_queue=do_queue).inverse()


class SX(GateBase):
    """Square root of X."""
    _num_params = 0
    _num_qubits = 1
    gate_code = GATE_CODES["SX"]

    def get_matrix(self):
        return 0.5 * np.array([[1 + 1j, 1 - 1j], [1 - 1j, 1 + 1j]])

    def adjoint(self, do_queue=False):
        return SX(qubits=self.qubits, do_queue=do_queue).inverse()


# ---------------------------------------------------------------------------
# Fixed multi-qubit gates (first listed qubit is the control / most significant)
# ---------------------------------------------------------------------------

class CNOT(GateBase):
    """Controlled X; qubits = [control, target]."""
    _num_params = 0
    _num_qubits = 2
    gate_code = GATE_CODES["CNOT"]

    def get_matrix(self):
        return np.array([[1, 0, 0, 0], [0, 1, 0, 0], [0, 0, 0, 1], [0, 0, 1, 0]])

    def adjoint(self, do_queue=False):
        return CNOT(qubits=self.qubits, do_queue=do_queue)


class CZ(GateBase):
    """Controlled Z."""
    _num_params = 0
    _num_qubits = 2
    gate_code = GATE_CODES["CZ"]

    def get_matrix(self):
        return np.array([[1, 0, 0, 0], [0, 1, 0, 0], [0, 0, 1, 0], [0, 0, 0, -1]])

    def adjoint(self, do_queue=False):
        return CZ(qubits=self.qubits, do_queue=do_queue)


class CY(GateBase):
    """Controlled Y."""
    _num_params = 0
    _num_qubits = 2
    gate_code = GATE_CODES["CY"]

    def get_matrix(self):
        return np.array([[1, 0, 0, 0], [0, 1, 0, 0], [0, 0, 0, -1j], [0, 0, 1j, 0]])

    def adjoint(self, do_queue=False):
        return CY(qubits=self.qubits, do_queue=do_queue)


class SWAP(GateBase):
    """Exchange of two qubits."""
    _num_params = 0
    _num_qubits = 2
    gate_code = GATE_CODES["SWAP"]

    def get_matrix(self):
        return np.array([[1, 0, 0, 0], [0, 0, 1, 0], [0, 1, 0, 0], [0, 0, 0, 1]])

    def adjoint(self, do_queue=False):
        return SWAP(qubits=self.qubits, do_queue=do_queue)


class CSWAP(GateBase):
    """Controlled SWAP (Fredkin); qubits = [control, a, b]."""
    _num_params = 0
    _num_qubits = 3
    gate_code = GATE_CODES["CSWAP"]

    def get_matrix(self):
        m = np.eye(8)
        m[[5, 6]] = m[[6, 5]]
        return m

    def adjoint(self, do_queue=False):
        return CSWAP(qubits=self.qubits, do_queue=do_queue)


class Toffoli(GateBase):
    """Doubly controlled X; qubits = [control, control, target]."""
    _num_params = 0
    _num_qubits = 3
    gate_code = GATE_CODES["Toffoli"]

    def get_matrix(self):
        m = np.eye(8)
        m[[6, 7]] = m[[7, 6]]
        return m

    def adjoint(self, do_queue=False):
        return Toffoli(qubits=self.qubits, do_queue=do_queue)


# ---------------------------------------------------------------------------
# Parametrised gates
# ---------------------------------------------------------------------------

def _rx(t):
    c, s = math.cos(t / 2), math.sin(-t / 2)
    return np.array([[c, 1j * s], [1j * s, c]])


def _ry(t):
    c, s = math.cos(t / 2), math.sin(t / 2)
    return np.array([[c, -s], [s, c]])


def _rz(t):
    p = cmath.exp(-0.5j * t)
    return np.array([[p, 0], [0, p.conjugate()]])


def _controlled(block):
    m = np.eye(4, dtype=complex)
    m[2:, 2:] = block
    return m


class RX(GateBase):
    """exp(-i phi X / 2)."""
    _num_params = 1
    _num_qubits = 1
    gate_code = GATE_CODES["RX"]

    def get_matrix(self):
        return _rx(self._fval(0))

    def adjoint(self, do_queue=False):
        return RX(-self.parameters[0], qubits=self.qubits, do_queue=do_queue)

    @property
    def generator(self):
        return [-0.5, PauliX(qubits=self.qubits, do_queue=False)]


class RY(GateBase):
    """exp(-i phi Y / 2)."""
    _num_params = 1
    _num_qubits = 1
    gate_code = GATE_CODES["RY"]

    def get_matrix(self):
        return _ry(self._fval(0))

    def adjoint(self, do_queue=False):
        return RY(-self.parameters[0], qubits=self.qubits, do_queue=do_queue)

    @property
    def generator(self):
        return [-0.5, PauliY(qubits=self.qubits, do_queue=False)]


class RZ(GateBase):
    """exp(-i phi Z / 2)."""
    _num_params = 1
    _num_qubits = 1
    gate_code = GATE_CODES["RZ"]

    def get_matrix(self):
        return _rz(self._fval(0))

    def adjoint(self, do_queue=False):
        return RZ(-self.parameters[0], qubits=self.qubits, do_queue=do_queue)

    @property
    def generator(self):
        return [-0.5, PauliZ(qubits=self.qubits, do_queue=False)]


class Rot(GateBase):
    """RZ(omega) RY(theta) RZ(phi) with parameters (phi, theta, omega) (qubit.py:1169-1189)."""
    _num_params = 3
    _num_qubits = 1
    gate_code = GATE_CODES["Rot"]

    def get_matrix(self):
        phi, theta, omega = self._fval(0), self._fval(1), self._fval(2)
        c, s = math.cos(theta / 2), math.sin(theta / 2)
        return np.array([
            [cmath.exp(-0.5j * (phi + omega)) * c, -cmath.exp(0.5j * (phi - omega)) * s],
            [cmath.exp(-0.5j * (phi - omega)) * s, cmath.exp(0.5j * (phi + omega)) * c],
        ])

    def adjoint(self, do_queue=False):
        phi, theta, omega = self.parameters
        return Rot(-omega, -theta, -phi, qubits=self.qubits, do_queue=do_queue)


class PhaseShift(GateBase):
    """diag(1, e^{i phi})."""
    _num_params = 1
    _num_qubits = 1
    gate_code = GATE_CODES["PhaseShift"]

    def get_matrix(self):
        return np.array([[1, 0], [0, cmath.exp(1j * self._fval(0))]])

    def adjoint(self, do_queue=False):
        return PhaseShift(-self.parameters[0], qubits=self.qubits, do_queue=do_queue)

    @property
    def generator(self):
        return [1, Unitary(np.array([[0.0, 0.0], [0.0, 1.0]], dtype=np.complex64),
                           qubits=self.qubits, do_queue=False)]


class ControlledPhaseShift(GateBase):
    """diag(1, 1, 1, e^{i phi}); qubits = [control, target]."""
    _num_params = 1
    _num_qubits = 2
    gate_code = GATE_CODES["ControlledPhaseShift"]

    def get_matrix(self):
        return np.diag([1, 1, 1, cmath.exp(1j * self._fval(0))])

    def adjoint(self, do_queue=False):
        return ControlledPhaseShift(-self.parameters[0], qubits=self.qubits, do_queue=do_queue)

    @property
    def generator(self):
        g = np.zeros((4, 4), dtype=np.complex64)
        g[3, 3] = 1.0
        return [1, Unitary(g.reshape([2, 2, 2, 2]), qubits=self.qubits, do_queue=False)]


class _ControlledRotation(GateBase):
    _num_params = 1
    _num_qubits = 2
    _block = None

    def get_parameter_shift(self):
        if self.grad_style == "param-shift":
            return CONTROL_GRAD_RECIPE
        raise ValueError(f"{self.name} : This quantum gate do not support parameter-shift method")

    def get_matrix(self):
        return _controlled(type(self)._block(self._fval(0)))

    def adjoint(self, do_queue=False):
        return type(self)(-self.parameters[0], qubits=self.qubits, do_queue=do_queue)


class CRX(_ControlledRotation):
    """Controlled RX; qubits = [control, target]."""
    gate_code = GATE_CODES["CRX"]
    _block = staticmethod(_rx)

    @property
    def generator(self):
        g = np.zeros((4, 4), dtype=np.complex64)
        g[2, 3] = g[3, 2] = 1.0
        return [-0.5, Unitary(g.reshape([2, 2, 2, 2]), qubits=self.qubits, do_queue=False)]


class CRY(_ControlledRotation):
    """Controlled RY; qubits = [control, target]."""
    gate_code = GATE_CODES["CRY"]
    _block = staticmethod(_ry)

    @property
    def generator(self):
        g = np.zeros((4, 4), dtype=np.complex64)
        g[2, 3], g[3, 2] = -1j, 1j
        return [-0.5, Unitary(g.reshape([2, 2, 2, 2]), qubits=self.qubits, do_queue=False)]


class CRZ(_ControlledRotation):
    """Controlled RZ; qubits = [control, target]."""
    gate_code = GATE_CODES["CRZ"]
    _block = staticmethod(_rz)

    @property
    def generator(self):
        g = np.zeros((4, 4), dtype=np.complex64)
        g[2, 2], g[3, 3] = 1.0, -1.0
        return [-0.5, Unitary(g.reshape([2, 2, 2, 2]), qubits=self.qubits, do_queue=False)]


class Unitary(GateBase, ObservableBase):
    """User-supplied operator on ``len(qubits)`` qubits.

    ``matrix`` may be given as a 2^k x 2^k array or as a [2]*2k tensor.  The
    reference constructor fails on an undefined name (qubit.py:1719); here the
    eigen-decomposition is computed lazily so any square operator is accepted.
    """
    _num_params = 0
    _num_qubits = None
    gate_code = GATE_CODES["Unitary"]

    def __init__(self, matrix, qubits, do_queue=True, **kwargs):
        self._num_qubits = len(qubits)
        dim = 2 ** len(qubits)
        self._user_defined_matrix = np.asarray(matrix).reshape(dim, dim)
        self._eig = None
        super().__init__(qubits=qubits, do_queue=do_queue, **kwargs)

    def get_matrix(self):
        return self._user_defined_matrix

    def _eigen(self):
        if self._eig is None:
            vals, vecs = np.linalg.eig(self._user_defined_matrix)
            self._eig = (vals, np.linalg.inv(vecs))
        return self._eig

    @property
    def eigvals(self):
        return self._eigen()[0]

    def adjoint(self, do_queue=False):
        return Unitary(complex_conjugate(self.matrix), qubits=self.qubits, do_queue=do_queue)

    def diagonalizing_gates(self, do_queue=False):
        return Unitary(self._eigen()[1], qubits=self.qubits, do_queue=do_queue)


ops = {
    "Hadamard", "PauliX", "PauliY", "PauliZ", "I", "S", "T", "SX", "CNOT", "CZ", "CY",
    "SWAP", "CSWAP", "Toffoli", "RX", "RY", "RZ", "Rot", "PhaseShift",
    "ControlledPhaseShift", "CRX", "CRY", "CRZ", "Unitary",
}
obs = {"Hadamard", "PauliX", "PauliY", "PauliZ"}

__all__ = sorted(ops | obs) + ["OperatorBase", "GateBase", "ObservableBase", "CONTROL_GRAD_RECIPE"]
