"""Circuit families used by the benchmark and the parity tests.

* ``hea``: hardware-efficient ansatz, per layer RY(theta) RZ(phi) on every
  qubit then CNOT(0,1), CNOT(1,2), ..., CNOT(n-2,n-1); parameters ordered
  layer -> qubit -> (theta, phi).
* ``disordered_chain``: Trotterised 1-D disordered spin chain (the reference's
  many-body-localisation example, examples/comparison/
  Many_body_Localization_1D_JT.py:56-75, with its exchange terms as exact
  two-qubit unitaries): Neel state (X on odd qubits), then per Trotter step
  RZ(2 h_i dt) on every qubit and, on every bond (i, i+1) in order,
  exp(-i J dt Z Z) and exp(-i J dt (X X + Y Y)).  The RZ angles are the
  trainable parameters (step -> qubit); ``chain_rows`` maps disorder fields
  h[B, n] to parameter rows.
* ``lattice``: L x L grid, per layer RX(theta) RZ(phi) on every qubit, then CZ
  on all horizontal bonds, then on all vertical bonds; one <Z_i> per qubit.
"""

import numpy as np
import scipy.linalg as sl

from . import measurement as _m
from . import operators as _g
from .circuit import Circuit

_X = np.array([[0, 1], [1, 0]], dtype=np.complex128)
_Y = np.array([[0, -1j], [1j, 0]], dtype=np.complex128)
_Z = np.array([[1, 0], [0, -1]], dtype=np.complex128)


def hea(n, layers, outputs="all", **compile_kwargs):
    """Compiled HEA; outputs "all" (<Z_i> for every qubit) or "z0" (<Z_0>)."""

    def circuit_def(*p):
        it = iter(p)
        for _ in range(layers):
            for q in range(n):
                _g.RY(next(it), qubits=[q])
                _g.RZ(next(it), qubits=[q])
            for q in range(n - 1):
                _g.CNOT(qubits=[q, q + 1])
        if outputs == "z0":
            return _m.expval(_g.PauliZ(qubits=[0]))
        return [_m.expval(_g.PauliZ(qubits=[q])) for q in range(n)]

    return Circuit(circuit_def, n, *([0.1] * (2 * n * layers))).compilecircuit(backend="pytorch", **compile_kwargs)


def exchange_unitaries(J=1.0, dt=0.1):
    """exp(-i J dt Z Z) and exp(-i J dt (X X + Y Y)) as 4x4 matrices (first qubit most significant)."""
    zz = sl.expm(-1j * J * dt * np.kron(_Z, _Z))
    xy = sl.expm(-1j * J * dt * (np.kron(_X, _X) + np.kron(_Y, _Y)))
    return zz, xy


def disordered_chain(n=40, steps=8, observe=20, J=1.0, dt=0.1, **compile_kwargs):
    """Compiled chain measuring <Z_observe>; parameters are the RZ angles (steps x n)."""
    zz, xy = exchange_unitaries(J, dt)

    def circuit_def(*p):
        it = iter(p)
        for q in range(1, n, 2):
            _g.PauliX(qubits=[q])
        for _ in range(steps):
            for q in range(n):
                _g.RZ(next(it), qubits=[q])
            for q in range(n - 1):
                _g.Unitary(zz, qubits=[q, q + 1])
                _g.Unitary(xy, qubits=[q, q + 1])
        return _m.expval(_g.PauliZ(qubits=[observe]))

    return Circuit(circuit_def, n, *([0.1] * (n * steps))).compilecircuit(backend="pytorch", **compile_kwargs)


def chain_rows(h, steps=8, dt=0.1):
    """Parameter rows [B, steps * n] (RZ angle 2 h_i dt, repeated every step) from fields h [B, n]."""
    h = np.asarray(h, dtype=np.float64)
    return np.tile(2.0 * h * dt, (1, steps)).astype(np.float32)


def lattice(L=5, layers=6, observe=None, **compile_kwargs):
    """Compiled L x L lattice circuit with one <Z_i> per qubit (L*L outputs), or per qubit of
    ``observe`` when given."""
    n = L * L
    observed = list(range(n)) if observe is None else list(observe)

    def circuit_def(*p):
        it = iter(p)
        for _ in range(layers):
            for q in range(n):
                _g.RX(next(it), qubits=[q])
                _g.RZ(next(it), qubits=[q])
            for r in range(L):
                for c in range(L - 1):
                    _g.CZ(qubits=[r * L + c, r * L + c + 1])
            for r in range(L - 1):
                for c in range(L):
                    _g.CZ(qubits=[r * L + c, (r + 1) * L + c])
        return [_m.expval(_g.PauliZ(qubits=[q])) for q in observed]

    return Circuit(circuit_def, n, *([0.1] * (2 * n * layers))).compilecircuit(backend="pytorch", **compile_kwargs)
