"""CPU oracle for the state-vector path — TEST INFRASTRUCTURE ONLY.

A restatement of the reference PyTorch backend's state-vector algorithm, used
by tests/, by ``__graft_entry__.smoke()`` and as the CPU baseline in
``bench.py``.  Nothing in the product package may import it.

It follows, step for step:
* gate matrices: tedq/QInterpreter/operators/qubit.py (RX :990-997, RY
  :1048-1055, RZ :1106-1112, Rot :1169-1189, PhaseShift :1253-1257,
  ControlledPhaseShift :1327-1334, CRX :1429-1436, CRY :1531-1538,
  CRZ :1631-1642, fixed gates :68-940);
* per-gate tensordot axes and inverse permutations:
  tedq/backends/compiled_circuit.py:126-202;
* the propagation loop tensordot(gate, state, axes).permute(perms):
  tedq/backends/pytorch_backend.py:358-380, initial |0..0>: :513-522;
* measurements: tedq/backends/pytorch_backend.py:393-498;
* parameter-shift recipes: tedq/QInterpreter/operators/ops_abc.py:261-280 and
  qubit.py:32-42 (controlled rotations).

``dtype`` selects complex128 (the parity oracle) or complex64 (the
reference's own precision, used for the CPU baseline).  Autograd through the
loop gives the reference's ``back_prop`` gradients.

A gate is a tuple ``(name, qubits, params)`` with ``params`` a list of real
0-d tensors or floats; an observable is ``(name, qubits)`` or a 2^k x 2^k
matrix with qubits; a measurement is ``("expval", obs_or_list)``,
``("probs", qubits_or_None)`` or ``("state",)``.
"""

import math

import numpy as np
import torch

_H = 1.0 / math.sqrt(2.0)
FIXED = {
    "I": [[1, 0], [0, 1]],
    "Hadamard": [[_H, _H], [_H, -_H]],
    "PauliX": [[0, 1], [1, 0]],
    "PauliY": [[0, -1j], [1j, 0]],
    "PauliZ": [[1, 0], [0, -1]],
    "S": [[1, 0], [0, 1j]],
    "T": [[1, 0], [0, complex(math.cos(math.pi / 4), math.sin(math.pi / 4))]],
    "SX": [[0.5 + 0.5j, 0.5 - 0.5j], [0.5 - 0.5j, 0.5 + 0.5j]],
    "CNOT": [[1, 0, 0, 0], [0, 1, 0, 0], [0, 0, 0, 1], [0, 0, 1, 0]],
    "CZ": [[1, 0, 0, 0], [0, 1, 0, 0], [0, 0, 1, 0], [0, 0, 0, -1]],
    "CY": [[1, 0, 0, 0], [0, 1, 0, 0], [0, 0, 0, -1j], [0, 0, 1j, 0]],
    "SWAP": [[1, 0, 0, 0], [0, 0, 1, 0], [0, 1, 0, 0], [0, 0, 0, 1]],
    "CSWAP": np.eye(8)[[0, 1, 2, 3, 4, 6, 5, 7]].tolist(),
    "Toffoli": np.eye(8)[[0, 1, 2, 3, 4, 5, 7, 6]].tolist(),
}
PARAM_COUNT = {"RX": 1, "RY": 1, "RZ": 1, "Rot": 3, "PhaseShift": 1, "ControlledPhaseShift": 1,
               "CRX": 1, "CRY": 1, "CRZ": 1}

_M1 = _H * (math.sqrt(2.0) + 1.0) / 4.0
_M2 = _H * (math.sqrt(2.0) - 1.0) / 4.0
TWO_TERM = [[0.5, 1.0, math.pi / 2], [-0.5, 1.0, -math.pi / 2]]
FOUR_TERM = [[_M1, 1.0, math.pi / 2], [-_M1, 1.0, -math.pi / 2],
             [-_M2, 1.0, 3 * math.pi / 2], [_M2, 1.0, -3 * math.pi / 2]]


def _t(x, rdtype):
    if isinstance(x, torch.Tensor):
        x = x.reshape(())
        if x.is_complex():
            x = x.real
        return x.to(rdtype)
    return torch.tensor(float(x), dtype=rdtype)


def _cplx(x, cdtype):
    return torch.complex(x, torch.zeros_like(x)).to(cdtype)


def gate_matrix(name, params, cdtype=torch.complex128):
    """2^k x 2^k matrix of a gate as a differentiable torch tensor."""
    rdtype = torch.float64 if cdtype == torch.complex128 else torch.float32
    if name in FIXED:
        return torch.tensor(FIXED[name], dtype=cdtype)
    p = [_t(v, rdtype) for v in params]
    one = torch.ones((), dtype=cdtype)
    zero = torch.zeros((), dtype=cdtype)
    if name in ("RX", "CRX", "RY", "CRY"):
        c, s = _cplx(torch.cos(p[0] / 2), cdtype), _cplx(torch.sin(p[0] / 2), cdtype)
        if name.endswith("RX"):
            blk = [[c, -1j * s], [-1j * s, c]]
        else:
            blk = [[c, -s], [s, c]]
    elif name in ("RZ", "CRZ"):
        e0 = torch.exp(-0.5j * _cplx(p[0], cdtype))
        e1 = torch.exp(0.5j * _cplx(p[0], cdtype))
        blk = [[e0, zero], [zero, e1]]
    elif name in ("PhaseShift", "ControlledPhaseShift"):
        blk = [[one, zero], [zero, torch.exp(1j * _cplx(p[0], cdtype))]]
    elif name == "Rot":
        phi, th, om = (_cplx(v, cdtype) for v in p)
        c, s = torch.cos(th / 2), torch.sin(th / 2)
        blk = [[torch.exp(-0.5j * (phi + om)) * c, -torch.exp(0.5j * (phi - om)) * s],
               [torch.exp(-0.5j * (phi - om)) * s, torch.exp(0.5j * (phi + om)) * c]]
    else:
        raise KeyError(name)
    b = torch.stack([torch.stack(r) for r in blk])
    if name.startswith("C"):
        top = torch.cat([torch.eye(2, dtype=cdtype), torch.zeros((2, 2), dtype=cdtype)], 1)
        bot = torch.cat([torch.zeros((2, 2), dtype=cdtype), b], 1)
        return torch.cat([top, bot], 0)
    return b


def _perms_single(q, n):
    return list(range(1, q + 1)) + [0] + list(range(q + 1, n))


def _perms_multi(qbts, n):
    order = list(qbts) + [x for x in range(n) if x not in qbts]
    return [order.index(i) for i in range(n)]


def apply_gate(state, mat, qubits, n):
    """One step of the reference loop: tensordot(gate, state, axes).permute(perms)."""
    k = len(qubits)
    g = mat.reshape([2] * (2 * k))
    if k == 1:
        out = torch.tensordot(g, state, ([1], list(qubits)))
        return out.permute(_perms_single(qubits[0], n))
    out = torch.tensordot(g, state, (list(range(k, 2 * k)), list(qubits)))
    return out.permute(_perms_multi(qubits, n))


def propagate(n, gates, cdtype=torch.complex128, init=None):
    """Final state [2]*n after applying ``gates`` to |0..0> (or ``init``)."""
    if init is None:
        state = torch.zeros([2] * n, dtype=cdtype)
        state.view(-1)[0] = 1.0
    else:
        state = torch.as_tensor(np.asarray(init).reshape([2] * n)).to(cdtype)
    for name, qubits, params in gates:
        if name == "Unitary":
            mat = torch.as_tensor(np.asarray(params)).to(cdtype)
        else:
            mat = gate_matrix(name, params, cdtype)
        state = apply_gate(state, mat, list(qubits), n)
    return state


def _obs_matrix(obs, cdtype):
    name, qubits = obs[0], obs[1]
    if isinstance(name, str):
        return gate_matrix(name, [], cdtype), list(qubits)
    return torch.as_tensor(np.asarray(name)).to(cdtype), list(qubits)


def measure(state, n, measurements, cdtype=torch.complex128):
    """Results per measurement, following tedq/backends/pytorch_backend.py:393-498."""
    axes = list(range(n))
    results = []
    for meas in measurements:
        kind = meas[0]
        if kind == "expval":
            obs = meas[1]
            if isinstance(obs, list):
                ms = state
                for ob in obs:
                    m, qb = _obs_matrix(ob, cdtype)
                    ms = torch.tensordot(m.reshape(2, 2), ms, dims=([1], qb)).permute(_perms_single(qb[0], n))
                results.append(torch.tensordot(torch.conj(state), ms, dims=(axes, axes)).real.squeeze())
            else:
                m, qb = _obs_matrix(obs, cdtype)
                k = len(qb)
                t = torch.tensordot(m.reshape([2] * (2 * k)), state, dims=(list(range(k, 2 * k)), qb))
                t = t.permute(_perms_multi(qb, n))
                results.append(torch.tensordot(torch.conj(state), t, dims=(axes, axes)).real.squeeze())
        elif kind == "probs":
            p = torch.abs(state) ** 2
            qs = meas[1]
            if qs is None:
                results.append(p)
            else:
                rest = [x for x in axes if x not in qs]
                results.append(torch.sum(p, dim=rest) if rest else p)
        elif kind == "state":
            results.append(state)
        else:
            raise KeyError(kind)
    return results


def execute(n, gates, measurements, cdtype=torch.complex128, init=None):
    """Stacked results, as the reference's ``execute`` (pytorch_backend.py:385-391)."""
    return torch.stack(measure(propagate(n, gates, cdtype, init), n, measurements, cdtype), 0)


# ---------------------------------------------------------------------------
# independent checks of the restatement
# ---------------------------------------------------------------------------

def dense_operator(n, name, qubits, params=()):
    """Brute-force 2^n x 2^n operator of one gate via Kronecker embedding (numpy, complex128)."""
    if name == "Unitary":
        m = np.asarray(params, dtype=np.complex128)
    else:
        m = gate_matrix(name, params, torch.complex128).detach().numpy()
    k = len(qubits)
    dim = 2 ** n
    full = np.zeros((dim, dim), dtype=np.complex128)
    others = [q for q in range(n) if q not in qubits]
    for col in range(dim):
        bits = [(col >> (n - 1 - q)) & 1 for q in range(n)]
        sub_in = 0
        for q in qubits:
            sub_in = (sub_in << 1) | bits[q]
        for sub_out in range(2 ** k):
            amp = m[sub_out, sub_in]
            if amp == 0:
                continue
            ob = list(bits)
            for i, q in enumerate(qubits):
                ob[q] = (sub_out >> (k - 1 - i)) & 1
            row = 0
            for q in range(n):
                row = (row << 1) | ob[q]
            full[row, col] += amp
    del others
    return full


def dense_state(n, gates):
    """|psi> by multiplying dense 2^n operators (independent of the tensordot loop)."""
    psi = np.zeros(2 ** n, dtype=np.complex128)
    psi[0] = 1.0
    for name, qubits, params in gates:
        psi = dense_operator(n, name, list(qubits), params) @ psi
    return psi


def shift_recipe(name):
    if name in ("CRX", "CRY", "CRZ"):
        return FOUR_TERM
    return TWO_TERM


def param_shift_jacobian(n, gate_fn, theta, measurements, cdtype=torch.complex128):
    """d outputs / d theta by the reference's shift rules; ``gate_fn(theta)`` -> gate list and names per column."""
    theta = np.asarray(theta, dtype=np.float64)
    gates, names = gate_fn(theta)
    cols = []
    for p in range(theta.size):
        g = 0.0
        for c, a, s in shift_recipe(names[p]):
            th = theta * a
            th = th.copy()
            th.reshape(-1)[p] += s
            g = g + c * execute(n, gate_fn(th)[0], measurements, cdtype).detach().numpy()
        cols.append(np.asarray(g).reshape(-1))
    return np.stack(cols, axis=1)


def finite_diff_jacobian(n, gate_fn, theta, measurements, h=1e-6):
    theta = np.asarray(theta, dtype=np.float64)
    cols = []
    for p in range(theta.size):
        tp, tm = theta.copy(), theta.copy()
        tp.reshape(-1)[p] += h
        tm.reshape(-1)[p] -= h
        fp = execute(n, gate_fn(tp)[0], measurements).detach().numpy()
        fm = execute(n, gate_fn(tm)[0], measurements).detach().numpy()
        cols.append(((fp - fm) / (2 * h)).reshape(-1))
    return np.stack(cols, axis=1)


# ---------------------------------------------------------------------------
# benchmark workloads
# ---------------------------------------------------------------------------

def hea_gates(n, theta):
    """Hardware-efficient ansatz: per layer RY(theta[l,q,0]) RZ(theta[l,q,1]) on every qubit,
    then CNOT(0,1), CNOT(1,2), ..., CNOT(n-2,n-1).  ``theta``: [layers, n, 2]."""
    gates, names = [], []
    L = theta.shape[0]
    for layer in range(L):
        for q in range(n):
            gates.append(("RY", [q], [theta[layer, q, 0]]))
            gates.append(("RZ", [q], [theta[layer, q, 1]]))
            names += ["RY", "RZ"]
        for q in range(n - 1):
            gates.append(("CNOT", [q, q + 1], []))
    return gates, names
