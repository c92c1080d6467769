"""Pin the CPU oracle before trusting it.

Checks, none of which retype the oracle's own formulas:
* the reference's golden constants (test/test_pytorch_backend.py:386-586 and
  the composite-cost gradients of test/test_jax_backend.py:385-428);
* closed forms: RX(a)RY(b)|0> gives <Z> = cos a cos b;
* dense brute force: Kronecker-embedded 2^n x 2^n operators (a construction
  independent of the tensordot/permute loop) on random circuits with every
  gate type, including non-adjacent and reversed qubit orders;
* gradient identities: autograd == parameter shift == central differences;
* every gate matrix by identities that fix it: Pauli algebra on the
  computational basis, Pauli exponentials (scipy expm) for the rotations,
  Rot = RZ RY RZ, S^2 = Z, T^2 = S, SX^2 = X, H = (X+Z)/sqrt 2, CZ / CY / SWAP
  from CNOT, controlled rotations as |0><0| x I + |1><1| x R, and Toffoli /
  CSWAP as classical logic on all 8 basis states.  The dense brute force above
  builds its operators from these same matrices, so these checks are what
  pins them.
"""

import math

import numpy as np
import pytest
import torch

from oracle import sv_oracle as O

A, B = 0.54, 0.12


def _golden_circuit(a, b):
    return [("RX", [0], [a]), ("RY", [0], [b])]


def test_golden_expval_probs_state():
    # test_pytorch_backend.py:392-409
    r = O.execute(2, _golden_circuit(A, B), [("expval", ("PauliZ", [0])), ("expval", ("PauliX", [1]))])
    assert np.allclose(np.round(r.numpy(), 5), np.round([0.85154057, 0.0], 5))
    # :412-430 probs (complex64 reference path, compared at 5 d.p. as the reference does)
    p = O.execute(2, _golden_circuit(A, B), [("probs", None)], cdtype=torch.complex64)[0]
    assert np.array_equal(np.round(p.numpy(), 5),
                          np.round(np.array([[0.9257702, 0.], [0.07422972, 0.]], dtype=np.float32), 5))
    # :433-451 partial probs
    pp = O.execute(2, _golden_circuit(A, B), [("probs", [1])])[0]
    assert np.allclose(pp.numpy(), [1.0, 0.0])
    # :454-472 state
    st = O.execute(2, _golden_circuit(A, B), [("state",)], cdtype=torch.complex64)[0]
    want = np.array([[0.9620366 + 0.01599429j, 0], [0.05779156 - 0.26625147j, 0]], dtype=np.complex64)
    assert np.array_equal(np.round(st.numpy(), 5), np.round(want, 5))


def test_golden_backprop_gradients():
    # test_pytorch_backend.py:483-530
    a = torch.tensor(A, dtype=torch.float64, requires_grad=True)
    b = torch.tensor(B, dtype=torch.float64, requires_grad=True)
    r = O.execute(2, _golden_circuit(a, b), [("expval", ("PauliZ", [0])), ("expval", ("PauliX", [1]))])
    r[0].backward()
    assert round(a.grad.item(), 5) == round(-0.5104387, 5)
    assert round(b.grad.item(), 5) == round(-0.10267819, 5)


def test_golden_composite_cost_gradient():
    # test_jax_backend.py:385-428: cost = |-(probs . state) + exp(sum probs(q0)) + 3.3 <Z0>|
    a = torch.tensor(A, dtype=torch.float64, requires_grad=True)
    b = torch.tensor(B, dtype=torch.float64, requires_grad=True)
    res = O.measure(O.propagate(2, _golden_circuit(a, b)), 2,
                    [("expval", ("PauliZ", [0])), ("probs", None), ("probs", [0]), ("state",)])
    cost = -torch.tensordot(res[1].to(torch.complex128), res[3]) + torch.exp(res[2].sum()) + 3.3 * res[0]
    cost = torch.abs(cost)
    cost.backward()
    assert np.allclose(np.round([a.grad.item(), b.grad.item()], 5), np.round([-1.3297417, -0.30148667], 5))


def test_closed_form_rx_ry():
    rng = np.random.default_rng(0)
    for a, b in rng.uniform(0, 2 * math.pi, size=(5, 2)):
        r = O.execute(1, _golden_circuit(a, b), [("expval", ("PauliZ", [0]))])
        assert abs(r.item() - math.cos(a) * math.cos(b)) < 1e-12


ALL_GATES = ["Hadamard", "PauliX", "PauliY", "PauliZ", "S", "T", "SX", "I", "RX", "RY", "RZ", "Rot",
             "PhaseShift", "CNOT", "CZ", "CY", "SWAP", "ControlledPhaseShift", "CRX", "CRY", "CRZ",
             "CSWAP", "Toffoli"]


def random_circuit(n, depth, rng, gate_names=ALL_GATES):
    gates = []
    for _ in range(depth):
        name = gate_names[rng.integers(len(gate_names))]
        k = 3 if name in ("CSWAP", "Toffoli") else 2 if (name.startswith("C") or name == "SWAP") else 1
        if k > n:
            continue
        qubits = [int(q) for q in rng.choice(n, size=k, replace=False)]
        params = [float(x) for x in rng.uniform(-math.pi, math.pi, size=O.PARAM_COUNT.get(name, 0))]
        gates.append((name, qubits, params))
    return gates


@pytest.mark.parametrize("seed", range(4))
def test_oracle_matches_dense_bruteforce(seed):
    rng = np.random.default_rng(seed)
    n = 4 + seed % 2
    gates = random_circuit(n, 30, rng)
    psi = O.propagate(n, gates).numpy().reshape(-1)
    assert np.allclose(psi, O.dense_state(n, gates), atol=1e-12)


def test_gradient_identities():
    rng = np.random.default_rng(7)
    n = 3
    theta = rng.uniform(0, 2 * math.pi, size=(2, n, 2))

    def fn(t):
        gates, names = O.hea_gates(n, t)
        gates.insert(3, ("CRX", [0, 2], [t[0, 0, 0]]))
        names = list(names)
        return gates, names
    meas = [("expval", ("PauliZ", [0])), ("expval", ("PauliX", [2]))]
    t = torch.tensor(theta, dtype=torch.float64, requires_grad=True)
    out = O.execute(n, O.hea_gates(n, t)[0], meas)
    jac_ad = np.stack([torch.autograd.grad(out[i], t, retain_graph=True)[0].numpy().reshape(-1)
                       for i in range(len(meas))])
    jac_ps = O.param_shift_jacobian(n, lambda th: O.hea_gates(n, th), theta, meas)
    jac_fd = O.finite_diff_jacobian(n, lambda th: O.hea_gates(n, th), theta, meas)
    assert np.allclose(jac_ad, jac_ps, atol=1e-10)
    assert np.allclose(jac_ad, jac_fd, atol=1e-6)
    # controlled rotations use the four-term rule
    ct = np.array([0.37])

    def crx(th):
        return [("Hadamard", [0], []), ("CRX", [0, 1], [th[0]])], ["CRX"]
    m2 = [("expval", ("PauliY", [1]))]
    assert np.allclose(O.param_shift_jacobian(2, crx, ct, m2), O.finite_diff_jacobian(2, crx, ct, m2), atol=1e-7)


def test_golden_notebook_values_and_mse_gradient():
    """examples/Expected_value_and_the_gradient_of_parameters.ipynb: values and MSE gradient as printed
    (4 decimals; the printed parameters are themselves rounded, hence the 5e-4 tolerance)."""
    import json
    import os
    g = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_values.json")))
    case = g["notebook_expval_and_gradient"]
    t = torch.tensor(case["params_printed_4dp"], dtype=torch.float64, requires_grad=True)
    gates = [("RX", [0], [t[0]]), ("RY", [1], [t[1]]), ("RZ", [2], [t[2]]),
             ("CNOT", [0, 1], []), ("CNOT", [1, 2], []), ("CNOT", [2, 0], [])]
    ev = O.execute(3, gates, [("expval", ("PauliZ", [q])) for q in range(3)])
    assert np.max(np.abs(ev.detach().numpy() - case["expval_printed_4dp"])) < 5e-4
    loss = torch.nn.functional.mse_loss(ev, torch.tensor(case["mse_target"], dtype=torch.float64))
    loss.backward()
    assert np.max(np.abs(t.grad.numpy() - case["grad_of_mse_printed_4dp"])) < 5e-4


# ---------------------------------------------------------------------------
# Gate matrices pinned by identities the mathematics fixes (no reference number
# exists for most of them; reference test_pytorch_backend.py:138-360 only
# compares the reference's numpy and torch copies with each other).  Each check
# is chosen so that a plausible slip (a swapped parameter, a flipped sign, the
# control taken as the target) fails it.
# ---------------------------------------------------------------------------

def _m(name, *params):
    return O.gate_matrix(name, list(params), torch.complex128).numpy()


def _kron(*ms):
    out = np.eye(1)
    for m in ms:
        out = np.kron(out, m)
    return out


_I2 = np.eye(2)
_P0 = np.diag([1.0, 0.0])
_P1 = np.diag([0.0, 1.0])


def _basis(bits):
    v = np.zeros(2 ** len(bits))
    v[int("".join(str(b) for b in bits), 2)] = 1.0
    return v


def test_pauli_and_clifford_matrices():
    X, Y, Z = _m("PauliX"), _m("PauliY"), _m("PauliZ")
    # computational basis: Z|0> = |0>, Z|1> = -|1>, X|0> = |1>
    assert np.allclose(Z @ _basis([0]), _basis([0])) and np.allclose(Z @ _basis([1]), -_basis([1]))
    assert np.allclose(X @ _basis([0]), _basis([1])) and np.allclose(X @ _basis([1]), _basis([0]))
    # Pauli algebra fixes Y including its sign: XY = iZ, Y hermitian and involutory
    assert np.allclose(X @ Y, 1j * Z)
    for P in (X, Y, Z):
        assert np.allclose(P, P.conj().T) and np.allclose(P @ P, _I2)
    assert np.allclose(_m("I"), _I2)
    assert np.allclose(_m("Hadamard"), (X + Z) / math.sqrt(2.0))
    S, T, SX = _m("S"), _m("T"), _m("SX")
    # S = PhaseShift(pi/2), T = PhaseShift(pi/4) (the definitions), hence S^2 = Z, T^2 = S
    assert np.allclose(S, _m("PhaseShift", math.pi / 2)) and np.allclose(T, _m("PhaseShift", math.pi / 4))
    assert np.allclose(S @ S, Z) and np.allclose(T @ T, S)
    # SX is the principal square root of X: SX^2 = X and SX = e^{i pi/4} RX(pi/2)
    assert np.allclose(SX @ SX, X)
    assert np.allclose(SX, np.exp(0.25j * math.pi) * _m("RX", math.pi / 2))


@pytest.mark.parametrize("theta", [0.0, 0.3, -1.7, 2.9, 5.5])
def test_rotation_matrices_are_pauli_exponentials(theta):
    from scipy.linalg import expm
    X, Y, Z = _m("PauliX"), _m("PauliY"), _m("PauliZ")
    assert np.allclose(_m("RX", theta), expm(-0.5j * theta * X))
    assert np.allclose(_m("RY", theta), expm(-0.5j * theta * Y))
    assert np.allclose(_m("RZ", theta), expm(-0.5j * theta * Z))
    # PhaseShift(phi) = diag(1, e^{i phi}) = e^{i phi/2} RZ(phi)
    assert np.allclose(_m("PhaseShift", theta), np.exp(0.5j * theta) * _m("RZ", theta))
    assert np.allclose(_m("PhaseShift", theta) @ _basis([0]), _basis([0]))


@pytest.mark.parametrize("phi,theta,omega", [(0.3, 0.4, 0.5), (-1.1, 2.2, 0.7), (2.5, -0.9, -2.0)])
def test_rot_is_rz_ry_rz(phi, theta, omega):
    # qubit.py:1169-1189: Rot(phi, theta, omega) = RZ(omega) RY(theta) RZ(phi), parameters in call order
    want = _m("RZ", omega) @ _m("RY", theta) @ _m("RZ", phi)
    assert np.allclose(_m("Rot", phi, theta, omega), want)
    # a swapped first and last parameter would differ here (RZ(phi) and RZ(omega) do not commute with RY)
    assert not np.allclose(_m("Rot", omega, theta, phi), want)


def test_cnot_control_first_and_two_qubit_identities():
    CNOT, H, S = _m("CNOT"), _m("Hadamard"), _m("S")
    # control is the first listed qubit (qubit.py:644): |a,b> -> |a, a xor b>
    for a in (0, 1):
        for b in (0, 1):
            assert np.allclose(CNOT @ _basis([a, b]), _basis([a, a ^ b]))
    # CZ = (I x H) CNOT (I x H); CY = (I x S) CNOT (I x S^dagger)
    assert np.allclose(_m("CZ"), _kron(_I2, H) @ CNOT @ _kron(_I2, H))
    assert np.allclose(_m("CY"), _kron(_I2, S) @ CNOT @ _kron(_I2, S.conj().T))
    # SWAP = CNOT_01 CNOT_10 CNOT_01, with CNOT_10 the CNOT controlled by the second qubit
    swap_h = _kron(H, H)
    cnot10 = swap_h @ CNOT @ swap_h
    assert np.allclose(_m("SWAP"), CNOT @ cnot10 @ CNOT)
    for a in (0, 1):
        for b in (0, 1):
            assert np.allclose(_m("SWAP") @ _basis([a, b]), _basis([b, a]))


@pytest.mark.parametrize("theta", [0.5, -2.3, 4.0])
def test_controlled_rotations_are_block_diagonal(theta):
    # qubit.py:1327-1334, 1429-1642: |0><0| x I + |1><1| x R, control first
    for cname, rname in (("CRX", "RX"), ("CRY", "RY"), ("CRZ", "RZ"), ("ControlledPhaseShift", "PhaseShift")):
        want = _kron(_P0, _I2) + _kron(_P1, _m(rname, theta))
        assert np.allclose(_m(cname, theta), want), cname
        # the target-first embedding differs at these angles, except for the controlled phase, which is
        # symmetric in its two qubits (diag(1, 1, 1, e^{i phi}) either way)
        if cname != "ControlledPhaseShift":
            assert not np.allclose(_m(cname, theta), _kron(_I2, _P0) + _kron(_m(rname, theta), _P1)), cname


def test_three_qubit_gates_on_all_basis_states():
    # Toffoli: controls first two qubits, X on the third; CSWAP: control first, swap the other two
    for bits in [(a, b, c) for a in (0, 1) for b in (0, 1) for c in (0, 1)]:
        a, b, c = bits
        assert np.allclose(_m("Toffoli") @ _basis(bits), _basis([a, b, c ^ (a & b)])), bits
        want = [a, c, b] if a else [a, b, c]
        assert np.allclose(_m("CSWAP") @ _basis(bits), _basis(want)), bits


def test_every_gate_is_unitary_and_shift_rules_match_derivatives():
    rng = np.random.default_rng(11)
    for name in ALL_GATES:
        params = rng.uniform(-math.pi, math.pi, size=O.PARAM_COUNT.get(name, 0))
        m = _m(name, *params)
        assert np.allclose(m @ m.conj().T, np.eye(m.shape[0])), name
    # the recipes (ops_abc.py:261-280, qubit.py:32-42) reproduce d/dtheta of a matrix element exactly for
    # the generators with eigenvalues +-1/2 (two terms) and {0, +-1/2} (four terms)
    for name in ("RX", "RY", "RZ", "PhaseShift", "CRX", "CRY", "CRZ", "ControlledPhaseShift"):
        th = 0.77
        h = 1e-6
        dim = _m(name, th).shape[0]
        # expectation values are quadratic in U; the recipe is exact for <U^dag O U>, so compare on one
        O_ = _kron(*([_m("PauliY")] * int(math.log2(dim))))
        psi = np.ones(dim) / math.sqrt(dim)
        f = lambda t: np.real(psi.conj() @ _m(name, t).conj().T @ O_ @ _m(name, t) @ psi)  # noqa: E731
        d = (f(th + h) - f(th - h)) / (2 * h)
        d_rec = sum(c * f(a * th + s) for c, a, s in O.shift_recipe(name))
        assert abs(d - d_rec) < 1e-7, name


def test_package_operator_matrices_equal_the_pinned_oracle():
    """The package's gate classes (used for observables and user matrices) carry the same matrices."""
    import tedq_amd as qai
    rng = np.random.default_rng(12)
    for name in ALL_GATES:
        params = [float(x) for x in rng.uniform(-math.pi, math.pi, size=O.PARAM_COUNT.get(name, 0))]
        k = 3 if name in ("CSWAP", "Toffoli") else 2 if (name.startswith("C") or name == "SWAP") else 1
        g = getattr(qai, name)(*params, qubits=list(range(k)), do_queue=False)
        assert np.allclose(np.asarray(g.matrix, dtype=np.complex128).reshape(2 ** k, 2 ** k), _m(name, *params)), name


def test_hea20_fixture_matches_the_oracle():
    """tests/golden/hea20_d10_rows.json (the GPU parity reference for the benchmarked circuit) still
    holds the oracle's numbers: values of both rows from a fresh propagation, read two ways (the
    oracle's tensordot measurement for three qubits and sum_x p_x (1 - 2 x_q) for all), and three
    gradient entries per row by central differences."""
    import json
    import os
    with open(os.path.join(os.path.dirname(__file__), "golden", "hea20_d10_rows.json")) as f:
        gold = json.load(f)
    n, L = gold["qubits"], gold["layers"]
    idx = torch.arange(2 ** n)

    def values(params, all_q=True):
        with torch.no_grad():
            psi = O.propagate(n, O.hea_gates(n, np.asarray(params).reshape(L, n, 2))[0])
        p = (psi.abs() ** 2).reshape(-1)
        z = np.array([float((p * (1 - 2 * ((idx >> (n - 1 - q)) & 1))).sum()) for q in range(n)])
        return psi, z

    for row in gold["rows"]:
        psi, z = values(row["params"])
        assert np.max(np.abs(z - row["values"])) < 1e-12
        tz = O.measure(psi, n, [("expval", ("PauliZ", [q])) for q in (0, 9, 19)])
        assert np.allclose([float(t) for t in tz], np.asarray(row["values"])[[0, 9, 19]], atol=1e-12)
        h = 1e-5
        for col in (0, 201, 399):
            tp, tm = np.array(row["params"]), np.array(row["params"])
            tp[col] += h
            tm[col] -= h
            fd = (values(tp)[1].sum() - values(tm)[1].sum()) / (2 * h)
            assert abs(fd - row["grad_of_sum"][col]) < 1e-7, col
