"""Pin the CPU oracle before trusting it.

Checks, none of which retype the oracle's own formulas:
* the reference's golden constants (test/test_pytorch_backend.py:386-586 and
  the composite-cost gradients of test/test_jax_backend.py:385-428);
* closed forms: RX(a)RY(b)|0> gives <Z> = cos a cos b;
* dense brute force: Kronecker-embedded 2^n x 2^n operators (a construction
  independent of the tensordot/permute loop) on random circuits with every
  gate type, including non-adjacent and reversed qubit orders;
* gradient identities: autograd == parameter shift == central differences.
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
