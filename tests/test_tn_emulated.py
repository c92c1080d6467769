"""Contraction-mode host logic on the CPU: descriptors, light-cone reduction,
paths and the reverse contraction tree, with the device calls emulated in
numpy (tests/tn_emulator.py) and checked against the complex128 oracle.

Tolerances: values 2e-5 and gradients 1e-4 absolute (complex64 leaves and
single-precision storage of intermediates against complex128).
"""

import numpy as np
import pytest
import scipy.linalg as sl
import torch

import tedq_amd as qai
from oracle import sv_oracle as O
from tedq_amd.tn.executor import TNExecutor, make_desc
from tedq_amd.tn.network import cancel_inverse_pairs, gen_tensor_networks  # noqa: F401
from tedq_amd.tn.path import greedy_path, search_path, sweep_path
from tn_emulator import contract_np, install


def test_descriptor_contraction_matches_einsum():
    rng = np.random.default_rng(0)
    letters = "abcdefgh"
    for trial in range(20):
        syms = list(letters)
        rng.shuffle(syms)
        na, nb = int(rng.integers(1, 6)), int(rng.integers(1, 6))
        la = syms[:na]
        lb = list(rng.choice(la, size=int(rng.integers(0, na + 1)), replace=False)) + syms[na:na + nb]
        rng.shuffle(lb)
        shared = [x for x in la if x in lb]
        keep_shared = [x for x in shared if rng.random() < 0.3]
        out = [x for x in la if x not in shared] + [x for x in lb if x not in shared] + keep_shared
        rng.shuffle(out)
        A = (rng.normal(size=2 ** len(la)) + 1j * rng.normal(size=2 ** len(la)))
        B = (rng.normal(size=2 ** len(lb)) + 1j * rng.normal(size=2 ** len(lb)))
        C = np.zeros((1, 2 ** len(out)), complex)
        contract_np(make_desc(la, lb, out), A[None], B[None], C)
        ref = np.einsum("".join(la) + "," + "".join(lb) + "->" + "".join(out),
                        A.reshape([2] * len(la)), B.reshape([2] * len(lb)))
        assert np.allclose(C.reshape(-1), np.asarray(ref).reshape(-1)), trial


def test_paths_cover_every_tensor_once():
    inputs = [["a", "b"], ["b", "c"], ["c", "d"], ["d", "a"], ["e"], ["e", "f"]]
    for path, sets in (greedy_path(inputs, ["f"]), sweep_path(inputs, ["f"]), search_path(inputs, ["f"])[:2]):
        used = [x for st in path for x in st]
        assert sorted(used) == list(range(len(inputs) + len(path) - 1))
        assert sets[-1] == frozenset({"f"})


def test_absorbing_single_qubit_tensors_first():
    """absorb=4: every rank<=2 tensor is merged into a neighbour before the greedy rule runs, as long
    as the merged rank stays <= 4 (here every one can be), so no single-qubit tensor meets a larger
    intermediate later; both paths stay valid and end in the output's index set."""
    # a 6-qubit brick of 2-qubit gates (rank 4), each qubit line also carrying 1-qubit gates (rank 2)
    inputs, line, fresh = [], {q: f"q{q}_0" for q in range(6)}, [0]

    def idx():
        fresh[0] += 1
        return f"x{fresh[0]}"
    for q in range(6):
        inputs.append([line[q]])                                    # |0> (rank 1)
    for layer in range(4):
        for q in range(6):
            o = idx()
            inputs.append([line[q], o])                             # 1-qubit gate
            line[q] = o
        for q in range(layer % 2, 5, 2):
            o1, o2 = idx(), idx()
            inputs.append([line[q], line[q + 1], o1, o2])           # 2-qubit gate
            line[q], line[q + 1] = o1, o2
    output = [line[q] for q in range(6)]
    p0, s0 = greedy_path(inputs, output)
    p1, s1 = greedy_path(inputs, output, absorb=4)
    n = len(inputs)
    for p, s in ((p0, s0), (p1, s1)):
        used = [x for st in p for x in st]
        assert sorted(used) == list(range(n + len(p) - 1))
        assert s[-1] == frozenset(output)
    touches_small = [min(len(s1[i]), len(s1[j])) <= 2 for i, j in p1]
    prefix = touches_small.index(False)
    assert prefix > 0 and not any(touches_small[prefix:])          # small tensors only in the prefix
    assert all(len(s1[n + k]) <= 4 for k in range(prefix))         # ... whose results stay gate-sized


def test_lazy_absorption_keeps_single_qubit_tensors_off_large_intermediates():
    """absorb=-4: a rank<=2 tensor that the greedy rule would contract into a rank>=8 intermediate is
    merged into a gate-sized neighbour instead (here one always exists); the plain rule does apply
    single-qubit tensors to the large state, so the two paths differ where it matters."""
    n_q = 10
    line = {q: f"s{q}" for q in range(n_q)}
    inputs = [[line[q] for q in range(n_q)]]                         # a 10-qubit state (rank 10)
    fresh = [0]

    def idx():
        fresh[0] += 1
        return f"x{fresh[0]}"
    for layer in range(3):
        for q in range(n_q):
            o = idx()
            inputs.append([line[q], o])
            line[q] = o
        for q in range(layer % 2, n_q - 1, 2):
            o1, o2 = idx(), idx()
            inputs.append([line[q], line[q + 1], o1, o2])
            line[q], line[q + 1] = o1, o2
    output = [line[q] for q in range(n_q)]
    n = len(inputs)
    for absorb, expect_small_on_large in ((0, True), (-4, False)):
        p, s = greedy_path(inputs, output, 4.0, 0.0, 0, absorb)
        assert sorted(x for st in p for x in st) == list(range(n + len(p) - 1))
        assert s[-1] == frozenset(output)
        small_on_large = any(min(len(s[i]), len(s[j])) <= 2 and max(len(s[i]), len(s[j])) >= 8 for i, j in p)
        assert small_on_large == expect_small_on_large, absorb


def _random_circuit(rng, n, depth):
    """A random circuit as (frontend builder, oracle gate list as a function of the parameter vector)."""
    spec = []
    for _ in range(depth):
        kind = rng.choice(["RX", "RY", "RZ", "Rot", "PhaseShift", "CRX", "CRZ", "CNOT", "CZ", "Hadamard", "S",
                           "Toffoli", "SWAP", "U2", "RXinv"])
        qs = list(rng.choice(n, size=3, replace=False))
        spec.append((str(kind), qs))
    u2 = sl.expm(-0.3j * np.kron(np.array([[0, 1], [1, 0]]), np.array([[1, 0], [0, -1]])))
    npar = {"RX": 1, "RY": 1, "RZ": 1, "Rot": 3, "PhaseShift": 1, "CRX": 1, "CRZ": 1, "RXinv": 1}

    def build(*p):
        it = iter(p)
        for kind, qs in spec:
            if kind == "U2":
                qai.Unitary(u2, qubits=qs[:2])
            elif kind == "RXinv":
                qai.RX(next(it), qubits=qs[:1]).inverse()
            elif kind in ("Toffoli",):
                qai.Toffoli(qubits=qs[:3])
            elif kind in ("CRX", "CRZ", "CNOT", "CZ", "SWAP"):
                getattr(qai, kind)(*[next(it) for _ in range(npar.get(kind, 0))], qubits=qs[:2])
            else:
                getattr(qai, kind)(*[next(it) for _ in range(npar.get(kind, 0))], qubits=qs[:1])

    def gates(theta):
        it = iter(theta)
        out = []
        for kind, qs in spec:
            if kind == "U2":
                out.append(("Unitary", qs[:2], u2))
            elif kind == "RXinv":
                out.append(("RX", qs[:1], [-next(it)]))
            elif kind == "Toffoli":
                out.append(("Toffoli", qs[:3], []))
            elif kind in ("CRX", "CRZ", "CNOT", "CZ", "SWAP"):
                out.append((kind, qs[:2], [next(it) for _ in range(npar.get(kind, 0))]))
            else:
                out.append((kind, qs[:1], [next(it) for _ in range(npar.get(kind, 0))]))
        return out

    P = sum(npar.get(k, 0) for k, _ in spec)
    return build, gates, P


@pytest.mark.parametrize("seed", range(6))
def test_emulated_executor_values_and_gradients(seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(3, 7))
    build, gates, P = _random_circuit(rng, n, int(rng.integers(8, 20)))
    q0, q1 = [int(x) for x in rng.choice(n, size=2, replace=False)]
    herm = np.array([[1.0, 0.5 - 0.2j], [0.5 + 0.2j, -0.3]])
    meas_kind = seed % 3

    def cd(*p):
        build(*p)
        if meas_kind == 0:
            return [qai.expval(qai.PauliZ(qubits=[q0])), qai.expval(qai.PauliX(qubits=[q1]))]
        if meas_kind == 1:
            return [qai.expval([qai.PauliZ(qubits=[q0]), qai.PauliY(qubits=[q1])]),
                    qai.expval(qai.Unitary(herm, qubits=[q1]))]
        return [qai.probs(qubits=sorted([q0, q1]))]

    theta = rng.uniform(0, 2 * np.pi, size=max(P, 1))[:P]
    cc = qai.Circuit(cd, n, *[torch.tensor(t) for t in theta]).compilecircuit(backend="pytorch", tn_mode=True)
    ex = install(TNExecutor(cc, max_bytes=(1 << 40) if seed % 2 else 64))   # odd: full storage, even: segments
    rows = torch.tensor(np.stack([theta, theta[::-1].copy()]), dtype=torch.float32)
    ups = [torch.ones(2) if s.kind == "fexp" else torch.ones(2, 4) for s in ex.specs]
    outs, grad = ex.value_and_grad(rows, ups)
    for b in range(2):
        th = torch.tensor(rows[b].numpy(), dtype=torch.float64, requires_grad=True)
        if meas_kind == 0:
            ms = [("expval", ("PauliZ", [q0])), ("expval", ("PauliX", [q1]))]
        elif meas_kind == 1:
            ms = [("expval", [("PauliZ", [q0]), ("PauliY", [q1])]), ("expval", (herm, [q1]))]
        else:
            ms = [("probs", sorted([q0, q1]))]
        res = O.measure(O.propagate(n, gates(list(th))), n, ms)
        for o, r in zip(outs, res):
            got = o[b].numpy().reshape(-1)
            want = r.detach().numpy().reshape(-1)
            assert np.max(np.abs(got - want)) < 2e-5
        if P:
            tot = sum(r.sum() for r in res)
            (g,) = torch.autograd.grad(tot, th)
            assert np.max(np.abs(grad[b].numpy() - g.numpy())) < 1e-4


def test_fused_gate_backward_dispatch_keeps_gradients():
    """Gate steps on tensors of >= 2^10 rows take the one-pass backward (tedq_tn_gate_backward, here its
    numpy twin); on a 3x4 lattice of 4 layers (such steps on rank-12 intermediates) value and gradient
    equal the oracle's (2e-5, 1e-4) with it on and off."""
    from tedq_amd.tn.executor import _gate_step
    R, C, layers, obs = 3, 4, 4, 5
    n = R * C
    bonds = [(r * C + c, r * C + c + 1) for r in range(R) for c in range(C - 1)]
    bonds += [(r * C + c, (r + 1) * C + c) for r in range(R - 1) for c in range(C)]

    def cd(*p):
        it = iter(p)
        for _ in range(layers):
            for q in range(n):
                qai.RX(next(it), qubits=[q])
                qai.RZ(next(it), qubits=[q])
            for a, b in bonds:
                qai.CZ(qubits=[a, b])
        return [qai.expval(qai.PauliZ(qubits=[obs]))]

    rng = np.random.default_rng(8)
    theta = rng.uniform(0, 2 * np.pi, size=2 * n * layers)
    rows = torch.tensor(theta[None], dtype=torch.float32)
    th = torch.tensor(rows[0].numpy(), dtype=torch.float64, requires_grad=True)
    gates, it = [], iter(th)
    for _ in range(layers):
        for q in range(n):
            gates += [("RX", [q], [next(it)]), ("RZ", [q], [next(it)])]
        gates += [("CZ", [a, b], []) for a, b in bonds]
    ref = O.execute(n, gates, [("expval", ("PauliZ", [obs]))])
    ref.sum().backward()
    for fused in (True, False):
        # without diagonal reduction, so that RZ / CZ stay gates and the large steps are plain gate steps
        cc = qai.Circuit(cd, n, *[torch.tensor(t) for t in theta]).compilecircuit(backend="pytorch", tn_mode=True,
                                                                                  tn_simplify=False)
        ex = install(TNExecutor(cc))
        ex.fused_backward = fused
        calls = []
        inner = ex._gate_backward  # pylint: disable=protected-access

        def counting(*a, inner=inner, calls=calls):
            calls.append(1)
            return inner(*a)
        ex._gate_backward = counting  # pylint: disable=protected-access
        outs, grad = ex.value_and_grad(rows, [torch.ones(1)])
        assert abs(float(outs[0][0]) - float(ref.detach())) < 2e-5
        assert np.max(np.abs(grad[0].numpy() - th.grad.numpy())) < 1e-4
        descs = ex._descs  # pylint: disable=protected-access
        assert sum(_gate_step(descs[st.fwd]) for st in ex.networks[0].steps) > 0
        assert (len(calls) > 0) == fused


def test_light_cone_reduction_keeps_value_and_shrinks_network():
    n, steps = 10, 2

    def cd(*p):
        it = iter(p)
        for q in range(1, n, 2):
            qai.PauliX(qubits=[q])
        for _ in range(steps):
            for q in range(n):
                qai.RZ(next(it), qubits=[q])
            for q in range(n - 1):
                qai.CNOT(qubits=[q, q + 1])
                qai.RY(next(it), qubits=[q + 1])
                qai.CNOT(qubits=[q, q + 1])
        return qai.expval(qai.PauliZ(qubits=[4]))

    P = steps * (2 * n - 1)
    rng = np.random.default_rng(5)
    theta = rng.uniform(0, 6.28, P)
    cc = qai.Circuit(cd, n, *[torch.tensor(t) for t in theta]).compilecircuit(backend="pytorch", tn_mode=True)
    tn = cc._tensor_networks[0]  # pylint: disable=protected-access
    red = cancel_inverse_pairs(tn)
    assert red.num_tensors < tn.num_tensors
    ex = install(TNExecutor(cc))
    outs, grad = ex.value_and_grad(torch.tensor(theta[None], dtype=torch.float32), [torch.ones(1)])
    th = torch.tensor(theta, requires_grad=True)
    gl = []
    it = iter(th)
    for q in range(1, n, 2):
        gl.append(("PauliX", [q], []))
    for _ in range(steps):
        for q in range(n):
            gl.append(("RZ", [q], [next(it)]))
        for q in range(n - 1):
            gl += [("CNOT", [q, q + 1], []), ("RY", [q + 1], [next(it)]), ("CNOT", [q, q + 1], [])]
    ref = O.execute(n, gl, [("expval", ("PauliZ", [4]))])
    ref.sum().backward()
    assert abs(float(outs[0][0]) - float(ref)) < 2e-5
    assert np.max(np.abs(grad[0].numpy() - th.grad.numpy())) < 1e-4
    # gates after the last interaction with qubit 4's past cancel: their gradients are exactly zero
    assert np.any(grad[0].numpy() == 0.0)


def test_gemm_shaped_steps_get_k_contiguous_operand_copies():
    """A chain wide enough for matrix-core steps: operand copies are planned and the value is unchanged."""
    n, steps, obs = 14, 5, 7
    from tedq_amd import workloads as W
    cc = W.disordered_chain(n, steps, obs, tn_mode=True)
    ex = install(TNExecutor(cc))
    assert ex._gemm, "expected GEMM-shaped steps"  # pylint: disable=protected-access
    rng = np.random.default_rng(4)
    rows = torch.tensor(W.chain_rows(rng.uniform(-5, 5, size=(1, n)), steps), dtype=torch.float32)
    outs, grad = ex.value_and_grad(rows, [torch.ones(1)])
    zz, xy = W.exchange_unitaries()
    th = torch.tensor(rows[0].numpy(), dtype=torch.float64, requires_grad=True)
    gates = [("PauliX", [q], []) for q in range(1, n, 2)]
    it = iter(th)
    for _ in range(steps):
        for q in range(n):
            gates.append(("RZ", [q], [next(it)]))
        for q in range(n - 1):
            gates += [("Unitary", [q, q + 1], zz), ("Unitary", [q, q + 1], xy)]
    ref = O.execute(n, gates, [("expval", ("PauliZ", [obs]))])
    ref.sum().backward()
    assert abs(float(outs[0][0]) - float(ref.detach())) < 2e-5
    assert np.max(np.abs(grad[0].numpy() - th.grad.numpy())) < 1e-4


def test_linear_paths_convert_to_ssa():
    from tedq_amd.tn.path import linear_to_ssa
    # opt_einsum convention: contract positions (i, j) of the current list, append the result
    assert linear_to_ssa([(0, 1), (0, 1), (0, 1)], 4) == [(0, 1), (2, 3), (4, 5)]
    assert linear_to_ssa([(1, 3), (0, 1), (0, 1)], 4) == [(1, 3), (0, 2), (4, 5)]
    assert linear_to_ssa([(2, 3), (0, 2), (0, 1)], 4) == [(2, 3), (0, 4), (1, 5)]


def _small_expval_circuit(**kw):
    def cd(*p):
        qai.RX(p[0], qubits=[0])
        qai.CRY(p[1], qubits=[0, 2])
        qai.CNOT(qubits=[2, 1])
        qai.RZ(p[2], qubits=[1])
        qai.Hadamard(qubits=[3])
        qai.CZ(qubits=[3, 1])
        return qai.expval(qai.PauliZ(qubits=[1]))
    return qai.Circuit(cd, 4, *([0.3] * 3)).compilecircuit(backend="pytorch", **kw)


def _small_reference(theta):
    th = torch.tensor(theta, dtype=torch.float64, requires_grad=True)
    gates = [("RX", [0], [th[0]]), ("CRY", [0, 2], [th[1]]), ("CNOT", [2, 1], []), ("RZ", [1], [th[2]]),
             ("Hadamard", [3], []), ("CZ", [3, 1], [])]
    ref = O.execute(4, gates, [("expval", ("PauliZ", [1]))])
    ref.sum().backward()
    return float(ref.detach()), th.grad.numpy()


def test_external_optimiser_paths_are_used_on_the_full_network():
    """use_cotengra=<module>: the engine calls HyperOptimizer(...).search(...) like the reference and
    contracts the unreduced network along the returned tree's SSA path."""
    calls = []

    class FakeTree:
        def __init__(self, inputs, output):
            self.ssa, _ = greedy_path(inputs, output)

        def get_ssa_path(self):
            return self.ssa

    class FakeCotengra:
        class HyperOptimizer:
            def __init__(self, **kw):
                calls.append(kw)

            def search(self, inputs, output, size_dict):  # pylint: disable=unused-argument
                return FakeTree(inputs, output)

    cc = _small_expval_circuit(use_cotengra=FakeCotengra)
    ex = install(TNExecutor(cc))
    tn = cc._tensor_networks[0]  # pylint: disable=protected-access
    assert calls and ex.path_report()[0]["strategy"] == "external"
    assert ex.networks[0].n_inputs == tn.num_tensors          # no light-cone reduction
    theta = np.array([0.7, -1.2, 0.4])
    outs, grad = ex.value_and_grad(torch.tensor(theta[None], dtype=torch.float32), [torch.ones(1)])
    v, g = _small_reference(theta)
    assert abs(float(outs[0][0]) - v) < 2e-5
    assert np.max(np.abs(grad[0].numpy() - g)) < 1e-4


def test_linear_paths_given_per_measurement():
    cc = _small_expval_circuit(tn_mode=True)
    n = cc._tensor_networks[0].num_tensors  # pylint: disable=protected-access
    ex = install(TNExecutor(cc, paths=[[(0, 1)] * (n - 1)]))
    theta = np.array([-0.3, 2.1, 1.0])
    outs, grad = ex.value_and_grad(torch.tensor(theta[None], dtype=torch.float32), [torch.ones(1)])
    v, g = _small_reference(theta)
    assert abs(float(outs[0][0]) - v) < 2e-5
    assert np.max(np.abs(grad[0].numpy() - g)) < 1e-4
    with pytest.raises(ValueError):
        TNExecutor(cc, paths=[[(0, 1)] * (n - 2)])


def _lattice_oracle(L, layers, theta):
    n = L * L
    it = iter(theta)
    gates = []
    for _ in range(layers):
        for q in range(n):
            gates.append(("RX", [q], [next(it)]))
            gates.append(("RZ", [q], [next(it)]))
        for r in range(L):
            for c in range(L - 1):
                gates.append(("CZ", [r * L + c, r * L + c + 1], []))
        for r in range(L - 1):
            for c in range(L):
                gates.append(("CZ", [r * L + c, (r + 1) * L + c], []))
    return gates


def test_sliced_contraction_equals_the_unsliced_one_and_the_oracle():
    """slicing_opts (the reference's hyper_opt key): the engine slices bond indices on its frozen path
    until every intermediate fits target_size, runs the sliced network once per slice (selector vectors
    in the constant arena) and adds values and gradients; they equal the unsliced run and the oracle.
    On the 3x3 lattice of 2 layers the <Z_4> network's intermediates exceed 2^6 elements; reaching that
    target takes several sliced (hyper-)indices."""
    from tedq_amd import workloads as W
    L, layers, obs = 3, 2, [4, 0]
    n = L * L
    full = install(W.lattice(L, layers, observe=obs, tn_mode=True).executor())
    sl_ex = install(W.lattice(L, layers, observe=obs, tn_mode=True,
                              hyper_opt={"slicing_opts": {"target_size": 2 ** 6}}).executor())
    rep_full, rep_sl = full.path_report(), sl_ex.path_report()
    assert len(sl_ex.networks[0].sliced) >= 2
    assert max(r["largest_rank"] for r in rep_sl) <= 6 < max(r["largest_rank"] for r in rep_full)
    rng = np.random.default_rng(11)
    theta = rng.uniform(0, 2 * np.pi, size=(2, 2 * n * layers))
    rows = torch.tensor(theta, dtype=torch.float32)
    ups = [torch.full((2,), 0.5) for _ in obs]
    v_f, g_f = full.value_and_grad(rows, ups)
    v_s, g_s = sl_ex.value_and_grad(rows, ups)
    assert np.max(np.abs(torch.stack(v_f).numpy() - torch.stack(v_s).numpy())) < 2e-6
    assert np.max(np.abs(g_f.numpy() - g_s.numpy())) < 2e-6
    # forward only (no backward): the sliced roots added per slice
    run = sl_ex.run_forward(rows)
    assert np.max(np.abs(torch.stack(sl_ex.measure(run)).numpy() - torch.stack(v_f).numpy())) < 2e-6
    th = torch.tensor(theta[1], dtype=torch.float64, requires_grad=True)
    ref = O.execute(n, _lattice_oracle(L, layers, th), [("expval", ("PauliZ", [q])) for q in obs])
    (ref.sum() / 2).backward()
    assert np.max(np.abs(np.array([float(v[1]) for v in v_s]) - ref.detach().numpy())) < 2e-5
    assert np.max(np.abs(g_s[1].numpy() - th.grad.numpy())) < 1e-4


def test_slices_of_an_external_tree_are_honoured():
    """An external contraction tree that slices indices (cotengra's ``sliced_inds``) is contracted
    along its path with those indices sliced, not whole."""

    class SlicedTree:
        def __init__(self, inputs, output):
            self.ssa, sets = greedy_path(inputs, output)
            counts = {}
            for t in inputs:
                for ix in t:
                    counts[ix] = counts.get(ix, 0) + 1
            widest = max(sets[len(inputs):], key=len)
            self.sliced_inds = {ix: None for ix in sorted(widest) if counts.get(ix) == 2}

        def get_ssa_path(self):
            return self.ssa

    cc = _small_expval_circuit(tn_mode=True)
    tn = cc._tensor_networks[0]  # pylint: disable=protected-access
    tree = SlicedTree(tn.input_indices, tn.output_indices)
    assert tree.sliced_inds
    ex = install(TNExecutor(cc, paths=[tree]))
    assert ex.networks[0].sliced == sorted(tree.sliced_inds)
    theta = np.array([0.7, -1.2, 0.4])
    outs, grad = ex.value_and_grad(torch.tensor(theta[None], dtype=torch.float32), [torch.ones(1)])
    v, g = _small_reference(theta)
    assert abs(float(outs[0][0]) - v) < 2e-5
    assert np.max(np.abs(grad[0].numpy() - g)) < 1e-4


def test_choose_slices_reaches_its_targets():
    from tedq_amd.tn.slicing import choose_slices
    from tedq_amd import workloads as W
    ex = W.disordered_chain(16, 3, 8, tn_mode=True).executor()
    net = ex.networks[0]
    inputs = [frozenset(t) for t in net.lay[:net.n_inputs]]
    sets = path_sets_of(net)
    for target in (2 ** 8, 2 ** 6):
        s = choose_slices(net.path, sets, net.n_inputs, net.output, target_size=target)
        assert max(len(t - set(s)) for t in sets[net.n_inputs:]) <= int(np.log2(target))
    s = choose_slices(net.path, sets, net.n_inputs, net.output, target_slices=8)
    assert len(s) == 3
    assert not set(s) & set(net.output)
    assert len(inputs) == net.n_inputs


def path_sets_of(net):
    from tedq_amd.tn.path import path_sets
    return path_sets([list(t) for t in net.lay[:net.n_inputs]], net.output, net.path)


@pytest.mark.parametrize("seed", range(4))
def test_diagonal_reduction_keeps_values_and_gradients(seed):
    """tn_simplify: diagonal gates (RZ, PhaseShift, CRZ, controlled phase, CZ, Z, S, T, a constant diagonal
    Unitary, diagonal observables) are stored as diagonals and their wires merged into hyper-indices, and
    the control wires of controlled gates (CNOT, CRX, CRY, Toffoli, CSWAP) are merged the same way;
    values and gradients equal the unsimplified network's and the oracle's, for expectation values and
    probabilities (output indices keep their names)."""
    rng = np.random.default_rng(100 + seed)
    n = 5
    diag_u = np.diag(np.exp(1j * rng.uniform(0, 6, 4)))
    names = ["RX", "RZ", "PhaseShift", "CRZ", "ControlledPhaseShift", "CZ", "PauliZ", "S", "T", "Hadamard",
             "CNOT", "RY", "U2d", "CRX", "CRY", "Toffoli", "CSWAP"]
    npar = {"RX": 1, "RZ": 1, "PhaseShift": 1, "CRZ": 1, "ControlledPhaseShift": 1, "RY": 1, "CRX": 1, "CRY": 1}
    spec = []
    for _ in range(24):
        name = names[rng.integers(len(names))]
        k = (3 if name in ("Toffoli", "CSWAP") else
             2 if name in ("CRZ", "ControlledPhaseShift", "CZ", "CNOT", "U2d", "CRX", "CRY") else 1)
        spec.append((name, [int(q) for q in rng.choice(n, size=k, replace=False)]))
    P = sum(npar.get(nm, 0) for nm, _ in spec)
    q0, q1 = [int(x) for x in rng.choice(n, size=2, replace=False)]
    kind = seed % 2

    def cd(*p):
        it = iter(p)
        for name, qs in spec:
            if name == "U2d":
                qai.Unitary(diag_u, qubits=qs)
            else:
                getattr(qai, name)(*[next(it) for _ in range(npar.get(name, 0))], qubits=qs)
        if kind == 0:
            return [qai.expval(qai.PauliZ(qubits=[q0])), qai.expval([qai.PauliZ(qubits=[q0]), qai.PauliX(qubits=[q1])])]
        return [qai.probs(qubits=sorted([q0, q1]))]

    theta = rng.uniform(0, 2 * np.pi, size=max(P, 1))[:P]
    rows = torch.tensor(np.stack([theta, theta[::-1].copy()]), dtype=torch.float32)
    res = {}
    for simplify in (True, False):
        cc = qai.Circuit(cd, n, *[torch.tensor(t) for t in theta]).compilecircuit(
            backend="pytorch", tn_mode=True, tn_simplify=simplify)
        ex = install(TNExecutor(cc, diagonal_reduction=simplify))
        if simplify:
            assert sum(net.n_diagonal for net in ex.networks) > 0
            # every controlled gate's leaf lost its control pairs
            for idx, op in cc._operators.items():  # pylint: disable=protected-access
                name = op.name[:-4] if getattr(op, "is_inverse", False) else op.name
                if name in ("CNOT", "CRX", "CRY", "CSWAP"):
                    assert ex._diag_ops[idx] == 1  # pylint: disable=protected-access
                elif name == "Toffoli":
                    assert ex._diag_ops[idx] == 2  # pylint: disable=protected-access
        ups = [torch.ones(2) if s.kind == "fexp" else torch.ones(2, 4) for s in ex.specs]
        res[simplify] = ex.value_and_grad(rows, ups)
    (v1, g1), (v0, g0) = res[True], res[False]
    for a, b in zip(v1, v0):
        assert np.max(np.abs(a.numpy() - b.numpy())) < 2e-6
    if P:
        assert np.max(np.abs(g1.numpy() - g0.numpy())) < 2e-6
    th = torch.tensor(rows[0].numpy(), dtype=torch.float64, requires_grad=True)
    it = iter(th)
    gates = []
    for name, qs in spec:
        if name == "U2d":
            gates.append(("Unitary", qs, diag_u))
        else:
            gates.append((name, qs, [next(it) for _ in range(npar.get(name, 0))]))
    ms = ([("expval", ("PauliZ", [q0])), ("expval", [("PauliZ", [q0]), ("PauliX", [q1])])] if kind == 0
          else [("probs", sorted([q0, q1]))])
    out = O.measure(O.propagate(n, gates), n, ms)
    for o, r in zip(v1, out):
        assert np.max(np.abs(o[0].numpy().reshape(-1) - r.detach().numpy().reshape(-1))) < 2e-5
    if P:
        (gref,) = torch.autograd.grad(sum(r.sum() for r in out), th)
        assert np.max(np.abs(g1[0].numpy() - gref.numpy())) < 1e-4


def test_diagonal_reduction_is_chosen_by_the_device_model():
    """With tn_simplify (default) the executor searches paths with and without diagonal reduction and
    keeps the variant of lower modelled device time; tn_simplify=False never reduces."""
    from tedq_amd import workloads as W
    ex = W.disordered_chain(16, 3, 8, tn_mode=True).executor()
    rep = ex.simplify_report
    assert rep is not None and ex.simplify == (rep["model_s_with"] < rep["model_s_without"])
    assert (sum(net.n_diagonal for net in ex.networks) > 0) == ex.simplify
    off = W.disordered_chain(16, 3, 8, tn_mode=True, tn_simplify=False).executor()
    assert not off.simplify and off.simplify_report is None
