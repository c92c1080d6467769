"""Contraction mode on the GPU: the step kernels, whole networks and their
gradients against the complex128 oracle and against the state-vector engine.

Tolerances (complex64 leaves and intermediates, fp32 accumulation):
* one contraction step: 2e-5 relative to the largest output magnitude (k up to 2^16);
* expectation values, probabilities and gradients against the complex128
  oracle, and contraction mode against the state-vector engine on the same
  circuit: 1e-5 relative to the largest magnitude (max|x - ref| / max|ref|).
Contraction paths are this build's own (the reference's optimisers are not
available), so parity is on values, not on path or intermediate order.
"""

import ctypes

import numpy as np
import pytest
import torch

import tedq_amd as qai
from oracle import sv_oracle as O
from tedq_amd import workloads as W
from tedq_amd.engine import hip
from tedq_amd.tn.executor import TNExecutor, make_desc

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda", 0)
TOL = 1e-5


def relerr(x, ref):
    x, ref = np.asarray(x), np.asarray(ref)
    return float(np.max(np.abs(x - ref)) / max(np.max(np.abs(ref)), 1e-12))


def _run_step(la, lb, lc, A, B, batch, shared_b=False, gemm_form=0):
    desc = make_desc(la, lb, lc, gemm_form)
    dd = torch.from_numpy(desc.copy()).to(DEV)
    lib = hip.load()
    hdesc = desc.ctypes.data_as(hip.c_vp)
    ws = lib.tedq_tn_workspace_bytes(hdesc, batch)
    assert ws >= 0
    wst = torch.empty(max(ws, 1), dtype=torch.uint8, device=DEV)
    a = torch.from_numpy(A.astype(np.complex64)).to(DEV)
    b = torch.from_numpy(B.astype(np.complex64)).to(DEV)
    c = torch.empty((batch, 2 ** len(lc)), dtype=torch.complex64, device=DEV)
    hip.call("tedq_tn_contract", hip.ptr(a), a.shape[1], hip.ptr(b), 0 if shared_b else b.shape[1], hip.ptr(c),
             c.shape[1], batch, hdesc, ctypes.c_void_p(dd.data_ptr()), hip.ptr(wst), ws,
             hip.stream_handle(DEV))
    torch.cuda.synchronize()
    return c.cpu().numpy(), ws


CASES = [
    # (A indices, B indices, C indices): small, GEMM-sized, split-k inner product, Hadamard index, outer product
    ("abcd", "cdef", "abef"),
    ("abcdefgh" + "ijkl", "ijkl" + "mnopqrst", "abcdefghmnopqrst"),
    ("abcdefghijklmnop", "abcdefghijklmnop", ""),
    ("xabc", "xbcd", "xad"),
    ("ab", "cd", "acbd"),
    ("qabcdefgh", "hgfedcbaz", "zq"),
    ("abcdefghijklmnopqr", "rqponmlkjihgfedcba", ""),
    # GEMM-sized with n only 64 wide (64 x 64 tiles), with a kept-shared index (128 x 128 tiles),
    # and with summed indices scattered through both operands' addresses
    ("abcdefgh" + "ijklmn", "ijklmn" + "opqrst", "abcdefghopqrst"),
    ("x" + "abcdefg" + "ijkl", "x" + "ijkl" + "mnopqrs", "mnopqrsxabcdefg"),
    ("iabcjdefghkl", "mnopiqrjklst", "abcdefghmnopqrst"),
    # gate-environment shapes (few outputs, long sums): plain, with a kept-shared index, and with the
    # summed indices in opposite orders and the kept ones in the middle of the operands
    ("ab" + "cdefghijklmnop", "qr" + "cdefghijklmnop", "abqr"),
    ("x" + "a" + "bcdefghijklm", "x" + "n" + "bcdefghijklm", "nxa"),
    ("cdaefbghijklmnopqrst", "tsrqponmlkjihgfedc" + "xy", "abxy"),
    # gates on a large tensor: two-qubit gate (A large), one-qubit gate with the large tensor second,
    # gate indices in the middle of both layouts, and a large operand with the gate's indices at both ends
    ("abcdefghijkl" + "xy", "xyuv", "abcdefghijkl" + "uv"),
    ("xy", "x" + "abcdefghijkl", "y" + "abcdefghijkl"),
    ("abcdexfghijkl", "xu", "abcudefghijkl"),
    ("uvxy", "abcdefxghijkly", "uabcdefghijklv"),
    # a tensor times a scalar into another layout (tiled transpose): reversed, low and high bits swapped,
    # and the scalar first
    ("abcdefghijklmn", "", "nmlkjihgfedcba"),
    ("abcdefghijklmn", "", "jklmnabcdefghi"),
    ("", "abcdefghijklm", "dgmbjelahcfik"),
    # a 64-wide GEMM with few tiles and a long sum (k split over 16 workgroups per tile)
    ("abcdef" + "ghijklmnop", "ghijklmnop" + "qrstuvw", "abcdefqrstuvw"),
    # one 128x128 tile and a long sum (k split over 64 workgroups)
    ("abcdefg" + "hijklmnopqrs", "hijklmnopqrs" + "tuvwxyz", "abcdefgtuvwxyz"),
]


@pytest.mark.parametrize("case", range(len(CASES)))
def test_contract_step_matches_einsum(case):
    la, lb, lc = (list(s) for s in CASES[case])
    rng = np.random.default_rng(case)
    batch = 3
    shared_b = case == 4
    A = rng.normal(size=(batch, 2 ** len(la))) + 1j * rng.normal(size=(batch, 2 ** len(la)))
    B = rng.normal(size=(1 if shared_b else batch, 2 ** len(lb))) + 1j * rng.normal(
        size=(1 if shared_b else batch, 2 ** len(lb)))
    got, _ = _run_step(la, lb, lc, A, B, batch, shared_b)
    expr = "".join(la) + "," + "".join(lb) + "->" + "".join(lc)
    for i in range(batch):
        ref = np.einsum(expr, A[i].reshape([2] * len(la)), B[0 if shared_b else i].reshape([2] * len(lb)))
        ref = np.asarray(ref).reshape(-1)
        scale = max(np.max(np.abs(ref)), 1.0)
        assert np.max(np.abs(got[i] - ref)) / scale < 2e-5, (case, i)


def _chain_gates(n, steps, theta):
    zz, xy = W.exchange_unitaries()
    gates = [("PauliX", [q], []) for q in range(1, n, 2)]
    it = iter(theta)
    for _ in range(steps):
        for q in range(n):
            gates.append(("RZ", [q], [next(it)]))
        for q in range(n - 1):
            gates.append(("Unitary", [q, q + 1], zz))
            gates.append(("Unitary", [q, q + 1], xy))
    return gates


def test_chain_values_and_gradients_vs_oracle():
    n, steps, obs = 8, 3, 3
    cc = W.disordered_chain(n, steps, obs, tn_mode=True)
    ex = cc.executor()
    assert isinstance(ex, TNExecutor)
    rng = np.random.default_rng(2)
    h = rng.uniform(-5, 5, size=(4, n))
    rows = torch.tensor(W.chain_rows(h, steps), device=DEV)
    outs, grad = ex.value_and_grad(rows, [torch.ones(4, device=DEV)])
    for b in range(4):
        th = torch.tensor(rows[b].cpu().numpy(), dtype=torch.float64, requires_grad=True)
        ref = O.execute(n, _chain_gates(n, steps, list(th)), [("expval", ("PauliZ", [obs]))])
        ref.sum().backward()
        assert abs(float(outs[0][b]) - float(ref.detach())) < TOL
        assert relerr(grad[b].cpu().numpy(), th.grad.numpy()) < TOL


def test_small_steps_batched_per_level_match_single_launches():
    """The forward's gate-merging steps run one launch per tree level (tedq_tn_contract_small_batch);
    the result equals the one-launch-per-step forward to 2e-5 (different summation order only)."""
    cc = W.disordered_chain(16, 4, 8, tn_mode=True)
    ex = cc.executor()
    net = ex.networks[0]
    plan = ex._small_plan(net)  # pylint: disable=protected-access
    assert len(plan["offs"]) > len(net.steps) // 2 and len(plan["bounds"]) < len(plan["offs"])
    rng = np.random.default_rng(5)
    rows = torch.tensor(W.chain_rows(rng.uniform(-5, 5, size=(6, 16)), 4), device=DEV)
    with torch.no_grad():
        batched = ex.measure(ex.run_forward(rows))[0].cpu().numpy()
        ex.small_batch = False
        single = ex.measure(ex.run_forward(rows))[0].cpu().numpy()
    assert np.max(np.abs(batched - single)) < 2e-5


def test_chain_tn_equals_state_vector_engine():
    n, steps, obs = 20, 4, 10
    tn = W.disordered_chain(n, steps, obs, tn_mode=True).executor()
    sv = W.disordered_chain(n, steps, obs).executor()
    rng = np.random.default_rng(3)
    rows = torch.tensor(W.chain_rows(rng.uniform(-5, 5, size=(8, n)), steps), device=DEV)
    ups = [torch.ones(8, device=DEV)]
    tv, tg = tn.value_and_grad(rows, ups)
    sv_v, sv_g = sv.value_and_grad(rows, ups)
    assert relerr(tv[0].cpu().numpy(), sv_v[0].float().cpu().numpy()) < TOL
    assert relerr(tg.cpu().numpy(), sv_g.cpu().numpy()) < TOL


def test_c3_full_size_light_cone_property():
    """40-qubit chain, 8 steps, 32 disorder seeds: <Z_20> in [-1, 1], and equal to the same
    network on the first 29 qubits (every bond right of qubit 28 lies outside Z_20's light cone)."""
    rng = np.random.default_rng(2)
    h = rng.uniform(-5, 5, size=(32, 40))
    full = W.disordered_chain(40, 8, 20, tn_mode=True).executor()
    rows = torch.tensor(W.chain_rows(h, 8), device=DEV)
    with torch.no_grad():
        v = full.measure(full.run_forward(rows))[0].cpu().numpy()
    assert np.all(np.isfinite(v)) and np.all(np.abs(v) <= 1.0 + 1e-5)
    cut = W.disordered_chain(29, 8, 20, tn_mode=True).executor()
    rows29 = torch.tensor(W.chain_rows(h[:, :29], 8), device=DEV)
    with torch.no_grad():
        v29 = cut.measure(cut.run_forward(rows29))[0].cpu().numpy()
    assert np.max(np.abs(v - v29)) < 2e-5


def test_lattice_gradient_vs_param_shift_and_oracle():
    L, layers = 3, 2
    n = L * L
    cc = W.lattice(L, layers, tn_mode=True)
    ex = cc.executor()
    rng = np.random.default_rng(3)
    theta = rng.uniform(0, 2 * np.pi, size=(1, 2 * n * layers)).astype(np.float32)
    rows = torch.tensor(theta, device=DEV)
    ups = [torch.full((1,), 1.0 / n, device=DEV) for _ in range(n)]
    outs, grad = ex.value_and_grad(rows, ups)
    jac = cc.jacobian_param_shift(rows)                     # [outputs, P] from shifted contractions
    ps = (jac.sum(0) / n).cpu().numpy()
    assert relerr(grad[0].cpu().numpy(), ps) < TOL
    th = torch.tensor(theta[0], dtype=torch.float64, requires_grad=True)
    gates = []
    it = iter(th)
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
    ref = O.execute(n, gates, [("expval", ("PauliZ", [q])) for q in range(n)])
    (ref.sum() / n).backward()
    got = np.array([float(o[0]) for o in outs])
    assert relerr(got, ref.detach().numpy()) < TOL
    assert relerr(grad[0].cpu().numpy(), th.grad.numpy()) < TOL


def test_compiled_call_autograd_in_contraction_mode():
    """The drop-in call path: compiled(*params) with back_prop in contraction mode."""
    n = 5

    def cd(*p):
        qai.RX(p[0], qubits=[0])
        qai.CRZ(p[1], qubits=[0, 3])
        qai.Rot(p[2], p[3], p[4], qubits=[3])
        qai.CNOT(qubits=[3, 1])
        qai.Toffoli(qubits=[0, 1, 4])
        return [qai.expval(qai.PauliZ(qubits=[1])), qai.expval(qai.PauliX(qubits=[3]))]

    p = torch.tensor([0.3, 1.1, -0.4, 0.9, 2.0], device=DEV, requires_grad=True)
    cc = qai.Circuit(cd, n, *torch.zeros(5)).compilecircuit(backend="pytorch", tn_mode=True)
    out = cc(p)
    out.sum().backward()
    th = torch.tensor(p.detach().cpu().numpy(), dtype=torch.float64, requires_grad=True)
    gates = [("RX", [0], [th[0]]), ("CRZ", [0, 3], [th[1]]), ("Rot", [3], [th[2], th[3], th[4]]),
             ("CNOT", [3, 1], []), ("Toffoli", [0, 1, 4], [])]
    ref = O.execute(n, gates, [("expval", ("PauliZ", [1])), ("expval", ("PauliX", [3]))])
    ref.sum().backward()
    assert relerr(out.detach().cpu().numpy(), ref.detach().numpy()) < TOL
    assert relerr(p.grad.cpu().numpy(), th.grad.numpy()) < TOL


def test_external_linear_path_on_the_full_network():
    """Paths found elsewhere (opt_einsum-style linear, one per measurement) drive the contraction."""
    n, steps, obs = 8, 2, 3
    cc0 = W.disordered_chain(n, steps, obs, tn_mode=True)
    nt = cc0._tensor_networks[0].num_tensors  # pylint: disable=protected-access
    cc = W.disordered_chain(n, steps, obs, tn_mode=True, tn_paths=[[(0, 1)] * (nt - 1)])
    ex = cc.executor()
    assert ex.path_report()[0]["strategy"] == "external"
    rng = np.random.default_rng(8)
    rows = torch.tensor(W.chain_rows(rng.uniform(-5, 5, size=(3, n)), steps), device=DEV)
    outs, grad = ex.value_and_grad(rows, [torch.ones(3, device=DEV)])
    for b in range(3):
        th = torch.tensor(rows[b].cpu().numpy(), dtype=torch.float64, requires_grad=True)
        ref = O.execute(n, _chain_gates(n, steps, list(th)), [("expval", ("PauliZ", [obs]))])
        ref.sum().backward()
        assert abs(float(outs[0][b]) - float(ref.detach())) < TOL
        assert relerr(grad[b].cpu().numpy(), th.grad.numpy()) < TOL


@pytest.mark.parametrize("kind", ["probs", "state", "hermitian"])
def test_probs_state_and_matrix_observables(kind):
    """Contraction mode for the other measurement kinds: probabilities (ket/bra diagonal outputs),
    the state (ket only, complex output), and a user Hermitian matrix; values and gradients."""
    n = 5
    herm = np.array([[0.3, 0.2 - 0.5j], [0.2 + 0.5j, -0.8]])

    def cd(*p):
        qai.RX(p[0], qubits=[0])
        qai.CRY(p[1], qubits=[0, 3])
        qai.Rot(p[2], p[3], p[4], qubits=[2])
        qai.CNOT(qubits=[2, 4])
        qai.CZ(qubits=[4, 1])
        qai.RY(p[5], qubits=[1])
        if kind == "probs":
            return qai.probs(qubits=[1, 3])
        if kind == "state":
            return qai.state()
        return qai.expval(qai.Unitary(herm, qubits=[4]))

    theta = np.array([0.4, -1.1, 0.9, 2.2, -0.6, 1.3])
    cc = qai.Circuit(cd, n, *([0.1] * 6)).compilecircuit(backend="pytorch", tn_mode=True)
    p = torch.tensor(theta, dtype=torch.float32, device=DEV, requires_grad=True)
    out = cc(p)
    th = torch.tensor(theta, dtype=torch.float64, requires_grad=True)
    gates = [("RX", [0], [th[0]]), ("CRY", [0, 3], [th[1]]), ("Rot", [2], [th[2], th[3], th[4]]),
             ("CNOT", [2, 4], []), ("CZ", [4, 1], []), ("RY", [1], [th[5]])]
    if kind == "probs":
        ref = O.execute(n, gates, [("probs", [1, 3])])
        w = torch.arange(4, dtype=torch.float64).reshape(1, 2, 2) + 1.0
    elif kind == "state":
        ref = O.execute(n, gates, [("state",)])
        w = None
    else:
        ref = O.execute(n, gates, [("expval", (herm, [4]))])
        w = torch.ones(1, dtype=torch.float64)
    got = out.detach().cpu().numpy()
    assert relerr(got, ref.detach().numpy()) < TOL
    if kind == "state":
        # a real loss of the complex output: sum of Re(psi) weighted + |psi_3|^2
        lw = torch.linspace(-1, 1, 2 ** n, dtype=torch.float64).reshape(out.shape)
        (out.real * lw.to(DEV, torch.float32)).sum().backward()
        (ref.real * lw).sum().backward()
    else:
        (out * w.to(DEV, torch.float32)).sum().backward()
        (ref * w).sum().backward()
    assert relerr(p.grad.cpu().numpy(), th.grad.numpy()) < TOL


def test_gemm_three_product_form_matches():
    """The 3-product (Gauss) complex GEMM and the 64 x 64-tile form, selected per step in the descriptor,
    give the same GEMM-shaped contraction as the default 4-product form and einsum (2e-5 relative)."""
    from tedq_amd.tn.executor import GEMM_FORMS
    rng = np.random.default_rng(1)
    la, lb, lc = list("abcdefghijkl"), list("ijklmnopqrst"), list("abcdefghmnopqrst")
    A = rng.normal(size=(2, 4096)) + 1j * rng.normal(size=(2, 4096))
    B = rng.normal(size=(2, 4096)) + 1j * rng.normal(size=(2, 4096))
    ref = [np.einsum("abcdefghijkl,ijklmnopqrst->abcdefghmnopqrst", A[i].reshape([2] * 12),
                     B[i].reshape([2] * 12)).reshape(-1) for i in range(2)]
    got = {}
    for name, form in GEMM_FORMS.items():
        got[name], _ = _run_step(la, lb, lc, A, B, 2, gemm_form=form)
        err = max(np.max(np.abs(got[name][i] - ref[i])) / np.max(np.abs(ref[i])) for i in range(2))
        print(f"gemm form {name}: relerr {err:.2e}", flush=True)
        assert err < 2e-5
    assert np.max(np.abs(got["3m"] - got["4m"])) > 0  # a different kernel really ran


def test_fused_gate_backward_matches_two_environment_steps():
    """tedq_tn_gate_backward (both environments of a gate step from one read of E_C) gives the same
    gradient as the two separate environment contractions (1e-5) and the oracle's (1e-5 relative), on a 3x4
    lattice of 4 layers whose paths apply gates to rank-12 intermediates."""
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

    rng = np.random.default_rng(9)
    theta = rng.uniform(0, 2 * np.pi, size=(3, 2 * n * layers))
    # without diagonal reduction, so that RZ / CZ stay gates and the large steps are plain gate steps
    cc = qai.Circuit(cd, n, *[torch.tensor(t) for t in theta[0]]).compilecircuit(backend="pytorch", tn_mode=True,
                                                                                 tn_simplify=False)
    ex = TNExecutor(cc)
    assert sum(_gate_step(ex._descs[st.fwd]) for st in ex.networks[0].steps) > 0  # pylint: disable=protected-access
    rows = torch.tensor(theta, dtype=torch.float32, device=DEV)
    ups = [torch.ones(3, device=DEV)]
    _, g_fused = ex.value_and_grad(rows, ups)
    ex.fused_backward = False
    _, g_split = ex.value_and_grad(rows, ups)
    assert float((g_fused - g_split).abs().max()) < 1e-5
    for b in range(3):
        th = torch.tensor(rows[b].cpu().numpy(), dtype=torch.float64, requires_grad=True)
        gates, it = [], iter(th)
        for _ in range(layers):
            for q in range(n):
                gates += [("RX", [q], [next(it)]), ("RZ", [q], [next(it)])]
            gates += [("CZ", [a, bb], []) for a, bb in bonds]
        O.execute(n, gates, [("expval", ("PauliZ", [obs]))]).sum().backward()
        assert relerr(g_fused[b].cpu().numpy(), th.grad.numpy()) < TOL


def _lattice_oracle_gates(L, layers, theta):
    n = L * L
    gates = []
    it = iter(theta)
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


def test_checkpointed_backward_segments_vs_oracle():
    """The reverse contraction tree with checkpointed segments (boundary tensors stored, the inside of
    every segment recomputed) on the HIP kernels: a keep budget far below the network's intermediates
    forces several segments; values and gradients equal the unsegmented run and the oracle."""
    L, layers = 3, 3
    n = L * L
    cc = W.lattice(L, layers, tn_mode=True)
    whole = cc.executor()
    seg = TNExecutor(cc, max_bytes=1 << 12)
    counts = [len(seg._segments(net, 2, DEV)) for net in seg.networks]  # pylint: disable=protected-access
    print(f"segments per network: {counts}", flush=True)
    assert min(counts) > 1
    rng = np.random.default_rng(5)
    theta = rng.uniform(0, 2 * np.pi, size=(2, 2 * n * layers)).astype(np.float32)
    rows = torch.tensor(theta, device=DEV)
    ups = [torch.full((2,), 1.0 / n, device=DEV) for _ in range(n)]
    v_w, g_w = whole.value_and_grad(rows, ups)
    v_s, g_s = seg.value_and_grad(rows, ups)
    assert np.max(np.abs(torch.stack(v_w).cpu().numpy() - torch.stack(v_s).cpu().numpy())) < 1e-6
    assert np.max(np.abs(g_w.cpu().numpy() - g_s.cpu().numpy())) < 1e-6
    th = torch.tensor(theta[1], dtype=torch.float64, requires_grad=True)
    ref = O.execute(n, _lattice_oracle_gates(L, layers, th), [("expval", ("PauliZ", [q])) for q in range(n)])
    (ref.sum() / n).backward()
    got = np.array([float(v[1]) for v in v_s])
    e_v = np.max(np.abs(got - ref.detach().numpy()))
    e_g = np.max(np.abs(g_s[1].cpu().numpy() - th.grad.numpy())) / np.max(np.abs(th.grad.numpy()))
    print(f"segmented backward: values abs err {e_v:.2e}, gradients relerr {e_g:.2e}", flush=True)
    assert e_v < 2e-5 and e_g < 1e-5


def test_c4_full_size_lattice_tn_vs_state_vector_engine():
    """The 5x5 lattice of 6 layers at its stated size: all 25 <Z_i> and the 300 gradients of their mean
    from the contraction path (reverse tree) against the state-vector engine (25 qubits, adjoint
    sweep), and 24 of the gradients against the parameter-shift rule evaluated as batched rows of the
    state-vector engine."""
    L, layers, n = 5, 6, 25
    rng = np.random.default_rng(3)
    theta = rng.uniform(0, 2 * np.pi, size=(1, 2 * n * layers)).astype(np.float32)
    rows = torch.tensor(theta, device=DEV)
    ups = [torch.full((1,), 1.0 / n, device=DEV) for _ in range(n)]
    sv = W.lattice(L, layers)
    v_sv, g_sv = sv.executor().value_and_grad(rows, ups)
    print("state-vector engine done", flush=True)
    ex = W.lattice(L, layers, tn_mode=True).executor()
    print("contraction paths built", flush=True)
    v_tn, g_tn = ex.value_and_grad(rows, ups)
    torch.cuda.synchronize()
    vs = np.array([float(v[0]) for v in v_sv])
    vt = np.array([float(v[0]) for v in v_tn])
    e_v = np.max(np.abs(vs - vt))
    e_g = np.max(np.abs(g_sv[0].cpu().numpy() - g_tn[0].cpu().numpy())) / np.max(np.abs(g_sv[0].cpu().numpy()))
    print(f"C4 full size: values abs diff {e_v:.2e}, gradients rel diff {e_g:.2e}", flush=True)
    assert e_v < 2e-5 and e_g < 1e-5
    # parameter shift (RX / RZ: two terms, +-pi/2) on every 12th parameter, all shifted rows in one launch
    cols = list(range(0, 2 * n * layers, 12))
    shifted = []
    for p in cols:
        for s in (np.pi / 2, -np.pi / 2):
            r = theta[0].astype(np.float64).copy()
            r[p] += s
            shifted.append(r)
    with torch.no_grad():
        out = sv.executor().measure(sv.executor().run_forward(torch.tensor(np.array(shifted), device=DEV,
                                                                           dtype=torch.float64)))
    mean = torch.stack(out, 1).double().mean(1).cpu().numpy()
    ps = 0.5 * (mean[0::2] - mean[1::2])
    e_ps = np.max(np.abs(ps - g_tn[0].cpu().numpy()[cols])) / np.max(np.abs(g_sv[0].cpu().numpy()))
    print(f"C4 parameter shift vs contraction gradients: rel diff {e_ps:.2e}", flush=True)
    assert e_ps < 1e-5


def test_sliced_contraction_on_the_device():
    """slicing_opts={"target_size": 2^8} on a 3x4 lattice of 3 layers (<Z_5>, <Z_6>): every slice runs
    the frozen sliced path on the HIP kernels with its selector vectors, and the summed values and
    gradients equal the unsliced run (2e-6) and the oracle (1e-5 relative)."""
    L, layers, obs = 3, 3, [5, 6]
    n = 12

    def lat(**kw):
        def cd(*p):
            it = iter(p)
            for _ in range(layers):
                for q in range(n):
                    qai.RX(next(it), qubits=[q])
                    qai.RZ(next(it), qubits=[q])
                for r in range(L):
                    for c in range(3):
                        qai.CZ(qubits=[r * 4 + c, r * 4 + c + 1])
                for r in range(L - 1):
                    for c in range(4):
                        qai.CZ(qubits=[r * 4 + c, (r + 1) * 4 + c])
            return [qai.expval(qai.PauliZ(qubits=[q])) for q in obs]
        return qai.Circuit(cd, n, *([0.1] * (2 * n * layers))).compilecircuit(backend="pytorch", tn_mode=True, **kw)
    full = lat().executor()
    sliced = lat(hyper_opt={"slicing_opts": {"target_size": 2 ** 8}}).executor()
    print("sliced indices per network:", [len(net.sliced) for net in sliced.networks],
          "largest rank", [r["largest_rank"] for r in full.path_report()], "->",
          [r["largest_rank"] for r in sliced.path_report()], flush=True)
    assert all(net.sliced for net in sliced.networks)
    rng = np.random.default_rng(8)
    theta = rng.uniform(0, 2 * np.pi, size=(2, 2 * n * layers)).astype(np.float32)
    rows = torch.tensor(theta, device=DEV)
    ups = [torch.full((2,), 0.5, device=DEV) for _ in obs]
    v_f, g_f = full.value_and_grad(rows, ups)
    v_s, g_s = sliced.value_and_grad(rows, ups)
    assert np.max(np.abs(torch.stack(v_f).cpu().numpy() - torch.stack(v_s).cpu().numpy())) < 2e-6
    assert np.max(np.abs(g_f.cpu().numpy() - g_s.cpu().numpy())) < 2e-6
    with torch.no_grad():
        v_fw = sliced.measure(sliced.run_forward(rows))
    assert np.max(np.abs(torch.stack(v_fw).cpu().numpy() - torch.stack(v_f).cpu().numpy())) < 2e-6
    th = torch.tensor(theta[0], dtype=torch.float64, requires_grad=True)
    gates = []
    it = iter(th)
    for _ in range(layers):
        for q in range(n):
            gates.append(("RX", [q], [next(it)]))
            gates.append(("RZ", [q], [next(it)]))
        for r in range(L):
            for c in range(3):
                gates.append(("CZ", [r * 4 + c, r * 4 + c + 1], []))
        for r in range(L - 1):
            for c in range(4):
                gates.append(("CZ", [r * 4 + c, (r + 1) * 4 + c], []))
    ref = O.execute(n, gates, [("expval", ("PauliZ", [q])) for q in obs])
    (ref.sum() / 2).backward()
    got = np.array([float(v[0]) for v in v_s])
    e_v, e_g = relerr(got, ref.detach().numpy()), relerr(g_s[0].cpu().numpy(), th.grad.numpy())
    print(f"sliced vs oracle: values relerr {e_v:.2e}, gradients relerr {e_g:.2e}", flush=True)
    assert e_v < TOL and e_g < TOL


def test_diagonal_reduction_on_the_device():
    """Diagonal reduction forced on (RZ and CZ stored as diagonals, their wires merged into hyper-indices)
    gives the unreduced run's values and gradients (2e-6) and the oracle's (1e-5 relative) on a 3x4
    lattice of 3 layers, <Z_5> and <Z_6>."""
    L, layers, obs = 3, 3, [5, 6]
    n = 12

    def cd(*p):
        it = iter(p)
        for _ in range(layers):
            for q in range(n):
                qai.RX(next(it), qubits=[q])
                qai.RZ(next(it), qubits=[q])
            for r in range(L):
                for c in range(3):
                    qai.CZ(qubits=[r * 4 + c, r * 4 + c + 1])
            for r in range(L - 1):
                for c in range(4):
                    qai.CZ(qubits=[r * 4 + c, (r + 1) * 4 + c])
        return [qai.expval(qai.PauliZ(qubits=[q])) for q in obs]
    cc = qai.Circuit(cd, n, *([0.1] * (2 * n * layers))).compilecircuit(backend="pytorch", tn_mode=True)
    red = TNExecutor(cc, diagonal_reduction=True)
    plain = TNExecutor(cc, diagonal_reduction=False)
    assert sum(net.n_diagonal for net in red.networks) > 0
    rng = np.random.default_rng(17)
    theta = rng.uniform(0, 2 * np.pi, size=(2, 2 * n * layers)).astype(np.float32)
    rows = torch.tensor(theta, device=DEV)
    ups = [torch.full((2,), 0.5, device=DEV) for _ in obs]
    v_r, g_r = red.value_and_grad(rows, ups)
    v_p, g_p = plain.value_and_grad(rows, ups)
    assert np.max(np.abs(torch.stack(v_r).cpu().numpy() - torch.stack(v_p).cpu().numpy())) < 2e-6
    assert np.max(np.abs(g_r.cpu().numpy() - g_p.cpu().numpy())) < 2e-6
    th = torch.tensor(theta[1], dtype=torch.float64, requires_grad=True)
    gates = []
    it = iter(th)
    for _ in range(layers):
        for q in range(n):
            gates.append(("RX", [q], [next(it)]))
            gates.append(("RZ", [q], [next(it)]))
        for r in range(L):
            for c in range(3):
                gates.append(("CZ", [r * 4 + c, r * 4 + c + 1], []))
        for r in range(L - 1):
            for c in range(4):
                gates.append(("CZ", [r * 4 + c, (r + 1) * 4 + c], []))
    ref = O.execute(n, gates, [("expval", ("PauliZ", [q])) for q in obs])
    (ref.sum() / 2).backward()
    e_v = relerr(np.array([float(v[1]) for v in v_r]), ref.detach().numpy())
    e_g = relerr(g_r[1].cpu().numpy(), th.grad.numpy())
    print(f"diagonal reduction vs oracle: values relerr {e_v:.2e}, gradients relerr {e_g:.2e}", flush=True)
    assert e_v < TOL and e_g < TOL


def test_control_wire_reduction_on_the_device():
    """Control wires of CNOT, CRX, CRY, Toffoli and CSWAP merged into hyper-indices (leaves stored as
    [controls][out][in]) next to diagonal RZ / CRZ: values and gradients equal the unreduced run's (2e-6)
    and the oracle's (1e-5 relative) on a 10-qubit circuit with two observables."""
    n = 10
    rng = np.random.default_rng(29)
    names = ["RY", "RZ", "CNOT", "CRX", "CRY", "CRZ", "Toffoli", "CSWAP", "Hadamard"]
    npar = {"RY": 1, "RZ": 1, "CRX": 1, "CRY": 1, "CRZ": 1}
    spec = [("RY", [q]) for q in range(n)]
    for _ in range(50):
        name = names[rng.integers(len(names))]
        k = 3 if name in ("Toffoli", "CSWAP") else 2 if name.startswith("C") else 1
        spec.append((name, [int(q) for q in rng.choice(n, size=k, replace=False)]))
    P = sum(npar.get(nm, 0) for nm, _ in spec)
    obs = [2, 7]

    def cd(*p):
        it = iter(p)
        for name, qs in spec:
            getattr(qai, name)(*[next(it) for _ in range(npar.get(name, 0))], qubits=qs)
        return [qai.expval(qai.PauliZ(qubits=[q])) for q in obs]
    cc = qai.Circuit(cd, n, *([0.1] * P)).compilecircuit(backend="pytorch", tn_mode=True)
    red = TNExecutor(cc, diagonal_reduction=True)
    plain = TNExecutor(cc, diagonal_reduction=False)
    assert any(0 < d < len(cc._operators[i].qubits) for i, d in red._diag_ops.items())  # pylint: disable=protected-access
    theta = rng.uniform(0, 2 * np.pi, size=(2, P)).astype(np.float32)
    rows = torch.tensor(theta, device=DEV)
    ups = [torch.ones(2, device=DEV) for _ in obs]
    v_r, g_r = red.value_and_grad(rows, ups)
    v_p, g_p = plain.value_and_grad(rows, ups)
    assert np.max(np.abs(torch.stack(v_r).cpu().numpy() - torch.stack(v_p).cpu().numpy())) < 2e-6
    assert np.max(np.abs(g_r.cpu().numpy() - g_p.cpu().numpy())) < 2e-6
    th = torch.tensor(theta[1], dtype=torch.float64, requires_grad=True)
    it = iter(th)
    gates = [(name, qs, [next(it) for _ in range(npar.get(name, 0))]) for name, qs in spec]
    ref = O.execute(n, gates, [("expval", ("PauliZ", [q])) for q in obs])
    ref.sum().backward()
    got = np.array([float(v[1]) for v in v_r])
    assert relerr(got, ref.detach().numpy()) < 1e-5
    assert relerr(g_r[1].cpu().numpy(), th.grad.numpy()) < 1e-5
