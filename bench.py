"""Benchmark: circuit fwd+grad evaluations per second on one MI355X.

Workload (the batched state-vector configuration): hardware-efficient ansatz on
20 qubits, 10 layers of per-qubit RY(theta) RZ(phi) followed by the CNOT chain
0->1->...->19, 400 parameters, a batch of 256 parameter rows ~ U[0, 2pi) from a
fixed seed, outputs <Z_i> for all 20 qubits, forward plus adjoint gradient of
sum_i <Z_i> (upstream dy = ones).  One step = one batched forward + backward;
inputs are resident in HBM when the timed region starts.

Extra fields on the JSON line:
* roofline: the dominant pass kernel, algorithmic bytes per launch (16 B per
  amplitude for a forward pass, 32 B for an adjoint pass over psi and lambda,
  24 B when it reads psi from the forward's checkpoints, over the tiles the pass
  processes: from |0..0> the first passes skip the tiles that hold zeros, and a
  generated forward pass reads only amplitudes whose tile bits earlier passes
  held)
  divided by its mean launch time, measured with events on the launch stream
  during the timed steps, against 8 TB/s; traffic (FETCH_SIZE + WRITE_SIZE)
  and the VALU issue share come from the committed counter runs of the same
  shapes;
* cpu_baseline: the complex64 CPU restatement of the reference algorithm
  (tensordot + permute per gate, torch autograd backward) on the host cores,
  on a bounded sample of the same workload;
* plain_gpu: the same reference-style algorithm run on the GPU one parameter
  row per call (the unbatched single-circuit path).
"""

import argparse
import json
import os
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for _p in (ROOT, os.path.join(ROOT, "ted-q_amd")):
    if _p not in sys.path:
        sys.path.insert(0, _p)

import numpy as np  # noqa: E402
import torch  # noqa: E402

HBM_PEAK_GBS = 8000.0
_T0 = time.perf_counter()


def progress(msg):
    """One line on stderr per finished phase (the JSON result alone goes to stdout)."""
    print(f"[bench {time.perf_counter() - _T0:7.1f} s] {msg}", file=sys.stderr, flush=True)


class EventProbe:
    """Records a start/end event pair around every launch of a kernel kind."""

    def __init__(self):
        self.pairs = {}
        self._open = {}

    def begin(self, kind):
        ev = torch.cuda.Event(enable_timing=True)
        ev.record()
        self._open[kind] = ev

    def end(self, kind):
        ev = torch.cuda.Event(enable_timing=True)
        ev.record()
        self.pairs.setdefault(kind, []).append((self._open.pop(kind), ev))

    def mean_ms(self, kind):
        pairs = self.pairs.get(kind, [])
        if not pairs:
            return None, 0
        return sum(a.elapsed_time(b) for a, b in pairs) / len(pairs), len(pairs)


def hea_circuit(n, layers):
    import tedq_amd as qai  # pylint: disable=import-outside-toplevel

    def circuit_def(*p):
        it = iter(p)
        for _ in range(layers):
            for q in range(n):
                qai.RY(next(it), qubits=[q])
                qai.RZ(next(it), qubits=[q])
            for q in range(n - 1):
                qai.CNOT(qubits=[q, q + 1])
        return [qai.expval(qai.PauliZ(qubits=[q])) for q in range(n)]
    return qai.Circuit(circuit_def, n, *([0.1] * (2 * n * layers)))


# ---------------------------------------------------------------------------
# plain single-circuit GPU path: the reference algorithm (one tensordot +
# permute per gate, autograd backward) in complex64, one parameter row per call
# ---------------------------------------------------------------------------
def plain_eval(n, layers, theta_row, device):
    th = theta_row.reshape(layers, n, 2)
    state = torch.zeros([2] * n, dtype=torch.complex64, device=device)
    state.view(-1)[0] = 1
    cnot = torch.tensor([[1, 0, 0, 0], [0, 1, 0, 0], [0, 0, 0, 1], [0, 0, 1, 0]], dtype=torch.complex64,
                        device=device).reshape(2, 2, 2, 2)
    for layer in range(layers):
        for q in range(n):
            t = th[layer, q, 0]
            c, s = torch.cos(t / 2), torch.sin(t / 2)
            ry = torch.stack([c, -s, s, c]).to(torch.complex64).reshape(2, 2)
            p = th[layer, q, 1]
            e = torch.exp(-0.5j * p.to(torch.complex64))
            rz = torch.stack([e, torch.zeros_like(e), torch.zeros_like(e), e.conj()]).reshape(2, 2)
            for g in (ry, rz):
                state = torch.tensordot(g, state, ([1], [q]))
                state = state.permute(list(range(1, q + 1)) + [0] + list(range(q + 1, n)))
        for q in range(n - 1):
            state = torch.tensordot(cnot, state, ([2, 3], [q, q + 1]))
            order = [q, q + 1] + [x for x in range(n) if x not in (q, q + 1)]
            state = state.permute([order.index(i) for i in range(n)])
    probs = (state.abs() ** 2).reshape(-1)
    idx = torch.arange(2 ** n, device=device)
    outs = [(probs * (1 - 2 * ((idx >> (n - 1 - q)) & 1))).sum() for q in range(n)]
    return torch.stack(outs)


def time_plain_gpu(n, layers, theta, device, budget_s=8.0):
    """Rows per second of plain_eval + autograd backward, one row per call.

    Warmed with one single-layer call (library handles and every kernel shape
    the full call uses), then timed over whole calls until ``budget_s`` has
    passed.  Returns (evals/s, calls, forward s, backward s) of the timed calls."""
    rows = torch.tensor(theta, dtype=torch.float32, device=device)
    r = rows[0, :2 * n].clone().requires_grad_(True)
    plain_eval(n, 1, r, device).sum().backward()
    torch.cuda.synchronize()
    done, fwd_s, bwd_s = 0, 0.0, 0.0
    t0 = time.perf_counter()
    for i in range(rows.shape[0]):
        r = rows[i].clone().requires_grad_(True)
        ta = time.perf_counter()
        out = plain_eval(n, layers, r, device)
        torch.cuda.synchronize()
        tb = time.perf_counter()
        out.sum().backward()
        torch.cuda.synchronize()
        fwd_s += tb - ta
        bwd_s += time.perf_counter() - tb
        done += 1
        progress(f"plain single-circuit path: call {done} done (forward {tb - ta:.1f} s, "
                 f"backward {time.perf_counter() - tb:.1f} s)")
        if time.perf_counter() - t0 > budget_s:
            break
    return done / (time.perf_counter() - t0), done, fwd_s / done, bwd_s / done


def engine_batch1(compiled, rows, budget_s=5.0):
    """The engine's own single-circuit rate: one parameter row per execute_batch call (forward +
    adjoint gradient of sum_i <Z_i>), same compiled program and kernels as the batched run."""
    def one(i):
        r = rows[i:i + 1].detach().requires_grad_(True)
        compiled.execute_batch(r).sum().backward()
        return r.grad
    for i in range(3):
        one(i)
    torch.cuda.synchronize()
    done = 0
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < budget_s and done < rows.shape[0]:
        one(done)
        torch.cuda.synchronize()
        done += 1
    el = time.perf_counter() - t0
    return {"evals_per_s": done / el, "ms_per_eval": el / done * 1e3, "calls": done,
            "how": "one row per call, synchronised after each, generated kernels of the batched run"}


def cpu_baseline(n, layers, theta, budget_s=15.0, check=None):
    """The oracle's complex64 restatement on the host cores (bounded sample).

    ``check`` = (row indices, GPU values [r, n], GPU gradients [r, P]) of the
    same workload: those rows are also run through the oracle in complex128 and
    the relative errors max|x - x128| / max(max|x128|, 1e-12) are returned."""
    from oracle import sv_oracle as O  # pylint: disable=import-outside-toplevel
    # the host threads this process may use: the OMP_NUM_THREADS share when set, else its CPU affinity
    cores = int(os.environ.get("OMP_NUM_THREADS") or 0) or len(os.sched_getaffinity(0))
    torch.set_num_threads(cores)
    meas = [("expval", ("PauliZ", [q])) for q in range(n)]
    error = None
    if check is not None:
        idx, vals, grads = check
        ev = eg = 0.0
        for r, i in enumerate(idx):
            t = torch.tensor(theta[i].reshape(layers, n, 2), dtype=torch.float64, requires_grad=True)
            ref = O.execute(n, O.hea_gates(n, t)[0], meas)
            ref.sum().backward()
            rv, rg = ref.detach().numpy(), t.grad.reshape(-1).numpy()
            ev = max(ev, float(np.abs(vals[r] - rv).max() / max(np.abs(rv).max(), 1e-12)))
            eg = max(eg, float(np.abs(grads[r] - rg).max() / max(np.abs(rg).max(), 1e-12)))
        error = {"rows": [int(i) for i in idx], "values_rel": ev, "gradients_rel": eg,
                 "reference": "oracle in complex128 (torch autograd), same parameter rows"}
    done = 0
    t0 = time.perf_counter()
    for i in range(theta.shape[0]):
        t = torch.tensor(theta[i].reshape(layers, n, 2), dtype=torch.float32, requires_grad=True)
        out = O.execute(n, O.hea_gates(n, t)[0], meas, cdtype=torch.complex64)
        out.sum().backward()
        done += 1
        if time.perf_counter() - t0 > budget_s:
            break
    el = time.perf_counter() - t0
    model = ""
    try:
        model = [ln.split(":", 1)[1].strip() for ln in subprocess.run(["lscpu"], capture_output=True, text=True)
                 .stdout.splitlines() if ln.startswith("Model name")][0]
    except (OSError, IndexError):
        pass
    out = {"value": done / el, "unit": "evals/s", "cores": cores, "kind": "port",
           "sample": f"{done} of {theta.shape[0]} parameter rows, {n}q d{layers} fwd+grad, complex64 "
                     f"tensordot+permute per gate with torch autograd, {el:.1f} s",
           "cpu_model": model}
    if error is not None:
        out["error_vs_complex128"] = error
    return out


def hea_z0_circuit(n, layers):
    import tedq_amd as qai  # pylint: disable=import-outside-toplevel

    def circuit_def(*p):
        it = iter(p)
        for _ in range(layers):
            for q in range(n):
                qai.RY(next(it), qubits=[q])
                qai.RZ(next(it), qubits=[q])
            for q in range(n - 1):
                qai.CNOT(qubits=[q, q + 1])
        return qai.expval(qai.PauliZ(qubits=[0]))
    return qai.Circuit(circuit_def, n, *([0.1] * (2 * n * layers)))


def small_circuit_leg(dev, reps=50):
    """The reference's CPU-runnable case: 8-qubit ansatz, 4 layers, <Z_0>, one parameter row, fwd + grad,
    on the engine (one compiled call per evaluation) and in the oracle's complex64 restatement on the
    host (the reference algorithm, run in full)."""
    from oracle import sv_oracle as O  # pylint: disable=import-outside-toplevel
    n, L = 8, 4
    theta = np.random.default_rng(0).uniform(0, 2 * np.pi, size=(L, n, 2))
    compiled = hea_z0_circuit(n, L).compilecircuit(backend="pytorch")
    t = torch.tensor(theta.reshape(-1), dtype=torch.float32, device=dev)

    def gpu_eval():
        p = t.detach().requires_grad_(True)
        compiled(p).sum().backward()
        return p.grad
    for _ in range(5):
        gpu_eval()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        g_gpu = gpu_eval()
    torch.cuda.synchronize()
    gpu_ms = (time.perf_counter() - t0) / reps * 1e3

    def cpu_eval():
        p = torch.tensor(theta, dtype=torch.float32, requires_grad=True)
        out = O.execute(n, O.hea_gates(n, p)[0], [("expval", ("PauliZ", [0]))], cdtype=torch.complex64)
        out.sum().backward()
        return p.grad
    cpu_eval()
    t0 = time.perf_counter()
    for _ in range(reps):
        g_cpu = cpu_eval()
    cpu_ms = (time.perf_counter() - t0) / reps * 1e3
    diff = float((g_gpu.cpu().double() - g_cpu.reshape(-1).double()).abs().max())
    return {"workload": "hardware-efficient ansatz 8q 4 layers, <Z_0>, one parameter row, fwd + grad",
            "engine_ms_per_eval": gpu_ms, "cpu_port_ms_per_eval": cpu_ms, "grad_max_diff_vs_cpu_c64": diff}


def serving_stream(n_req, seed=4, rate=None):
    """Requests of the mixed stream: n ~ U{6..24}, layers ~ U{2..20}, params ~ U[0, 2pi)."""
    from tedq_amd.serving.engine import Request  # pylint: disable=import-outside-toplevel
    rng = np.random.default_rng(seed)
    ns = rng.integers(6, 25, size=n_req)
    ls = rng.integers(2, 21, size=n_req)
    t = np.cumsum(rng.exponential(1.0 / rate, size=n_req)) if rate else np.zeros(n_req)
    return [Request(i, (int(n), int(L)), rng.uniform(0, 2 * np.pi, 2 * int(n) * int(L)).astype(np.float32),
                    float(t[i])) for i, (n, L) in enumerate(zip(ns, ls))]


def serving_leg(n_req, warm, ragged=True, common_rate=None):
    """Closed-loop capacity, Poisson stream at 0.7x capacity (p50/p99), and the unbatched path.

    With ``ragged`` the engine batches requests of different structures that share a tile shape into
    one launch per pass index; without it only identical structures share launches.  ``common_rate``:
    also serve a Poisson stream at that rate (the top-level latency figures then are that stream's,
    and the 0.7x-capacity stream's go under ``high_load``)."""
    from tedq_amd.serving.engine import ServingEngine, latency_stats  # pylint: disable=import-outside-toplevel
    compiled = {}

    def factory(key):
        if key not in compiled:
            compiled[key] = hea_z0_circuit(*key).compilecircuit(backend="pytorch")
        return compiled[key]

    eng = ServingEngine(factory, hbm_budget=64 << 30, max_batch=64, n_streams=4, max_inflight=8, ragged=ragged)
    closed = serving_stream(n_req)
    keys = sorted({r.key for r in closed} | {r.key for r in serving_stream(n_req, seed=5, rate=1.0)})
    eng.prepare(keys, buckets=(1, 2, 4), log=progress)
    # every request simulates the backward light cone of Z_0 only (trailing gates folded into the
    # observable, qubits left without gates dropped; exact)
    sim_n = float(np.mean([factory(r.key).executor().n for r in closed]))
    wide_n = float(np.mean([r.key[0] for r in closed]))
    eng.run(serving_stream(warm, seed=99))
    eng.completed.clear()
    eng.batch_sizes.clear()
    eng.run(closed)
    cap = latency_stats(closed)
    closed_batch = float(np.mean(eng.batch_sizes))
    progress(f"serving ({'ragged' if ragged else 'same-structure'}): closed-loop capacity "
             f"{cap['evals_per_s']:.0f} evals/s, mean batch {closed_batch:.1f}")
    rate = 0.7 * cap["evals_per_s"]
    eng.completed.clear()
    eng.batch_sizes.clear()
    stream = serving_stream(n_req, seed=5, rate=rate)
    eng.run(stream)
    st = latency_stats(stream)
    mean_batch = float(np.mean(eng.batch_sizes))
    progress(f"serving: Poisson stream p50 {st['p50_ms']:.2f} ms, p99 {st['p99_ms']:.2f} ms, "
             f"mean batch {mean_batch:.2f}")
    high = None
    if common_rate is not None:
        high = {"p50_ms": st["p50_ms"], "p99_ms": st["p99_ms"], "evals_per_s": st["evals_per_s"],
                "arrival_rate": rate, "mean_batch": mean_batch}
        rate = common_rate
        eng.completed.clear()
        eng.batch_sizes.clear()
        stream = serving_stream(n_req, seed=5, rate=rate)
        eng.run(stream)
        st = latency_stats(stream)
        mean_batch = float(np.mean(eng.batch_sizes))
        progress(f"serving: Poisson stream at {rate:.0f}/s p50 {st['p50_ms']:.2f} ms, p99 {st['p99_ms']:.2f} ms, "
                 f"mean batch {mean_batch:.2f}")
    if not ragged:
        return {"p50_ms": st["p50_ms"], "p99_ms": st["p99_ms"], "evals_per_s": st["evals_per_s"],
                "arrival_rate": rate, "capacity_evals_per_s": cap["evals_per_s"], "mean_batch": mean_batch,
                "closed_loop_mean_batch": closed_batch}
    single = ServingEngine(factory, hbm_budget=64 << 30, max_batch=1, n_streams=1, max_inflight=1,
                           graph_max_qubits=0)
    sub = serving_stream(min(n_req, 400), seed=6)
    single.run(sub)
    un = latency_stats(sub)
    progress(f"serving: unbatched {un['evals_per_s']:.0f} evals/s")
    out = {"p50_ms": st["p50_ms"], "p99_ms": st["p99_ms"], "mean_ms": st["mean_ms"],
           "evals_per_s": st["evals_per_s"], "arrival_rate": rate, "requests": st["requests"],
           "capacity_evals_per_s": cap["evals_per_s"], "mean_batch": mean_batch,
           "closed_loop_mean_batch": closed_batch, "ragged_batches": eng.ragged_batches,
           "unbatched_evals_per_s": un["evals_per_s"], "graphs_captured": len(eng._graphs),
           "batching": "ragged once a full batch (64) of a tile shape waits: requests of different "
                       "structures sharing a tile shape run in one launch per pass index; below that, and for "
                       "identical structures, the oldest request's structure runs at once on captured graphs",
           "stream": ("Poisson at 0.7x the same-structure engine's closed-loop capacity (high_load: at 0.7x "
                      "this engine's own)" if high else "Poisson at 0.7x closed-loop capacity")
                     + "; n~U{6..24}, layers~U{2..20}; <Z_0> + full gradient",
           "qubits_mean": wide_n, "qubits_simulated_mean": sim_n}
    if high:
        out["high_load"] = high
    return out


def measured_traffic(kernel, n, layers, batch, ex):
    """Mean HBM bytes per launch of ``kernel`` from the committed FETCH_SIZE / WRITE_SIZE counter runs of
    this workload (profiles/sv_pass_traffic_hea20_b256.json: one record per profiled adjoint shape), when
    the run uses a profiled shape; else None."""
    path = os.path.join(ROOT, "profiles", "sv_pass_traffic_hea20_b256.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        recs = json.load(f)
    ckpt = ex._checkpoints_fit(torch.device("cuda", 0), batch, True)  # pylint: disable=protected-access
    for rec in recs if isinstance(recs, list) else [recs]:
        w = rec["workload"]
        if (rec["kernel"] == kernel and w["qubits"] == n and w["layers"] == layers and w["batch"] == batch
                and w["tile_bits_bwd"] == ex.bwd.plan.k and w["reg_bits_bwd"] == ex.bwd.plan.r
                and w.get("forward_checkpoints", False) == ckpt):
            return rec["traffic_bytes_per_launch"]
    return None


def measured_valu_share(kernel, n, layers, batch, ex):
    """Time-weighted VALU issue share of ``kernel`` (SQ_INSTS_VALU x 4 cycles over the SIMDs' cycles) from the
    committed SQ counter runs of this workload (profiles/sv_pass_sq_counters_hea20_b256.json), when the run
    uses a profiled shape pair; else None.  A pass kernel whose HBM fraction is low because it is busy
    issuing arithmetic shows it here."""
    path = os.path.join(ROOT, "profiles", "sv_pass_sq_counters_hea20_b256.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        recs = json.load(f)
    # a kernel family's share depends on its own tile shape (and, for the adjoint, on checkpoints)
    adj = kernel == "sv_pass_adj"
    shape = (ex.bwd.plan.k, ex.bwd.plan.r) if adj else (ex.fwd.plan.k, ex.fwd.plan.r)
    ckpt = ex._checkpoints_fit(torch.device("cuda", 0), batch, True)  # pylint: disable=protected-access
    for rec in reversed(recs):
        w = rec["workload"]
        own = (w["tile_bits_bwd"], w["reg_bits_bwd"]) if adj else (w["tile_bits"], w["reg_bits"])
        if ((w["qubits"], w["layers"], w["batch"]) == (n, layers, batch) and own == shape
                and (not adj or w["forward_checkpoints"] == ckpt)):
            return rec["valu_issue_share"].get(kernel)
    return None


def tn_leg(c4=True, c3=True):
    """Contraction mode: the 40-qubit disordered chain (forward, batch 32) and the 5x5 lattice (fwd + grad)."""
    from tedq_amd import workloads as W  # pylint: disable=import-outside-toplevel
    dev = torch.device("cuda", 0)
    out = {}
    if c3:
        out.update(_chain_leg(W, dev))
    if c4:
        out.update(_lattice_leg(W, dev))
    return out


def tn_bound_s(ex, batch, with_grad):
    """Roofline time of the executed contraction steps: per step the larger of 8 flops per complex
    multiply-add at the 157.3 TFLOP/s FP32 matrix peak and its operand + result bytes at 8 TB/s
    (forward steps, plus both environment steps of the backward when ``with_grad``)."""
    total = 0.0
    for net in ex.networks:
        for st in net.steps:
            rows = [st.fwd] + ([r for r in (st.bwd_a, st.bwd_b) if r >= 0] if with_grad else [])
            for r in rows:
                nh, nm, nn, nk = (int(v) for v in ex._descs[r][:4])  # pylint: disable=protected-access
                flops = 8.0 * batch * 2.0 ** (nh + nm + nn + nk)
                byts = 8.0 * batch * (2.0 ** (nh + nm + nk) + 2.0 ** (nh + nn + nk) + 2.0 ** (nh + nm + nn))
                total += max(flops / 157.3e12, byts / 8e12)
    return total


def _chain_leg(W, dev):
    out = {}
    t0 = time.perf_counter()
    ex = W.disordered_chain(40, 8, 20, tn_mode=True).executor()
    build_s = time.perf_counter() - t0
    rng = np.random.default_rng(2)
    rows = torch.tensor(W.chain_rows(rng.uniform(-5, 5, size=(32, 40)), 8), device=dev)
    with torch.no_grad():
        ex.measure(ex.run_forward(rows))
        torch.cuda.synchronize()
        reps = 10
        t0 = time.perf_counter()
        for _ in range(reps):
            v = ex.measure(ex.run_forward(rows))[0]
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / reps
    info = ex.path_report()[0]
    progress(f"contraction: 40-qubit chain {dt * 1e3:.1f} ms per batch of 32")
    out["chain40"] = {"workload": "1-D disordered chain 40q, 8 Trotter steps, <Z_20>, batch 32 seeds, forward",
                      "evals_per_s": 32 / dt, "ms_per_batch": dt * 1e3, "path_build_s": build_s,
                      "steps": info["steps"], "largest_rank": info["largest_rank"], "macs_per_eval": info["macs"],
                      "achieved_TFLOPs": 8.0 * info["macs"] * 32 / dt / 1e12,
                      "roofline": {"bound_ms": tn_bound_s(ex, 32, False) * 1e3,
                                   "frac": tn_bound_s(ex, 32, False) / dt,
                                   "model": "per step max(8 flops/MAC at 157.3 TFLOP/s, bytes at 8 TB/s)"},
                      "mean_value": float(v.mean())}
    return out


def _lattice_leg(W, dev):
    out = {}
    torch.cuda.empty_cache()  # the backward keeps up to ~94 GB of intermediates per network
    t0 = time.perf_counter()
    ex = W.lattice(5, 6, tn_mode=True).executor()
    build_s = time.perf_counter() - t0
    progress(f"contraction: 5x5 lattice paths found in {build_s:.1f} s")
    rng = np.random.default_rng(3)
    rows = torch.tensor(rng.uniform(0, 2 * np.pi, size=(1, 300)), dtype=torch.float32, device=dev)
    ups = [torch.full((1,), 1.0 / 25, device=dev) for _ in range(25)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    outs, grad = ex.value_and_grad(rows, ups)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    rep = ex.path_report()
    progress(f"contraction: 5x5 lattice value and gradient in {dt:.1f} s")
    out["lattice5x5"] = {"workload": "5x5 lattice 6 layers, mean <Z_i> over 25 networks, fwd + reverse-tree grad",
                         "s_per_eval": dt, "path_build_s": build_s,
                         "steps": sum(r["steps"] for r in rep),
                         "largest_rank": max(r["largest_rank"] for r in rep),
                         "macs_fwd": sum(r["macs"] for r in rep),
                         "roofline": {"bound_s": tn_bound_s(ex, 1, True), "frac": tn_bound_s(ex, 1, True) / dt,
                                      "model": "per forward and environment step max(8 flops/MAC at 157.3 "
                                               "TFLOP/s, bytes at 8 TB/s)"},
                         "cost": float(sum(float(o[0]) for o in outs) / 25),
                         "grad_finite": bool(torch.isfinite(grad).all())}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--qubits", type=int, default=20)
    ap.add_argument("--layers", type=int, default=10)
    ap.add_argument("--batch", type=int, default=256)
    ap.add_argument("--tile-bits", type=int, default=12)
    ap.add_argument("--reg-bits", type=int, default=4)
    ap.add_argument("--tile-bits-bwd", type=int, default=None)
    ap.add_argument("--reg-bits-bwd", type=int, default=None)
    ap.add_argument("--skip-cpu", action="store_true")
    ap.add_argument("--skip-plain", action="store_true")
    ap.add_argument("--skip-serving", action="store_true")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--skip-tn", action="store_true")
    ap.add_argument("--no-jit", action="store_true", help="keep the interpreted pass kernels")
    ap.add_argument("--skip-c4", action="store_true", help="leave out the 5x5 lattice (contraction mode)")
    ap.add_argument("--skip-c3", action="store_true", help="leave out the 40-qubit chain (contraction mode)")
    ap.add_argument("--tn-only", action="store_true", help="run only the contraction-mode leg and print its JSON")
    ap.add_argument("--plain-only", action="store_true",
                    help="time one call of the plain per-gate torch path (--qubits, --layers) and print its JSON")
    ap.add_argument("--serve-requests", type=int, default=1500)
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a 1 GiB device-to-device copy (counter calibration for traffic)")
    args = ap.parse_args()
    if args.copy_calibration:
        src = torch.ones(2 ** 28, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": src.numel() * 4, "copy_bytes_written": dst.numel() * 4}))
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this engine is measured on one GPU; multi-GPU runs are not part of this "
                                   "build (batches shard across GPUs without communication)", "n_gpus": args.gpus}))
        sys.exit(2)

    import tedq_amd  # noqa: F401  pylint: disable=import-outside-toplevel,unused-import
    if args.tn_only:
        print(json.dumps({"tensor_network": tn_leg(c4=not args.skip_c4, c3=not args.skip_c3)}), flush=True)
        return
    if args.plain_only:
        theta = np.random.default_rng(1).uniform(0, 2 * np.pi, size=(2, args.layers * args.qubits * 2))
        ps, cnt, fwd_s, bwd_s = time_plain_gpu(args.qubits, args.layers, theta, torch.device("cuda", 0),
                                               budget_s=0.0)
        print(json.dumps({"plain_gpu": {"value": ps, "unit": "evals/s", "calls": cnt, "forward_s": fwd_s,
                                        "backward_s": bwd_s, "qubits": args.qubits, "layers": args.layers}}),
              flush=True)
        return
    dev = torch.device("cuda", 0)
    n, L, B = args.qubits, args.layers, args.batch
    rng = np.random.default_rng(1)
    theta = rng.uniform(0, 2 * np.pi, size=(B, L * n * 2))
    compiled = hea_circuit(n, L).compilecircuit(backend="pytorch", tile_bits=args.tile_bits,
                                                 reg_bits=args.reg_bits, tile_bits_bwd=args.tile_bits_bwd,
                                                 reg_bits_bwd=args.reg_bits_bwd)
    ex = compiled.executor()
    rows = torch.tensor(theta, dtype=torch.float32, device=dev)

    tuning = None
    if not args.no_autotune:
        from tedq_amd.engine.autotune import autotune  # pylint: disable=import-outside-toplevel
        res = autotune(ex, rows, jit=not args.no_jit, log=progress)
        tuning = {"best": list(res["best"]), "fwd_ms": {f"{k}/{r}": v for (k, r), v in res["fwd"].items()},
                  "bwd_ms": {f"{k}/{r}": v for (k, r), v in res["bwd"].items()},
                  "checkpointed_fwd_plus_bwd_ms": {f"{a}/{b}+{c}/{d}": v
                                                   for (a, b, c, d), v in res["checkpointed"].items()}}

    jit_s = None
    if not args.no_jit:
        t0 = time.perf_counter()
        jit_used = ex.enable_jit(dev)
        jit_s = time.perf_counter() - t0
        ex.jit = True

    def step():
        r = rows.detach().requires_grad_(True)
        out = compiled.execute_batch(r)
        out.sum().backward()
        return r.grad

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    progress(f"warmup done ({args.warmup} steps); timing {args.steps} steps")
    probe = EventProbe()
    ex.probe = probe
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        g = step()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    ex.probe = None
    assert torch.isfinite(g).all()
    value = B * args.steps / elapsed
    progress(f"timed: {value:.1f} evals/s, {elapsed / args.steps * 1e3:.2f} ms per step")
    from tedq_amd.engine.sv import norm_deviation  # pylint: disable=import-outside-toplevel
    norm_dev = norm_deviation(ex)

    amps = B * (2 ** n)
    kernels = {}
    # algorithmic bytes count only the tiles a pass processes: from |0..0> the first passes skip the tiles
    # that hold zeros (zero masks), so their launches move a fraction of the state
    live = {"sv_pass_fwd": [2.0 ** -bin(z).count("1") for z in ex.fwd.zero_masks],
            "sv_pass_adj": [2.0 ** -bin(z).count("1") for z in ex.bwd.zero_masks]}
    # with per-pass checkpoints the adjoint reads psi and lambda and writes lambda only (24 B per amplitude)
    ckpt = ex._checkpoints_fit(dev, B, True)  # pylint: disable=protected-access
    # generated forward passes from |0..0> read only the amplitudes whose tile bits earlier passes held
    # (the first pass reads nothing): 8 B written plus 8 B x that fraction read per amplitude
    fresh_loads = ex.fwd._jit_fns(dev, False) is not None  # pylint: disable=protected-access
    per_pass = {"sv_pass_fwd": [], "sv_pass_adj": [(24 if ckpt else 32) * amps * f for f in live["sv_pass_adj"]]}
    seen = 0
    for i, (f, m) in enumerate(zip(live["sv_pass_fwd"], ex.fwd.enc.tile_masks)):
        rd = (0.0 if i == 0 else 2.0 ** -bin(m & ~seen).count("1")) if fresh_loads else 1.0
        per_pass["sv_pass_fwd"].append(8 * amps * f * (1 + rd))
        seen |= m
    for kind in ("sv_pass_fwd", "sv_pass_adj"):
        ms, cnt = probe.mean_ms(kind)
        if ms:
            frac = float(np.mean(live[kind]))
            nbytes = float(np.mean(per_pass[kind]))
            kernels[kind] = {"mean_ms": ms, "launches": cnt, "total_ms": ms * cnt, "bytes_per_launch": nbytes,
                             "live_tile_fraction": frac, "achieved_GBs": nbytes / (ms * 1e-3) / 1e9}
    dom = max(kernels, key=lambda k: kernels[k]["total_ms"])
    kd = kernels[dom]
    line = {
        "metric": "circuit fwd+grad evals/s (20q d10, batch 256) + p50/p99 latency, mixed stream",
        "value": value, "unit": "evals/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "c64", "data": "synthetic",
        "config": {"workload": f"hardware-efficient ansatz {n}q {L} layers, batch {B}, <Z_i> for all qubits, "
                               "fwd + adjoint grad", "qubits": n, "layers": L, "batch": B,
                   "params_per_circuit": 2 * n * L, "tile_bits": ex.fwd.plan.k, "reg_bits": ex.fwd.plan.r,
                   "tile_bits_bwd": ex.bwd.plan.k, "reg_bits_bwd": ex.bwd.plan.r,
                   "passes": ex.n_passes, "passes_bwd": len(ex.bwd.plan.passes),
                   # exact simplifications: the last RZ layer and CNOT chain folded into the Z-string
                   # observables; diagonal gates applied without their global phase (expectations only)
                   "folded_gates": ex.folded, "phase_free_forward": ex.fwd.plan.specialize,
                   "forward_checkpoints": ckpt,
                   "pass_kernels": "interpreted" if args.no_jit else
                   {"generated_fwd": jit_used[0], "generated_bwd": jit_used[1], "compile_s": jit_s}},
        "autotune": tuning,
        "roofline": {"kernel": dom, "bound": "hbm", "achieved": kd["achieved_GBs"], "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": kd["achieved_GBs"] / HBM_PEAK_GBS,
                     "traffic": measured_traffic(dom, n, L, B, ex),
                     "valu_issue_share": measured_valu_share(dom, n, L, B, ex),
                     "bytes_per_launch": kd["bytes_per_launch"], "mean_launch_ms": kd["mean_ms"]},
        "kernels": kernels,
        "norm_check": {"max_abs_deviation": norm_dev, "rows": B,
                       "how": "<psi|psi> of every row of the last timed step, fp64 compensated reduction on "
                              "the device in the measurement pass"},
        # speed figure: the sweeps an unfused gate-by-gate engine makes per evaluation (3 per gate in the
        # forward + adjoint, one per parameter, one for the measurement), 16 B per amplitude each;
        # fusion lets this exceed HBM bandwidth, so it is never a roofline fraction
        "unfused_equivalent": {"sweeps_per_eval": 3 * L * (3 * n - 1) + 2 * n * L + 1,
                               "TB_per_s": value * (3 * L * (3 * n - 1) + 2 * n * L + 1) * 16 * 2 ** n / 1e12},
    }
    if not args.skip_serving:
        same = serving_leg(args.serve_requests, warm=200, ragged=False)
        line["latency"] = serving_leg(args.serve_requests, warm=200, ragged=True, common_rate=same["arrival_rate"])
        line["latency"]["same_structure_only"] = same
    if not args.skip_tn:
        line["tensor_network"] = tn_leg(c4=not args.skip_c4)
    if not args.skip_plain:
        b1 = engine_batch1(compiled, rows)
        progress(f"engine, one row per call: {b1['evals_per_s']:.1f} evals/s")
        ps, cnt, fwd_s, bwd_s = time_plain_gpu(n, L, theta[:16], dev)
        progress(f"plain single-circuit path: {ps:.4f} evals/s")
        line["plain_gpu"] = {"value": ps, "unit": "evals/s", "sample": f"{cnt} single-row calls",
                             "forward_s": fwd_s, "backward_s": bwd_s,
                             "speedup_batched_over_plain": value / ps,
                             "engine_batch1": b1, "speedup_batched_over_engine_batch1": value / b1["evals_per_s"]}
    if not args.skip_cpu:
        # the engine's values and gradients of two rows of the benchmarked batch, for the error check
        idx = [0, min(B, 32) - 1]
        r = rows[idx].detach().clone().requires_grad_(True)
        vals = compiled.execute_batch(r)
        vals.sum().backward()
        check = (idx, vals.detach().double().cpu().numpy(), r.grad.double().cpu().numpy())
        line["cpu_baseline"] = cpu_baseline(n, L, theta[:32], check=check)
        progress(f"cpu baseline: {line['cpu_baseline']['value']:.3f} evals/s")
        try:
            line["cpu_baseline"]["small_circuit"] = sc = small_circuit_leg(dev)
            progress(f"small circuit: engine {sc['engine_ms_per_eval']:.2f} ms, "
                     f"cpu {sc['cpu_port_ms_per_eval']:.2f} ms per eval")
        except (RuntimeError, ValueError) as e:  # a side record: never lose the main line for it
            line["cpu_baseline"]["small_circuit"] = {"error": str(e)[:300]}
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
