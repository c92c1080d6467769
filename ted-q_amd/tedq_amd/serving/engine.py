"""Continuous-batching request loop for compiled-circuit evaluations on one GPU.

Each request is a circuit structure key plus one parameter row; it asks for
the circuit's measurements and the full parameter gradient.  The loop

* keeps one compiled device program per structure key (built on first use,
  then reused: state, tables and programs stay on the device);
* queues requests per key as they arrive and forms batches of structurally
  identical circuits from whatever is waiting when a slot frees up, oldest
  request first;
* with ``ragged=True`` queues by tile shape instead: once ``ragged_min``
  requests of a tile shape wait (a backlog: by default a full batch), a batch
  may mix circuit structures (depths and qubit counts), and every pass index
  of all of them runs as one ragged launch (tedq_amd/engine/ragged.py); with
  fewer waiting, or when the oldest requests share one structure, the oldest
  request's structure is served at once on the same-structure path (captured
  graphs), which costs the host far less per batch, so latency under moderate
  load stays that of the same-structure loop;
* admits a batch only while the bytes of all in-flight batches stay within an
  HBM budget (requests wait otherwise);
* runs batches on several HIP streams so their kernels overlap on the device,
  polls completion with events, and returns results to pinned host memory.

The reference has no serving component; its unit of work is one call of the
compiled circuit (tedq/backends/pytorch_backend.py:124-165), which is what a
single-row batch of this loop executes.
"""

import time
from collections import deque
from dataclasses import dataclass, field

import numpy as np
import torch


@dataclass
class Request:
    """One evaluation: structure key, parameter row (float32 numpy), arrival time in seconds."""
    rid: int
    key: tuple
    params: np.ndarray
    arrival: float
    value: np.ndarray = None
    grad: np.ndarray = None
    t_launch: float = None
    t_done: float = None

    @property
    def latency(self):
        return None if self.t_done is None else self.t_done - self.arrival


class GraphedFwdGrad:
    """One captured launch sequence (tables, passes, measurement, adjoint) for a fixed batch size.

    Inputs and outputs live in static device buffers; ``run`` copies a
    parameter block in, replays the graph on a stream and copies results to
    pinned host memory.  Only one run per instance may be in flight.
    """

    def __init__(self, ex, batch, device, stream):
        self.batch = batch
        self.n_out = len(ex.specs)
        self.rows = torch.full((batch, ex.n_params), 0.5, dtype=torch.float32, device=device)
        self.ups = [torch.ones(batch, device=device) for _ in ex.specs]
        self.host_in = torch.empty((batch, ex.n_params), dtype=torch.float32, pin_memory=True)
        self.host_vals = torch.empty((batch, self.n_out), dtype=torch.float32, pin_memory=True)
        self.host_grad = torch.empty((batch, ex.n_params), dtype=torch.float32, pin_memory=True)
        self.stream = stream
        with torch.cuda.stream(stream):
            for _ in range(2):                      # warm the allocator outside the capture
                ex.value_and_grad(self.rows, self.ups)
        stream.synchronize()
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph, stream=stream):
            outs, grad = ex.value_and_grad(self.rows, self.ups)
            self.vals = torch.stack([o.reshape(batch) for o in outs], 1).to(torch.float32)
            self.grad = grad
        self.nbytes = ex.memory_bytes(batch)

    def run(self, params_np):
        b = params_np.shape[0]
        self.host_in[:b].copy_(torch.from_numpy(params_np))
        with torch.cuda.stream(self.stream):
            self.rows.copy_(self.host_in, non_blocking=True)
            self.graph.replay()
            self.host_vals.copy_(self.vals, non_blocking=True)
            self.host_grad.copy_(self.grad, non_blocking=True)
            ev = torch.cuda.Event()
            ev.record(self.stream)
        return ev


@dataclass
class _Inflight:
    reqs: list
    event: object
    nbytes: int
    host_vals: torch.Tensor
    host_grad: torch.Tensor
    keep: list = field(default_factory=list)
    # ragged batches: per request (host values, host gradients, row) instead of the two tensors above
    slots: list = None


class ServingEngine:
    """Continuous batching of same-structure circuits under an HBM budget.

    ``factory(key)`` must return a compiled circuit (``PyTorchBackend``) for a
    structure key; its executor provides ``value_and_grad`` and
    ``memory_bytes``.  ``upstream(key, B)`` gives the upstream gradient of
    every measurement for a batch (default: ones, i.e. the gradient of the sum
    of all expectation values).
    """

    def __init__(self, factory, device=None, hbm_budget=32 << 30, max_batch=64, n_streams=4, max_inflight=8,
                 graph_max_qubits=22, ragged=False, ragged_min=None, ragged_max=64):
        self.factory = factory
        # queue by tile shape and mix structures in one launch sequence (see module docstring) once
        # ragged_min requests of a shape wait (None: a full batch, max_batch); a ragged batch holds at most
        # ragged_max requests (its host-side cost grows with the structures in it)
        self.ragged = ragged
        self.ragged_min = max_batch if ragged_min is None else ragged_min
        self.ragged_max = ragged_max
        self.ragged_batches = 0
        self.graph_max_qubits = graph_max_qubits
        self._graphs = {}
        self._graph_busy = set()
        self.graph_bytes = 0          # device memory held by captured graphs (counted in the budget)
        self.device = device or torch.device("cuda", torch.cuda.current_device())
        self.hbm_budget = hbm_budget
        self.max_batch = max_batch
        self.max_inflight = max_inflight
        self.streams = [torch.cuda.Stream(device=self.device) for _ in range(n_streams)]
        self._next_stream = 0
        self._compiled = {}
        self._queues = {}
        self._inflight = deque()
        self._bytes_in_use = 0
        self.completed = []
        self.batch_sizes = []

    # ------------------------------------------------------------------
    def executor(self, key):
        if key not in self._compiled:
            self._compiled[key] = self.factory(key).executor()
        return self._compiled[key]

    def _queue_key(self, key):
        if not self.ragged:
            return key
        from ..engine.ragged import shape_key  # pylint: disable=import-outside-toplevel
        return ("shape",) + shape_key(self.executor(key))

    def submit(self, req):
        self._queues.setdefault(self._queue_key(req.key), deque()).append(req)

    @property
    def pending(self):
        return sum(len(q) for q in self._queues.values())

    @property
    def inflight(self):
        return len(self._inflight)

    # ------------------------------------------------------------------
    def _pick_key(self):
        best, best_t = None, None
        for key, q in self._queues.items():
            if q and (best_t is None or q[0].arrival < best_t):
                best, best_t = key, q[0].arrival
        return best

    def _graph_for(self, key, ex, b):
        """A free captured runner for ``b`` rows (power-of-two bucket), or None."""
        if ex.n > self.graph_max_qubits:
            return None
        bucket = 1
        while bucket < b:
            bucket *= 2
        for i in range(4):
            gk = (key, bucket, i)
            if gk in self._graph_busy:
                continue
            if gk not in self._graphs:
                need = ex.memory_bytes(bucket)
                if self.graph_bytes + need > self.hbm_budget // 2:
                    return None
                stream = self.streams[len(self._graphs) % len(self.streams)]
                self._graphs[gk] = GraphedFwdGrad(ex, bucket, self.device, stream)
                self.graph_bytes += need
            return gk
        return None

    def prepare(self, keys, buckets=(1,), log=None):
        """Compile and capture ahead of traffic (keys above the graph threshold only compile).

        ``log``, if given, is called with a line of text after every 25 keys."""
        keys = list(keys)
        for i, key in enumerate(keys):
            ex = self.executor(key)
            for b in buckets:
                if b <= self.max_batch:
                    self._graph_for(key, ex, b)
            if log and ((i + 1) % 25 == 0 or i + 1 == len(keys)):
                log(f"prepared {i + 1}/{len(keys)} circuit structures")
        torch.cuda.synchronize(self.device)

    def _launch_graph(self, key, ex, now, q=None):
        q = self._queues[key] if q is None else q
        b = min(len(q), self.max_batch)
        gk = self._graph_for(key, ex, b)
        if gk is None:
            return False
        runner = self._graphs[gk]
        reqs = [q.popleft() for _ in range(b)]
        ev = runner.run(np.stack([r.params for r in reqs]).astype(np.float32))
        for r in reqs:
            r.t_launch = now
        self._graph_busy.add(gk)
        self._inflight.append(_Inflight(reqs, ev, 0, runner.host_vals, runner.host_grad, keep=[gk]))
        self.batch_sizes.append(b)
        return True

    def _launch(self, qkey, now):
        q = self._queues[qkey]
        if self.ragged:
            head = [q[i].key for i in range(min(len(q), self.max_batch))]
            if len(set(head)) > 1 and len(q) >= self.ragged_min:
                return self._launch_ragged(qkey, now)
            key = head[0]
            if len(set(head)) > 1:
                # light load: the oldest request's structure at once (with its twins)
                mine = deque(r for r in q if r.key == key)
                others = [r for r in q if r.key != key]
                launched = self._launch_same(key, mine, now)
                if launched:
                    self._queues[qkey] = deque(sorted(list(mine) + others, key=lambda r: r.arrival))
                return launched
        else:
            key = qkey
        return self._launch_same(key, q, now)

    def _launch_same(self, key, q, now):
        """A batch of the structure ``key`` from the head of ``q`` (requests of that structure only)."""
        ex = self.executor(key)
        if ex.n <= self.graph_max_qubits and self._launch_graph(key, ex, now, q):
            return True
        B = min(len(q), self.max_batch)
        if self.ragged:
            B = min(B, sum(1 for r in q if r.key == key)) if q[0].key == key else 1
        while B > 1 and self._bytes_in_use + self.graph_bytes + ex.memory_bytes(B) > self.hbm_budget:
            B //= 2
        need = ex.memory_bytes(B)
        if self._bytes_in_use + self.graph_bytes + need > self.hbm_budget and self._inflight:
            return False
        reqs = [q.popleft() for _ in range(B)]
        stream = self.streams[self._next_stream % len(self.streams)]
        self._next_stream += 1
        host = torch.from_numpy(np.stack([r.params for r in reqs]).astype(np.float32)).pin_memory()
        with torch.cuda.stream(stream):
            rows = host.to(self.device, non_blocking=True)
            ups = [torch.ones(B, device=self.device) for _ in ex.specs]
            outs, grad = ex.value_and_grad(rows, ups)
            vals = torch.stack([o.reshape(B) for o in outs], 1).to(torch.float32)
            host_vals = torch.empty(vals.shape, dtype=torch.float32, pin_memory=True)
            host_grad = torch.empty(grad.shape, dtype=grad.dtype, pin_memory=True)
            host_vals.copy_(vals, non_blocking=True)
            host_grad.copy_(grad, non_blocking=True)
            ev = torch.cuda.Event()
            ev.record(stream)
        for r in reqs:
            r.t_launch = now
        self._inflight.append(_Inflight(reqs, ev, need, host_vals, host_grad, keep=[host, rows, vals, grad]))
        self._bytes_in_use += need
        self.batch_sizes.append(B)
        return True

    def _launch_ragged(self, qkey, now):
        """One launch sequence for the oldest waiting requests of a tile shape, whatever their structures."""
        from ..engine.ragged import ragged_value_and_grad  # pylint: disable=import-outside-toplevel
        q = self._queues[qkey]
        reqs, need = [], 0
        while q and len(reqs) < min(self.max_batch, self.ragged_max):
            more = self.executor(q[0].key).memory_bytes(1)
            if reqs and self._bytes_in_use + self.graph_bytes + need + more > self.hbm_budget:
                break
            reqs.append(q.popleft())
            need += more
        if self._bytes_in_use + self.graph_bytes + need > self.hbm_budget and self._inflight:
            for r in reversed(reqs):
                q.appendleft(r)
            return False
        groups = {}
        for r in reqs:
            groups.setdefault(r.key, []).append(r)
        stream = self.streams[self._next_stream % len(self.streams)]
        self._next_stream += 1
        keep, slots, order = [], {}, []
        with torch.cuda.stream(stream):
            items = []
            for key, grs in groups.items():
                ex = self.executor(key)
                host = torch.from_numpy(np.stack([r.params for r in grs]).astype(np.float32)).pin_memory()
                rows = host.to(self.device, non_blocking=True)
                keep += [host, rows]
                items.append((ex, rows, [torch.ones(len(grs), device=self.device) for _ in ex.specs]))
            results, bufs = ragged_value_and_grad(items, self.device)
            keep += bufs
            for (key, grs), (outs, grad) in zip(groups.items(), results):
                vals = torch.stack([o.reshape(len(grs)) for o in outs], 1).to(torch.float32)
                host_vals = torch.empty(vals.shape, dtype=torch.float32, pin_memory=True)
                host_grad = torch.empty(grad.shape, dtype=grad.dtype, pin_memory=True)
                host_vals.copy_(vals, non_blocking=True)
                host_grad.copy_(grad, non_blocking=True)
                keep += [vals, grad]
                for i, r in enumerate(grs):
                    slots[r.rid] = (host_vals, host_grad, i)
            ev = torch.cuda.Event()
            ev.record(stream)
        for r in reqs:
            r.t_launch = now
            order.append(slots[r.rid])
        self._inflight.append(_Inflight(reqs, ev, need, None, None, keep=keep, slots=order))
        self._bytes_in_use += need
        self.batch_sizes.append(len(reqs))
        self.ragged_batches += 1
        return True

    def poll(self):
        """Retire finished batches; returns the number of requests completed."""
        done = 0
        still = deque()
        while self._inflight:
            b = self._inflight.popleft()
            if b.event.query():
                t = time.perf_counter()
                for i, r in enumerate(b.reqs):
                    hv, hg, row = (b.host_vals, b.host_grad, i) if b.slots is None else b.slots[i]
                    r.value = hv[row].numpy().copy()
                    r.grad = hg[row].numpy().copy()
                    r.t_done = t
                    self.completed.append(r)
                self._bytes_in_use -= b.nbytes
                if b.keep and isinstance(b.keep[0], tuple):
                    self._graph_busy.discard(b.keep[0])
                done += len(b.reqs)
            else:
                still.append(b)
        self._inflight = still
        return done

    def schedule(self, now):
        """Launch batches while slots and memory allow."""
        launched = 0
        while len(self._inflight) < self.max_inflight:
            key = self._pick_key()
            if key is None or not self._launch(key, now):
                break
            launched += 1
        return launched

    def run(self, requests, clock_start=None, idle_sleep=50e-6):
        """Replay ``requests`` (sorted by arrival, seconds from start) in real time until all complete."""
        t0 = time.perf_counter() if clock_start is None else clock_start
        for r in requests:
            r.arrival += t0
        i = 0
        total = len(requests)
        while len(self.completed) < total:
            now = time.perf_counter()
            while i < total and requests[i].arrival <= now:
                self.submit(requests[i])
                i += 1
            self.poll()
            self.schedule(now)
            if self._inflight or i < total:
                time.sleep(idle_sleep)
        for r in requests:
            r.arrival -= t0
            r.t_done -= t0
            r.t_launch -= t0
        return requests


def latency_stats(requests):
    lat = np.array([r.latency for r in requests])
    span = max(r.t_done for r in requests) - min(r.arrival for r in requests)
    return {"p50_ms": float(np.percentile(lat, 50) * 1e3), "p99_ms": float(np.percentile(lat, 99) * 1e3),
            "mean_ms": float(lat.mean() * 1e3), "evals_per_s": len(requests) / span, "requests": len(requests)}
