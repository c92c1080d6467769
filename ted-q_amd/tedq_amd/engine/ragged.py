"""Ragged batches: circuits of different structures evaluated by one launch sequence.

A serving batch may hold requests of many circuit structures (different
depths and qubit counts).  Launching each request's passes separately costs
one kernel launch per pass per request and leaves the GPU idle for small
circuits.  Here the passes of all circuits that share the tile shape (tile
bits k, register bits r; every circuit of at least 12 qubits with the default
shapes) run as ONE launch per pass index: ``tedq_sv_pass_ragged`` takes a
device table of per-circuit records (state, program of this pass, slot
matrices, partial rows, rest-bit mask, tile count), so circuit c's tiles are
workgroups [tile_end(c-1), tile_end(c)) of that launch.  Circuits with fewer
passes simply drop out of the later launches.  The adjoint sweep is batched
the same way; tables, measurements, adjoint seeds and gradient reductions run
per structure (batched over that structure's rows).

Every circuit runs the same programs, kernels and reductions as when it is
evaluated alone, so its results do not depend on what else is in the batch.
"""

import ctypes

import numpy as np
import torch

from . import hip

# one record per circuit and pass (RaggedRec in csrc/sv_passes.hip: 64 bytes)
RECORD = np.dtype([("psi", "<u8"), ("lam", "<u8"), ("prog", "<u8"), ("mats", "<u8"), ("partials", "<u8"),
                   ("rest_mask", "<u8"), ("tile_end", "<i8"), ("n", "<i4"), ("n_gslots", "<i4")])
assert RECORD.itemsize == 64


def shape_key(ex):
    """Circuits whose executors share this key can run in one ragged launch sequence."""
    return (ex.fwd.plan.k, ex.fwd.plan.r, ex.bwd.plan.k, ex.bwd.plan.r)


def _records(items, adjoint):
    """Record table of every pass index: (records array, [(start, count, tiles) per pass index])."""
    per_pass = []
    for it in items:
        prog = it["ex"].bwd if adjoint else it["ex"].fwd
        enc, plan = prog.enc, prog.plan
        n = it["ex"].n
        full = (1 << n) - 1
        offs = enc.adj_offsets if adjoint else enc.fwd_offsets
        masks = list(reversed(enc.tile_masks)) if adjoint else enc.tile_masks
        base = prog.buffers(it["dev"]).prog_base
        B = it["rows"].shape[0]
        amp = 8 << n
        mstride = plan.mat_stride * 8
        mats = it["mats_b" if adjoint else "mats_f"].data_ptr()
        G = plan.n_gslots
        prow_bytes = max(G, 1) * prog.partial_rows() * 32
        for i, (off, mask) in enumerate(zip(offs, masks)):
            while len(per_pass) <= i:
                per_pass.append([])
            for b in range(B):
                per_pass[i].append((it["psi"].data_ptr() + b * amp,
                                    it["lam"].data_ptr() + b * amp if adjoint else 0,
                                    base + 4 * off, mats + b * mstride,
                                    it["partials"].data_ptr() + b * prow_bytes if adjoint else 0,
                                    full & ~mask, 1 << (n - plan.k), n, G if adjoint else 0))
    rows, bounds = [], []
    for recs in per_pass:
        start = len(rows)
        end = 0
        for psi, lam, prg, mats, part, rest, tiles, n, g in recs:
            end += tiles
            rows.append((psi, lam, prg, mats, part, rest, end, n, g))
        bounds.append((start, len(recs), end))
    arr = np.array(rows, dtype=RECORD)
    return arr, bounds


def _launch(arr, bounds, adjoint, k, r, dev, st):
    """Copy the record table to the device and issue one ragged launch per pass index."""
    host = torch.from_numpy(arr.view(np.uint8).copy()).pin_memory()
    table = torch.empty(host.shape, dtype=torch.uint8, device=dev)
    table.copy_(host, non_blocking=True)
    for start, count, tiles in bounds:
        if not count:
            continue
        hip.call("tedq_sv_pass_ragged", int(adjoint), ctypes.c_void_p(table.data_ptr() + 64 * start), count, tiles,
                 k, r, st)
    return host, table


def ragged_value_and_grad(items, device):
    """Measurements and parameter gradients of several circuit structures in one launch sequence.

    ``items``: list of (executor, rows [B, P] on ``device``, upstream gradients per measurement); all
    executors must share ``shape_key`` and use the library's pass kernels (not generated ones).
    Returns [(outputs, grad [B, P])] in item order and the buffers the launches read (keep them alive
    until the work has finished).
    """
    if not items:
        return [], []
    keys = {shape_key(ex) for ex, _, _ in items}
    if len(keys) != 1:
        raise ValueError(f"ragged batch mixes tile shapes {sorted(keys)}")
    kf, rf, kb, rb = keys.pop()
    st = hip.stream_handle(device)
    its = []
    for ex, rows, ups in items:
        if rows.shape[1] != ex.n_params or ex.n < ex.fwd.plan.k:
            raise ValueError("ragged batch: parameter rows do not match their circuit")
        mats_f, _ = ex.fwd.tables(rows, False)
        psi = torch.empty((rows.shape[0], 2 ** ex.n), dtype=torch.complex64, device=device)
        _, init = ex._zmask_buffer(device)  # pylint: disable=protected-access
        if init is None:
            hip.call("tedq_sv_init_zero", hip.ptr(psi), rows.shape[0], ex.n, st)
        else:
            hip.call("tedq_sv_init_copy", hip.ptr(psi), hip.ptr(init), rows.shape[0], ex.n, st)
        its.append({"ex": ex, "rows": rows, "ups": ups, "dev": device, "mats_f": mats_f, "psi": psi})
    keep = []
    arr, bounds = _records(its, adjoint=False)
    keep += _launch(arr, bounds, False, kf, rf, device, st)
    for it in its:
        ex = it["ex"]
        it["outs"] = ex.measure(it["psi"])
        it["lam"] = ex.seed_adjoint(it["psi"], it["ups"])
        it["mats_b"], it["amats"] = ex.bwd.tables(it["rows"], True)
        G = ex.bwd.plan.n_gslots
        it["partials"] = torch.empty((it["rows"].shape[0], max(G, 1), ex.bwd.partial_rows(), 8),
                                     dtype=torch.float32, device=device)
    arr, bounds = _records(its, adjoint=True)
    keep += _launch(arr, bounds, True, kb, rb, device, st)
    results = []
    for it in its:
        ex, rows = it["ex"], it["rows"]
        B = rows.shape[0]
        prog = ex.bwd
        G = prog.plan.n_gslots
        grad = torch.empty((B, max(ex.n_params, 1)), dtype=rows.dtype, device=device)
        if ex.n_params:
            C = torch.empty((B, max(G, 1), 8), dtype=torch.float64, device=device)
            hip.call("tedq_sv_grad_reduce", hip.ptr(it["partials"]), B, G, prog.partial_rows(), hip.ptr(it["amats"]),
                     hip.ptr(prog.buffers(device).pgslot), ex.n_params, hip.ptr(C), hip.ptr(grad),
                     int(rows.dtype == torch.float64), st)
            keep.append(C)
        keep += [it["psi"], it["lam"], it["mats_f"], it["mats_b"], it["amats"], it["partials"]]
        results.append((it["outs"], grad[:, :ex.n_params]))
    return results, keep
