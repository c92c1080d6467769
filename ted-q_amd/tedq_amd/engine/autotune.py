"""Tile-shape autotuner for the state-vector passes.

For a compiled circuit and a representative parameter batch, times the
forward program and the adjoint program for a set of (tile bits k, register
bits r) shapes with device events and configures the executor with the
fastest of each.  The shapes trade pass count (larger k: fewer passes) against
LDS per workgroup (2^k x 8 B forward, 2^k x 16 B adjoint) and register
pressure (larger r: fewer LDS gather/scatter stages, more VGPRs).  A forward
and an adjoint with the same k share one pass partition, so the forward can
keep its per-pass states for the adjoint (checkpoints, which spare the adjoint
a quarter of its traffic); such pairs are timed too and win when their sum is
lower.
"""

import torch

FWD_SHAPES = ((12, 4), (13, 4), (13, 5), (12, 5), (11, 4), (12, 3))
BWD_SHAPES = ((12, 4), (12, 3), (13, 4), (11, 4), (11, 3), (10, 3), (10, 2))


def valid_shape(k, r, adjoint):
    """At most 512 threads per tile, LDS within 160 KiB, at most 16 registers per state in the adjoint."""
    return k - r <= 9 and (2 ** k) * (16 if adjoint else 8) <= 160 * 1024 and (not adjoint or r <= 4)


def _time(fn, reps):
    fn()
    torch.cuda.synchronize()
    t0 = torch.cuda.Event(enable_timing=True)
    t1 = torch.cuda.Event(enable_timing=True)
    t0.record()
    for _ in range(reps):
        fn()
    t1.record()
    t1.synchronize()
    return t0.elapsed_time(t1) / reps


def autotune(ex, rows, fwd_shapes=FWD_SHAPES, bwd_shapes=BWD_SHAPES, reps=2, jit=False, log=None):
    """Configure ``ex`` (an SVExecutor) with the fastest shapes for ``rows`` ([B, P] on the GPU).

    With ``jit`` every candidate runs its generated per-pass kernels (compiled
    once per shape and kept by the executor); shapes those cannot serve are
    timed with the interpreted kernels.  ``log``, if given, is called with a
    line of text after every timed candidate.
    Returns {"fwd": {shape: ms}, "bwd": {shape: ms}, "best": (k, r, k_bwd, r_bwd)}.
    """
    n = ex.n
    dev = rows.device
    fwd_shapes = [s for s in fwd_shapes if s[0] <= n and valid_shape(*s, False)] or [(n, min(4, n))]
    bwd_shapes = [s for s in bwd_shapes if s[0] <= n and valid_shape(*s, True)] or [(n, min(4, n))]
    kb0, rb0 = ex.bwd.plan.k, ex.bwd.plan.r
    keep_jit = ex.jit
    ex.jit = False              # compile explicitly below, only what each timing needs
    fwd_ms, bwd_ms = {}, {}
    for k, r in fwd_shapes:
        ex.configure(k, r, kb0, rb0)
        if jit:
            ex.enable_jit(dev, forward=True, adjoint=False)
        fwd_ms[(k, r)] = _time(lambda: ex.run_forward(rows), reps)
        if log:
            log(f"autotune forward {k}/{r}: {fwd_ms[(k, r)]:.2f} ms")
    kf, rf = min(fwd_ms, key=fwd_ms.get)
    ex.configure(kf, rf, kb0, rb0)
    psi = ex.run_forward(rows)
    ups = [torch.ones(rows.shape[0], device=rows.device) for _ in ex.specs]
    lam = ex.seed_adjoint(psi, ups)
    for k, r in bwd_shapes:
        ex.configure(kf, rf, k, r)
        if jit:
            ex.enable_jit(dev, forward=False, adjoint=True)
        bwd_ms[(k, r)] = _time(lambda: ex.run_backward(rows, psi.clone(), lam.clone()), reps)
        if log:
            log(f"autotune adjoint {k}/{r}: {bwd_ms[(k, r)]:.2f} ms")
    kb, rb = min(bwd_ms, key=bwd_ms.get)
    best, best_ms = (kf, rf, kb, rb), fwd_ms[(kf, rf)] + bwd_ms[(kb, rb)]
    # same-k pairs with checkpoints: the adjoint reads psi from the forward's per-pass states (timed with the
    # same two state copies as above, so that the sums compare)
    ckpt_ms = {}
    if jit:
        for k in sorted({s[0] for s in fwd_ms} & {s[0] for s in bwd_ms}):
            rk = min((s[1] for s in fwd_ms if s[0] == k), key=lambda r: fwd_ms[(k, r)])
            for kb2, rb2 in [s for s in bwd_ms if s[0] == k]:
                ex.configure(k, rk, kb2, rb2)
                ex.enable_jit(dev)
                psi_c = ex.run_forward(rows, need_grad=True)
                if getattr(psi_c, "tedq_checkpoints", None) is None:
                    continue
                lam_c = ex.seed_adjoint(psi_c, ups)
                ms = _time(lambda: (psi_c.clone(), ex.run_backward(rows, psi_c, lam_c.clone())), reps)
                ckpt_ms[(k, rk, kb2, rb2)] = fwd_ms[(k, rk)] + ms
                if log:
                    log(f"autotune checkpointed {k}/{rk} + {kb2}/{rb2}: adjoint {ms:.2f} ms")
                if ckpt_ms[(k, rk, kb2, rb2)] < best_ms:
                    best, best_ms = (k, rk, kb2, rb2), ckpt_ms[(k, rk, kb2, rb2)]
                del psi_c, lam_c
    ex.configure(*best)
    ex.jit = keep_jit
    return {"fwd": fwd_ms, "bwd": bwd_ms, "checkpointed": ckpt_ms, "best": best}
