"""Greedy pairwise contraction order for the circuit tensor networks.

The reference obtains its order from external optimisers (opt_einsum
``contract_expression(optimize='auto')`` in tedq/oe_wrapper/oe_wrapper.py:57,
JDOptTN or cotengra, tedq/backends/compiled_circuit.py:340-404), none of which
is available here, so the order is this module's own, deterministic greedy
rule, and a computed path is frozen with the network it belongs to:

* candidates are pairs of live tensors that share at least one index;
* the pair minimising ``size(result) - size(a) - size(b)`` is contracted
  first (sizes are element counts, every index has dimension 2); ties go to
  the smaller ``(min(id), max(id))`` pair of SSA ids;
* an index is summed when it occurs in the two operands, in no other live
  tensor and not in the requested output;
* when no pair shares an index, the remaining tensors are joined by outer
  products, smallest first.

The result is an SSA path: step s contracts ids (i, j) and creates id
``n_inputs + s``.
"""

import heapq
import math
import random


def _cost(a, b, keep_counts, out_set, alpha=1.0):
    union = a | b
    summed = {ix for ix in (a & b) if keep_counts[ix] == 2 and ix not in out_set}
    res = union - summed
    return (1 << len(res)) - alpha * ((1 << len(a)) + (1 << len(b))), frozenset(res)


def greedy_path(inputs, output, alpha=1.0, temperature=0.0, seed=0, absorb=0):
    """SSA contraction path and the index set of every intermediate.

    ``inputs``: list of index lists; ``output``: list of kept indices.
    Returns (path, sets) where ``sets[id]`` is the frozenset of indices of
    tensor ``id`` (inputs first, then one per step).  ``alpha`` weighs the
    operand sizes in the cost (the rule above is alpha = 1); a positive
    ``temperature`` perturbs each candidate's cost by seeded Gumbel noise
    scaled by its magnitude, so that a set of seeds explores nearby orders.
    With ``absorb`` > 0 the path first merges every tensor of rank <= 2
    (single-qubit gates, basis states) into a neighbour, in id order, while
    the merged tensor's rank stays <= ``absorb``: a one-qubit gate then rides
    on a two-qubit gate instead of being applied to a large intermediate.
    With ``absorb`` < 0 the same merge happens lazily: only when the greedy
    rule picks a rank <= 2 tensor against an intermediate of rank >= 8, and
    a neighbour of rank <= 4 gives a merged rank <= -``absorb``; the rest of
    the order is the plain greedy one.
    """
    rng = random.Random(seed)

    def noisy(c):
        if temperature <= 0.0 or c == 0:
            return c
        return c - temperature * abs(c) * math.log(-math.log(rng.random() or 1e-300))

    out_set = set(output)
    sets = [frozenset(t) for t in inputs]
    counts = {}
    for t in sets:
        for ix in t:
            counts[ix] = counts.get(ix, 0) + 1
    holders = {}
    for i, t in enumerate(sets):
        for ix in t:
            holders.setdefault(ix, set()).add(i)
    live = set(range(len(sets)))
    heap = []

    def push(i, j):
        if i > j:
            i, j = j, i
        c, _ = _cost(sets[i], sets[j], counts, out_set, alpha)
        heapq.heappush(heap, (noisy(c), i, j))

    seen = set()
    # candidate pairs are pushed in index order: with noise, the draw a pair gets depends on the
    # push order, which must not follow the (per-process) hash order of the index symbols
    for ix in sorted(holders):
        hs = sorted(holders[ix])
        for a in range(len(hs)):
            for b in range(a + 1, len(hs)):
                if (hs[a], hs[b]) not in seen:
                    seen.add((hs[a], hs[b]))
                    push(hs[a], hs[b])
    path = []

    def contract(i, j):
        _, res = _cost(sets[i], sets[j], counts, out_set, alpha)
        new = len(sets)
        sets.append(res)
        path.append((min(i, j), max(i, j)))
        live.difference_update((i, j))
        for ix in sets[i] | sets[j]:
            holders[ix].discard(i)
            holders[ix].discard(j)
            counts[ix] -= (ix in sets[i]) + (ix in sets[j])
        for ix in res:
            holders[ix].add(new)
            counts[ix] += 1
        live.add(new)
        return new, res

    if absorb > 0:
        changed = True
        while changed:
            changed = False
            for t in sorted(live):
                if t not in live or len(sets[t]) > 2:
                    continue
                neigh = sorted(set().union(*(holders[ix] for ix in sets[t])) - {t}) if sets[t] else []
                for u in neigh:
                    if len(_cost(sets[t], sets[u], counts, out_set)[1]) <= absorb:
                        contract(t, u)
                        changed = True
                        break
        heap.clear()
        seen = set()
        for ix in sorted(holders):
            hs = sorted(holders[ix])
            for a in range(len(hs)):
                for b in range(a + 1, len(hs)):
                    if (hs[a], hs[b]) not in seen:
                        seen.add((hs[a], hs[b]))
                        push(hs[a], hs[b])
    while heap:
        c, i, j = heapq.heappop(heap)
        if i not in live or j not in live:
            continue
        if absorb < 0:
            small, large = (i, j) if len(sets[i]) <= len(sets[j]) else (j, i)
            if len(sets[small]) <= 2 and len(sets[large]) >= 8:
                near = sorted(set().union(*(holders[ix] for ix in sets[small])) - {small, large})
                for u in near:
                    if len(sets[u]) <= 4 and len(_cost(sets[small], sets[u], counts, out_set)[1]) <= -absorb:
                        i, j = small, u
                        break
        new, res = contract(i, j)
        neigh = set()
        for ix in res:
            neigh |= holders[ix]
        neigh.discard(new)
        for m in sorted(neigh):
            push(m, new)
    # outer products of what is left, smallest first
    rest = sorted(live, key=lambda t: (len(sets[t]), t))
    while len(rest) > 1:
        i, j = rest[0], rest[1]
        _, res = _cost(sets[i], sets[j], counts, out_set)
        new = len(sets)
        sets.append(res)
        path.append((min(i, j), max(i, j)))
        rest = sorted(rest[2:] + [new], key=lambda t: (len(sets[t]), t))
    return path, sets


def sweep_path(inputs, output):
    """Contract the tensors into one accumulator in input order (circuit order)."""
    out_set = set(output)
    sets = [frozenset(t) for t in inputs]
    counts = {}
    for t in sets:
        for ix in t:
            counts[ix] = counts.get(ix, 0) + 1
    path = []
    acc = 0
    for t in range(1, len(inputs)):
        a, b = sets[acc], sets[t]
        _, res = _cost(a, b, counts, out_set)
        for ix in a | b:
            counts[ix] -= (ix in a) + (ix in b)
        for ix in res:
            counts[ix] += 1
        sets.append(res)
        path.append((acc, t))
        acc = len(sets) - 1
    return path, sets


# (alpha, temperature, seed, absorb) trials of the greedy rule, plus the sweep.  The plain greedy
# rule at four weights and four noisy variants, and the four weights again with single-qubit
# tensors merged into neighbours first (absorb 4) or when they would meet a large intermediate
# (absorb -4).  Which wins differs between networks: eager merging gives the 40-qubit chain 7.7x
# fewer multiply-adds; lazy merging halves the 5x5 lattice's large steps.
_GREEDY_SETTINGS = ((1.0, 0.0, 0), (2.0, 0.0, 0), (4.0, 0.0, 0), (8.0, 0.0, 0), (2.0, 0.05, 0), (2.0, 0.05, 1),
                    (8.0, 0.05, 0), (1.0, 0.2, 0))
SEARCH_TRIALS = (tuple(t + (0,) for t in _GREEDY_SETTINGS)
                 + tuple(t + (m,) for m in (4, -4) for t in _GREEDY_SETTINGS if t[1] == 0.0))


# device model used to rank candidate orders: HBM bytes at 5 TB/s, complex multiply-adds (8 flops)
# at 100 TFLOP/s on GEMM-shaped steps (matrix cores) and 20 TFLOP/s on the rest
_MODEL_BYTES_S = 5e12
_MODEL_GEMM_FLOPS = 1e14
_MODEL_VALU_FLOPS = 2e13


def path_seconds(path, sets, n_inputs):
    """Modelled device time of a path: per step the larger of its operand+result traffic and its
    arithmetic time (every step of a large network is memory- or compute-bound, so the number of
    multiply-adds alone undercounts steps that stream a large tensor past a small one)."""
    total = 0.0
    for s, (i, j) in enumerate(path):
        a, b, c = sets[i], sets[j], sets[n_inputs + s]
        nm, nn, nk = len((a - b) & c), len((b - a) & c), len((a & b) - c)
        gemm = nm >= 6 and nn >= 6 and nk >= 4
        byts = 8.0 * ((1 << len(a)) + (1 << len(b)) + (1 << len(c)))
        flops = 8.0 * (1 << len(a | b))
        total += max(byts / _MODEL_BYTES_S, flops / (_MODEL_GEMM_FLOPS if gemm else _MODEL_VALU_FLOPS))
    return total


def search_path(inputs, output, trials=SEARCH_TRIALS):
    """Deterministic path search: the cheapest of a fixed set of candidate orders.

    Candidates are the greedy rule under each (alpha, temperature, seed,
    absorb) of ``trials`` and the circuit-order sweep; the winner minimises
    the modelled device time (``path_seconds``), ties broken by the largest
    intermediate.  Returns (path, sets, info).
    """
    n = len(inputs)
    best = None
    cands = [("sweep", None)] + [("greedy", t) for t in trials]
    for kind, t in cands:
        if n < 2:
            path, sets = [], [frozenset(x) for x in inputs]
        elif kind == "sweep":
            path, sets = sweep_path(inputs, output)
        else:
            path, sets = greedy_path(inputs, output, *t)
        largest, _ = path_stats(path, sets, n)
        macs = sum(1 << len(sets[i] | sets[j]) for i, j in path)
        secs = path_seconds(path, sets, n)
        key = (secs, largest)
        if best is None or key < best[0]:
            best = (key, path, sets, {"strategy": kind, "trial": t, "macs": macs, "largest_rank": largest,
                                      "model_s": secs})
    return best[1], best[2], best[3]


def path_stats(path, sets, n_inputs):
    """log2 of the largest intermediate and log10 of the multiply-add count (complex MACs)."""
    import math  # pylint: disable=import-outside-toplevel
    largest = 0
    flops = 0
    for s, (i, j) in enumerate(path):
        res = sets[n_inputs + s]
        largest = max(largest, len(res))
        flops += 1 << len(sets[i] | sets[j])
    return largest, math.log10(max(flops, 1))


def linear_to_ssa(path, n_inputs):
    """opt_einsum-style path (positions in the shrinking operand list, result appended) -> SSA path."""
    ids = list(range(n_inputs))
    nxt = n_inputs
    ssa = []
    for step in path:
        if len(step) != 2:
            raise ValueError("only pairwise contraction steps are supported")
        i, j = sorted(step)
        a, b = ids[i], ids[j]
        del ids[j]
        del ids[i]
        ids.append(nxt)
        ssa.append((min(a, b), max(a, b)))
        nxt += 1
    return ssa


def path_sets(inputs, output, ssa):
    """Index sets of every tensor along an SSA path (validated: each tensor used exactly once)."""
    out_set = set(output)
    sets = [frozenset(t) for t in inputs]
    counts = {}
    for t in sets:
        for ix in t:
            counts[ix] = counts.get(ix, 0) + 1
    used = set()
    for i, j in ssa:
        if i in used or j in used or i == j or max(i, j) >= len(sets):
            raise ValueError(f"contraction path step ({i}, {j}) is not a valid SSA step")
        used |= {i, j}
        _, res = _cost(sets[i], sets[j], counts, out_set)
        for ix in sets[i] | sets[j]:
            counts[ix] -= (ix in sets[i]) + (ix in sets[j])
        for ix in res:
            counts[ix] += 1
        sets.append(res)
    if len(ssa) != max(len(inputs) - 1, 0) or (ssa and len(used) != len(sets) - 1):
        raise ValueError("contraction path does not contract every tensor into one")
    return sets


def external_path(tree, n_inputs):
    """SSA path from a contraction tree or path object of an external optimiser.

    Accepts what cotengra's ``HyperOptimizer.search`` and JDtensorPath return
    (objects with ``get_ssa_path()`` or ``get_path()``), opt_einsum-style
    objects with a ``path`` attribute (linear), or a plain list of pairs
    (linear).  The indices such a tree slices (``sliced_inds``) are read
    separately (tn/executor.py, ``sliced_indices``) and sliced the same way.
    """
    if hasattr(tree, "get_ssa_path"):
        return [tuple(int(x) for x in st) for st in tree.get_ssa_path()]
    if hasattr(tree, "get_path"):
        return linear_to_ssa([tuple(st) for st in tree.get_path()], n_inputs)
    if hasattr(tree, "path"):
        return linear_to_ssa([tuple(st) for st in tree.path], n_inputs)
    return linear_to_ssa([tuple(st) for st in tree], n_inputs)
