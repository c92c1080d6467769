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


def greedy_path(inputs, output, alpha=1.0, temperature=0.0, seed=0):
    """SSA contraction path and the index set of every intermediate.

    ``inputs``: list of index lists; ``output``: list of kept indices.
    Returns (path, sets) where ``sets[id]`` is the frozenset of indices of
    tensor ``id`` (inputs first, then one per step).  ``alpha`` weighs the
    operand sizes in the cost (the rule above is alpha = 1); a positive
    ``temperature`` perturbs each candidate's cost by seeded Gumbel noise
    scaled by its magnitude, so that a set of seeds explores nearby orders.
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
    for ix, hs in holders.items():
        hs = sorted(hs)
        for a in range(len(hs)):
            for b in range(a + 1, len(hs)):
                if (hs[a], hs[b]) not in seen:
                    seen.add((hs[a], hs[b]))
                    push(hs[a], hs[b])
    path = []
    while heap:
        c, i, j = heapq.heappop(heap)
        if i not in live or j not in live:
            continue
        _, res = _cost(sets[i], sets[j], counts, out_set, alpha)
        new = len(sets)
        sets.append(res)
        path.append((i, j))
        live -= {i, j}
        for ix in sets[i] | sets[j]:
            holders[ix].discard(i)
            holders[ix].discard(j)
            counts[ix] -= (ix in sets[i]) + (ix in sets[j])
        for ix in res:
            holders[ix].add(new)
            counts[ix] += 1
        live.add(new)
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


# (alpha, temperature, seed) trials of the greedy rule, plus the sweep
SEARCH_TRIALS = ((1.0, 0.0, 0), (2.0, 0.0, 0), (4.0, 0.0, 0), (8.0, 0.0, 0), (2.0, 0.05, 0), (2.0, 0.05, 1),
                 (8.0, 0.05, 0), (1.0, 0.2, 0))


def search_path(inputs, output, trials=SEARCH_TRIALS):
    """Deterministic path search: the cheapest of a fixed set of candidate orders.

    Candidates are the greedy rule under each (alpha, temperature, seed) of
    ``trials`` and the circuit-order sweep; the winner minimises the number
    of complex multiply-adds, ties broken by the largest intermediate.
    Returns (path, sets, info).
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
        key = (macs, largest)
        if best is None or key < best[0]:
            best = (key, path, sets, {"strategy": kind, "trial": t, "macs": macs, "largest_rank": largest})
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
    (linear).  Sliced trees are contracted whole.
    """
    if hasattr(tree, "get_ssa_path"):
        return [tuple(int(x) for x in st) for st in tree.get_ssa_path()]
    if hasattr(tree, "get_path"):
        return linear_to_ssa([tuple(st) for st in tree.get_path()], n_inputs)
    if hasattr(tree, "path"):
        return linear_to_ssa([tuple(st) for st in tree.path], n_inputs)
    return linear_to_ssa([tuple(st) for st in tree], n_inputs)
