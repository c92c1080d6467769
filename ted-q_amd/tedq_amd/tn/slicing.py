"""Sliced contraction: fix a few bond indices and sum the networks they leave.

The reference hands slicing to its path finders (``slicing_opts`` passed to
cotengra's HyperOptimizer or to JDtensorPath, tedq/backends/
compiled_circuit.py:340-393; doc/source/PathOptimizer.rst:38-63): a sliced
index is fixed to each of its values in turn, every intermediate loses that
index, and the results of the 2^s sliced contractions are added.  Peak
memory drops by up to 2^s at the price of repeated work.

Here a slice is expressed inside the network itself, so that every slice
runs the same frozen path, layouts and step descriptors: a bond index ``s``
shared by tensors T1 ... Tm (m = 2 for a plain bond, more for a hyper-index
left by diagonal reduction) is renamed in T2 ... Tm, and m selector vectors
e_v (|0> or |1>, one on each copy) are appended as leaves and contracted into
their tensors before the original path continues.  Summing over v the
product of the selectors gives back the bond (the delta tensor of the
hyper-index), so the sum over all assignments equals the unsliced network.  The selector
vectors live in the executor's constant arena; between slices only their
values change.

Index choice (``choose_slices``) is greedy on the frozen path: while the
largest intermediate exceeds the target (or there are fewer slices than
asked), slice the index that shrinks the largest intermediate most, ties
going to the one whose sliced total work is smallest, then to the smallest
symbol.  Output indices are never sliced.
"""

from .network import get_symbol


def _widest(inter, sliced):
    return max((len(t - sliced) for t in inter), default=0)


def _work(path, sets, sliced):
    """Complex multiply-adds of all slices together along ``path``."""
    per = sum(1 << len((sets[i] | sets[j]) - sliced) for i, j in path)
    return per << len(sliced)


def choose_slices(path, sets, n_inputs, output, target_size=None, target_slices=None, max_sliced=20):
    """Greedy slice selection on a path (see module docstring); returns a list of index symbols.

    ``target_size``: largest number of elements an intermediate may keep;
    ``target_slices``: least number of slices (a power of two is reached).
    """
    if target_size is None and target_slices is None:
        return []
    out = set(output)
    inter = [sets[n_inputs + s] for s in range(len(path))]
    counts = {}
    for t in sets[:n_inputs]:
        for ix in t:
            counts[ix] = counts.get(ix, 0) + 1
    bonds = {ix for ix, c in counts.items() if c >= 2 and ix not in out}
    sliced = set()
    while len(sliced) < max_sliced:
        cur = _widest(inter, sliced)
        size_ok = target_size is None or (1 << cur) <= target_size
        count_ok = target_slices is None or (1 << len(sliced)) >= target_slices
        if size_ok and count_ok:
            break
        cands = sorted({ix for t in inter if len(t - sliced) == cur for ix in t} & (bonds - sliced))
        if not cands:
            cands = sorted(bonds - sliced)
        if not cands:
            break
        best = None
        for ix in cands:
            trial = sliced | {ix}
            key = (_widest(inter, trial), _work(path, sets, trial), ix)
            if best is None or key < best:
                best = key
        sliced.add(best[2])
    return sorted(sliced)


def slice_network(inputs, sources, output, path, sliced):
    """The network with selector leaves for ``sliced`` indices and its SSA path.

    Returns (inputs, sources, path, selector_of_leaf) where the new sources
    of the selector leaves are ("select", j) for the j-th sliced index.
    """
    if not sliced:
        return [list(t) for t in inputs], list(sources), [tuple(p) for p in path]
    out = set(output)
    new_inputs = [list(t) for t in inputs]
    used = {ix for t in inputs for ix in t} | out
    nxt = 0
    holders = {}
    for t, ix_list in enumerate(inputs):
        for ix in ix_list:
            holders.setdefault(ix, []).append(t)
    caps = []           # (tensor it caps, index symbol, slice number)
    for j, s in enumerate(sliced):
        hs = holders.get(s, [])
        if s in out or len(hs) < 2 or len(set(hs)) != len(hs):
            raise ValueError(f"index {s!r} is not a bond between tensors and cannot be sliced")
        # every holder but the first gets its own copy of the index; a selector on each copy (the same
        # vector for all of them) fixes the value, summing the products over v gives back the bond (or
        # the hyper-index's delta)
        caps.append((hs[0], s, j))
        for t in hs[1:]:
            while get_symbol(nxt) in used:
                nxt += 1
            s2 = get_symbol(nxt)
            used.add(s2)
            new_inputs[t] = [s2 if x == s else x for x in new_inputs[t]]
            caps.append((t, s2, j))
    n0 = len(inputs)
    new_sources = list(sources)
    for _, sym, j in caps:
        new_inputs.append([sym])
        new_sources.append(("select", j))
    n1 = n0 + len(caps)
    cur = {t: t for t in range(n0)}      # current SSA id holding original leaf t
    new_path = []
    for c, (t, _, _) in enumerate(caps):
        new_path.append((min(cur[t], n0 + c), max(cur[t], n0 + c)))
        cur[t] = n1 + len(new_path) - 1
    shift = n1 + len(caps) - n0

    def remap(i):
        return cur[i] if i < n0 else i + shift
    for i, j in path:
        a, b = remap(i), remap(j)
        new_path.append((min(a, b), max(a, b)))
    return new_inputs, new_sources, new_path
