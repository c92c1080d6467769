"""Freeze the contraction paths and index layouts of three networks as fixtures.

The reference takes its contraction order from opt_einsum / cotengra /
JDtensorPath (tedq/backends/compiled_circuit.py:340-404, tedq/oe_wrapper/
oe_wrapper.py:57), none of which exists here, so path parity with the
reference cannot be pinned.  What is pinned instead is this engine's own
order: for each network, the SSA path, the index layout (address-bit order)
of every tensor, and a digest of every step descriptor the kernels read.  A
test rebuilds the executors on the CPU and compares against these files
bit-exactly, so any drift of the path search or of the layout rules shows.

Run from the repository root:  python tests/golden/make_tn_paths.py
"""

import hashlib
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
for _p in (ROOT, os.path.join(ROOT, "ted-q_amd")):
    if _p not in sys.path:
        sys.path.insert(0, _p)


def cases():
    """(name, description, builder) of the frozen networks."""
    from tedq_amd import workloads as W  # pylint: disable=import-outside-toplevel
    return [
        ("lattice3x3_l2", "3x3 lattice, 2 layers of RX RZ + CZ bonds, <Z_i> for all 9 qubits (9 networks)",
         lambda: W.lattice(3, 2, tn_mode=True).executor()),
        ("chain40_s8", "40-qubit disordered chain, 8 Trotter steps, <Z_20> (the batched contraction workload)",
         lambda: W.disordered_chain(40, 8, 20, tn_mode=True).executor()),
        ("lattice5x5_l6_z12", "5x5 lattice, 6 layers, <Z_12> (one network of the lattice workload)",
         lambda: W.lattice(5, 6, observe=[12], tn_mode=True).executor()),
    ]


def describe(ex):
    """Path, layouts and descriptor digests of every network of a built TNExecutor."""
    # pylint: disable=protected-access
    nets = []
    for net in ex.networks:
        h = hashlib.sha256()
        rows = []
        for st in net.steps:
            for r in (st.fwd, st.bwd_a, st.bwd_b):
                rows.append(r)
                if r >= 0:
                    h.update(ex._descs[r].tobytes())
                    g = ex._gemm.get(r)
                    if g is not None:
                        for rr in (g[0], g[2], g[4]):
                            if rr >= 0:
                                h.update(ex._descs[rr].tobytes())
        nets.append({
            "n_inputs": net.n_inputs,
            "strategy": [net.path_info["strategy"], net.path_info["trial"]],
            "path": [list(p) for p in net.path],
            "layouts": [" ".join(t) for t in net.lay],
            "descriptor_rows": rows,
            "descriptor_sha256": h.hexdigest(),
        })
    return nets


def fixture_path(name):
    return os.path.join(HERE, f"paths_{name}.json")


def main():
    for name, what, build in cases():
        rec = {"workload": what, "networks": describe(build())}
        with open(fixture_path(name), "w") as f:
            json.dump(rec, f, separators=(",", ":"))
        print(f"wrote {fixture_path(name)}: {len(rec['networks'])} networks, "
              f"{sum(len(n['path']) for n in rec['networks'])} steps")


if __name__ == "__main__":
    main()
