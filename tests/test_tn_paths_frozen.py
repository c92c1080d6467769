"""Contraction paths, tensor layouts and step descriptors stay exactly as frozen in tests/golden.

The path search and the layout rules are deterministic; these fixtures pin
them bit-exactly on three networks (a 3x3 lattice of 9 networks, the 40-qubit
chain, one network of the 5x5 lattice), so a change that alters an order, an
intermediate's address-bit order or a descriptor word fails here before it
reaches a GPU.  Regenerate with tests/golden/make_tn_paths.py only when such
a change is intended.  No GPU is used: executors are planned on the host.
"""

import json

import numpy as np
import pytest

from golden.make_tn_paths import cases, describe, fixture_path
from tedq_amd.tn.executor import make_desc


@pytest.mark.parametrize("case", cases(), ids=lambda c: c[0])
def test_paths_layouts_and_descriptors_match_the_frozen_fixture(case):
    name, _, build = case
    with open(fixture_path(name)) as f:
        want = json.load(f)["networks"]
    ex = build()
    got = json.loads(json.dumps(describe(ex)))
    assert len(got) == len(want)
    for i, (g, w) in enumerate(zip(got, want)):
        assert g["n_inputs"] == w["n_inputs"], (name, i)
        assert g["path"] == w["path"], (name, i, "path")
        assert g["layouts"] == w["layouts"], (name, i, "layouts")
        assert g["descriptor_rows"] == w["descriptor_rows"], (name, i)
        assert g["descriptor_sha256"] == w["descriptor_sha256"], (name, i, "descriptors")
        assert list(g["strategy"]) == list(w["strategy"]), (name, i)


def test_frozen_descriptors_follow_from_the_frozen_layouts():
    """Every forward step descriptor of the frozen 3x3 lattice is what the descriptor encoder makes of the
    operand and result layouts in the fixture (the kernels read operands through exactly those bits)."""
    name, _, build = cases()[0]
    with open(fixture_path(name)) as f:
        nets = json.load(f)["networks"]
    ex = build()
    # pylint: disable=protected-access
    for net, rec in zip(ex.networks, nets):
        lay = [t.split(" ") if t else [] for t in rec["layouts"]]
        for s, st in enumerate(net.steps):
            a, b = rec["path"][s]
            c = rec["n_inputs"] + s
            want = make_desc(lay[a], lay[b], lay[c], int(ex._descs[st.fwd][-1]))
            assert np.array_equal(ex._descs[st.fwd], want)
