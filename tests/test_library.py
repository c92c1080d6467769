"""The C-ABI library loads and exports every symbol declared in include/tedq_hip.h (no GPU needed)."""

import ctypes
import os
import re

import pytest

from tedq_amd.engine import hip

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "tedq_hip.h")


def declared_symbols():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?[a-z0-9_]+\*?\s+\*?(tedq_[a-z0-9_]+)\(", text, re.M)))


@pytest.mark.skipif(not os.path.exists(hip.library_path()), reason="library not built")
def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(hip.library_path())
    syms = declared_symbols()
    assert len(syms) >= 17
    for name in syms:
        assert hasattr(lib, name), name
    # the ctypes binding covers exactly the declared ABI
    assert set(hip.exported_symbols()) == set(syms)


@pytest.mark.skipif(not os.path.exists(hip.library_path()), reason="library not built")
def test_library_rejects_bad_arguments_without_gpu_work():
    lib = hip.load()
    assert lib.tedq_abi_version() == 1
    rc = lib.tedq_sv_init_zero(None, 4, 10, None)
    assert rc != 0
    assert b"bad arguments" in lib.tedq_last_error()
    assert lib.tedq_sv_reduce_chunks(20) == 64
