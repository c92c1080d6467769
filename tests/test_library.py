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


@pytest.mark.skipif(not os.path.exists(hip.library_path()), reason="library not built")
def test_no_alloc_expval_checks_its_array_size_on_the_host():
    """The no-allocation expectation entry point refuses inputs whose partial sums exceed its fixed 8 MiB
    device array before touching the device (the pointers below are never dereferenced)."""
    lib = hip.load()
    dummy = ctypes.c_void_p(16)
    rc = lib.tedq_sv_expval_z_noalloc(dummy, 160, 22, dummy, 32, dummy, None)
    assert rc != 0 and b"device scratch holds 8388608" in lib.tedq_last_error()
    rc = lib.tedq_sv_expval_z_noalloc(None, 4, 10, dummy, 2, dummy, None)
    assert rc != 0 and b"bad arguments" in lib.tedq_last_error()


@pytest.mark.skipif(not os.path.exists(hip.library_path()), reason="library not built")
def test_contraction_entry_points_validate_on_the_host():
    """Launch planning and argument checks of the contraction entry points run on the host: workspace
    sizes of split and unsplit steps, the gate-backward eligibility rule, and records of a small-step
    batch that leave their arenas are rejected before any launch."""
    import numpy as np
    from tedq_amd.tn.executor import _gate_step, make_desc
    lib = hip.load()

    def desc(la, lb, lc):
        return make_desc(list(la), list(lb), list(lc))

    def ptr(a):
        return a.ctypes.data_as(hip.c_vp)
    # gate environment summed over 2^16 terms: 2^11 workgroup partials of 16 outputs
    d = desc("ab" + "cdefghijklmnopqr", "st" + "cdefghijklmnopqr", "abst")
    assert lib.tedq_tn_workspace_bytes(ptr(d), 2) == 8 * 2 * 16 << 4
    # a one-tile 128x128 GEMM with a 2^12 sum splits k 64 ways; a wide one does not split
    d = desc("abcdefg" + "hijklmnopqrs", "hijklmnopqrs" + "tuvwxyz", "abcdefgtuvwxyz")
    assert lib.tedq_tn_workspace_bytes(ptr(d), 3) == 8 * 3 * (1 << 14) * 64
    d = desc("abcdefghij" + "klmn", "klmn" + "opqrstuvwx", "abcdefghijopqrstuvwx")
    assert lib.tedq_tn_workspace_bytes(ptr(d), 32) == 0
    # the fused gate backward accepts gates on >= 2^10 rows, and the Python rule agrees
    g = desc("abcdefghijkl" + "xy", "xyuv", "abcdefghijkl" + "uv")
    assert lib.tedq_tn_gate_backward_workspace(ptr(g), 1) == 8 * (1 << 2) * 16 and _gate_step(g)
    ng = desc("abcdefgh" + "xy", "xyuv", "abcdefgh" + "uv")
    assert lib.tedq_tn_gate_backward_workspace(ptr(ng), 1) == -1 and not _gate_step(ng)
    # a small-step record whose result leaves arena 2 is refused without a launch
    d = desc("ab", "bc", "ac")
    descs = np.ascontiguousarray(d[None, :])
    recs = np.array([[0, 0, 0, 0, 4, 2, 4, 0]], dtype=np.int64)       # result at offset 2, 4 elements
    fake = ctypes.c_void_p(16)                                           # never dereferenced on the host
    bases = (ctypes.c_void_p * 3)(16, 16, 16)
    strides = (ctypes.c_int64 * 3)(8, 0, 4)
    lens = (ctypes.c_int64 * 3)(8, 8, 4)                                 # arena 2 holds 4 elements only
    rc = lib.tedq_tn_contract_small_batch(bases, strides, lens, ptr(recs), fake, 1, 1, ptr(descs), fake, 1, None)
    assert rc != 0 and b"does not fit" in lib.tedq_last_error()


def test_generated_pass_kernels_compile_for_gfx950(tmp_path):
    """The per-pass kernel generator emits HIP that hipcc accepts for gfx950 (device code only)."""
    import shutil
    import subprocess

    import tedq_amd as qai
    from tedq_amd.engine import jit
    from tedq_amd.engine.sv import SVExecutor
    from tedq_amd.engine.lowering import lower_circuit

    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(hipcc):
        pytest.skip("hipcc not available")

    def cd(*p):
        qai.RX(p[0], qubits=[0])
        qai.CRY(p[1], qubits=[3, 1])
        qai.Rot(p[2], p[3], p[4], qubits=[5])
        qai.Toffoli(qubits=[0, 5, 7])
        qai.CZ(qubits=[2, 6])
        qai.RZ(p[5], qubits=[8])
        qai.PhaseShift(p[6], qubits=[4]).inverse()
        qai.CNOT(qubits=[8, 2])
        return [qai.expval(qai.PauliZ(qubits=[q])) for q in range(9)]

    cc = qai.Circuit(cd, 9, *([0.1] * 7)).compilecircuit(backend="pytorch")
    n, units, P, pool = lower_circuit(cc)
    ex = SVExecutor(n, units, P, pool, cc._measurements, k_max=6, r=3)  # pylint: disable=protected-access
    prog = ex.fwd
    assert jit.supported(n, prog.plan.k, prog.plan.r, True)
    csrc = os.path.join(ROOT, "ted-q_amd", "csrc")
    # the one-tile-per-workgroup kernels, and the adjoint kernels that loop over tiles with the fused seed
    for tag, kw in (("plain", {}), ("persistent", {"seed_masks": tuple(ex.zmasks), "persistent": True})):
        src = jit.program_source(prog.enc, len(prog.enc.fwd_offsets), len(prog.enc.adj_offsets), **kw)
        assert ("TILE += gridDim.x" in src) == (tag == "persistent")
        path = tmp_path / f"gen_{tag}.hip"
        path.write_text(src)
        out = subprocess.run([hipcc, "--offload-arch=gfx950", "-O3", "-std=c++17", "-Xclang", "-target-feature",
                              "-Xclang", "-packed-fp32-ops", "--cuda-device-only", "-c", "-I", csrc, str(path), "-o",
                              str(tmp_path / f"gen_{tag}.o")], capture_output=True, text=True, timeout=300)
        assert out.returncode == 0, out.stderr[-2000:]
