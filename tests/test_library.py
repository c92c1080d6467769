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
    src = jit.program_source(prog.enc, len(prog.enc.fwd_offsets), len(prog.enc.adj_offsets))
    path = tmp_path / "gen.hip"
    path.write_text(src)
    csrc = os.path.join(ROOT, "ted-q_amd", "csrc")
    out = subprocess.run([hipcc, "--offload-arch=gfx950", "-O3", "-std=c++17", "-Xclang", "-target-feature",
                          "-Xclang", "-packed-fp32-ops", "--cuda-device-only", "-c", "-I", csrc, str(path), "-o",
                          str(tmp_path / "gen.o")], capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stderr[-2000:]
