"""Build the gfx950 HIP engine library ``tedq_amd/_lib/libtedq_hip.so`` in-tree.

Every ``csrc/*.hip`` file is compiled with hipcc for gfx950 only and linked
into one shared library whose C ABI is declared in ``include/tedq_hip.h``.
Packed-FP32 instructions are disabled for all device code.
"""

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT_DIR = os.path.join(HERE, "tedq_amd", "_lib")
OUT = os.path.join(OUT_DIR, "libtedq_hip.so")
INCLUDE = os.path.join(os.path.dirname(HERE), "include")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics", "-Wall", "-Wno-unused-function",
    # keep per-bit switch arms separate: sinking their common code turns the
    # register-array indices into runtime values and the arrays into scratch
    "-mllvm", "-simplifycfg-sink-common=false",
]


def sources():
    return sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".hip"))


def build(verbose=True):
    """Compile and link the library; returns its path.  Raises on failure."""
    os.makedirs(OUT_DIR, exist_ok=True)
    objs = []
    for src in sources():
        obj = os.path.join(OUT_DIR, os.path.basename(src)[:-4] + ".o")
        cmd = [HIPCC, *[f for f in FLAGS if f != "-shared"], "-c", "-I", CSRC, "-I", INCLUDE, src, "-o", obj]
        if verbose:
            print("[build]", os.path.basename(src), flush=True)
        subprocess.run(cmd, check=True)
        objs.append(obj)
    tmp = OUT + ".tmp"
    # hiprtc compiles the per-circuit pass kernels at run time (csrc/sv_jit.hip)
    subprocess.run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", *objs, "-o", tmp,
                    "-L/opt/rocm/lib", "-lhiprtc", "-Wl,-rpath,/opt/rocm/lib"], check=True)
    os.replace(tmp, OUT)
    for obj in objs:
        os.remove(obj)
    return OUT


if __name__ == "__main__":
    sys.exit(0 if build() else 1)
