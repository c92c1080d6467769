"""ctypes binding of the gfx950 engine library (C ABI in include/tedq_hip.h).

The library is the only compute path: if it is missing or fails to load, every
entry point raises.  Device memory always comes from torch's caching
allocator; the library never allocates or frees.  Every call is enqueued on
the caller's current torch stream.
"""

import ctypes
import os

import torch

_LIB_PATH = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "_lib", "libtedq_hip.so")
_lib = None

c_int, c_i64, c_u64, c_float, c_vp = ctypes.c_int, ctypes.c_int64, ctypes.c_uint64, ctypes.c_float, ctypes.c_void_p

_SIGNATURES = {
    "tedq_abi_version": ([], c_int),
    "tedq_last_error": ([], ctypes.c_char_p),
    "tedq_device_info": ([c_vp], c_int),
    "tedq_sv_build_tables": ([c_vp, c_int, c_i64, c_int, c_vp, c_int, c_vp, c_vp, c_vp, c_i64, c_vp, c_vp, c_vp],
                             c_int),
    "tedq_sv_apply_pass": ([c_vp, c_i64, c_int, c_int, c_int, c_int, c_vp, c_u64, c_u64, c_vp, c_i64, c_vp], c_int),
    "tedq_sv_adjoint_pass": ([c_vp, c_vp, c_i64, c_int, c_int, c_int, c_int, c_vp, c_u64, c_u64, c_vp, c_i64, c_vp,
                              c_i64, c_int, c_vp], c_int),
    "tedq_sv_pass_ragged": ([c_int, c_vp, c_int, c_i64, c_int, c_int, c_vp], c_int),
    "tedq_sv_init_zero": ([c_vp, c_i64, c_int, c_vp], c_int),
    "tedq_sv_init_copy": ([c_vp, c_vp, c_i64, c_int, c_vp], c_int),
    "tedq_sv_reduce_chunks": ([c_int], c_i64),
    "tedq_sv_expval_z": ([c_vp, c_i64, c_int, c_vp, c_int, c_vp, c_vp, c_vp], c_int),
    "tedq_sv_expval_z_noalloc": ([c_vp, c_i64, c_int, c_vp, c_int, c_vp, c_vp], c_int),
    "tedq_sv_lambda_z": ([c_vp, c_vp, c_i64, c_int, c_vp, c_int, c_vp, c_vp], c_int),
    "tedq_sv_expval_factored": ([c_vp, c_i64, c_int, c_int, c_vp, c_vp, c_u64, c_vp, c_vp, c_vp], c_int),
    "tedq_sv_lambda_factored": ([c_vp, c_vp, c_i64, c_int, c_int, c_vp, c_vp, c_u64, c_vp, c_vp], c_int),
    "tedq_sv_probs": ([c_vp, c_i64, c_int, c_int, c_vp, c_vp, c_vp, c_vp], c_int),
    "tedq_sv_probs_chunks": ([c_int, c_int], c_i64),
    "tedq_sv_post_measure": ([c_vp, c_int, c_int, c_vp, c_vp, c_vp, c_vp], c_int),
    "tedq_sv_lambda_probs": ([c_vp, c_vp, c_i64, c_int, c_int, c_vp, c_vp, c_vp], c_int),
    "tedq_sv_axpy": ([c_vp, c_vp, c_i64, c_int, c_float, c_vp], c_int),
    "tedq_sum_partials": ([c_vp, c_i64, c_i64, c_int, c_vp, c_vp], c_int),
    "tedq_sv_grad_reduce": ([c_vp, c_i64, c_int, c_i64, c_vp, c_vp, c_int, c_vp, c_vp, c_int, c_vp], c_int),
    "tedq_jit_compile": ([ctypes.c_char_p, c_int, c_vp, c_vp, c_vp], c_int),
    "tedq_jit_function": ([c_vp, ctypes.c_char_p, c_vp], c_int),
    "tedq_jit_function_scratch": ([c_vp, c_vp], c_int),
    "tedq_jit_unload": ([c_vp], c_int),
    "tedq_jit_pass": ([c_vp, c_int, c_vp, c_vp, c_i64, c_int, c_int, c_int, c_u64, c_u64, c_vp, c_i64, c_vp, c_i64,
                       c_int, c_vp, c_int, c_vp], c_int),
    "tedq_tn_workspace_bytes": ([c_vp, c_i64], c_i64),
    "tedq_tn_contract": ([c_vp, c_i64, c_vp, c_i64, c_vp, c_i64, c_i64, c_vp, c_vp, c_vp, c_i64, c_vp], c_int),
    "tedq_tn_contract_small_batch": ([c_vp, c_vp, c_vp, c_vp, c_vp, c_int, c_i64, c_vp, c_vp, c_int, c_vp], c_int),
    "tedq_tn_gate_backward_workspace": ([c_vp, c_i64], c_i64),
    "tedq_tn_gate_backward": ([c_vp, c_i64, c_vp, c_i64, c_vp, c_i64, c_vp, c_i64, c_vp, c_i64, c_i64, c_vp, c_vp,
                               c_vp, c_i64, c_vp], c_int),
    "tedq_tn_gate_tensors": ([c_vp, c_int, c_i64, c_int, c_vp, c_int, c_vp, c_vp, c_i64, c_vp, c_i64, c_vp, c_vp],
                             c_int),
    "tedq_tn_param_grad": ([c_vp, c_i64, c_vp, c_i64, c_vp, c_int, c_vp, c_vp, c_i64, c_vp], c_int),
}


class EngineError(RuntimeError):
    """A library entry point returned a non-zero status."""


def library_path():
    return _LIB_PATH


def load():
    """Load (once) and return the engine library; raises if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_LIB_PATH):
        raise EngineError(
            f"tedq_amd HIP engine library not found at {_LIB_PATH}; build it with "
            f"`python ted-q_amd/build_hip.py` (or __graft_entry__.build())")
    lib = ctypes.CDLL(_LIB_PATH)
    for name, (args, res) in _SIGNATURES.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res
    if lib.tedq_abi_version() != 1:
        raise EngineError("tedq_amd HIP engine library has an unexpected ABI version")
    _lib = lib
    return lib


def exported_symbols():
    return list(_SIGNATURES)


def call(name, *args):
    """Invoke ``name`` and raise EngineError with the library's message on failure."""
    lib = load()
    rc = getattr(lib, name)(*args)
    if rc != 0:
        msg = lib.tedq_last_error().decode(errors="replace")
        raise EngineError(f"{name} failed ({rc}): {msg}")
    return rc


def ptr(t):
    """Raw device pointer of a tensor (None -> NULL)."""
    return None if t is None else ctypes.c_void_p(t.data_ptr())


def stream_handle(device):
    return ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)


def device_info():
    import numpy as np  # pylint: disable=import-outside-toplevel
    props = np.zeros(4, dtype=np.int32)
    call("tedq_device_info", props.ctypes.data_as(c_vp))
    return {"cu_count": int(props[0]), "lds_per_block": int(props[1]), "wave": int(props[2]),
            "arch": int(props[3])}
