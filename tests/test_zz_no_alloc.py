"""The no-allocation variant of the Z-string expectation entry point, called through the C ABI.

tedq_sv_expval_z_noalloc keeps its partial sums in the library's fixed device
array instead of a caller buffer.  It runs the same kernels in the same
order as tedq_sv_expval_z, so its results must be bit-identical to that entry
point's on the same state, also on a second call with a smaller input (the
array keeps the first call's contents, which must not leak into the second).
"""

import numpy as np
import pytest
import torch

from tedq_amd.engine import hip

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda", 0)
LIMIT = 8 << 20


def _state(batch, n, seed):
    g = torch.Generator(device="cpu").manual_seed(seed)
    re = torch.randn(batch, 1 << n, generator=g, dtype=torch.float64)
    im = torch.randn(batch, 1 << n, generator=g, dtype=torch.float64)
    psi = torch.complex(re, im)
    psi = psi / psi.abs().pow(2).sum(1, keepdim=True).sqrt()
    return psi.to(torch.complex64).to(DEV)


def _masks(n, m, seed):
    rng = np.random.default_rng(seed)
    masks = [0] + [int(rng.integers(1, 1 << n)) for _ in range(m - 1)]
    return torch.tensor(np.array(masks, dtype=np.uint64).view(np.int64), device=DEV)


def _with_buffer(psi, n, masks):
    batch, m = psi.shape[0], masks.numel()
    partial = torch.empty(batch * hip.load().tedq_sv_reduce_chunks(n) * m, dtype=torch.float64, device=DEV)
    out = torch.empty(batch, m, dtype=torch.float64, device=DEV)
    hip.call("tedq_sv_expval_z", hip.ptr(psi), batch, n, hip.ptr(masks), m, hip.ptr(partial), hip.ptr(out),
             hip.stream_handle(DEV))
    return out


def _no_alloc(psi, n, masks):
    batch, m = psi.shape[0], masks.numel()
    out = torch.empty(batch, m, dtype=torch.float64, device=DEV)
    hip.call("tedq_sv_expval_z_noalloc", hip.ptr(psi), batch, n, hip.ptr(masks), m, hip.ptr(out),
             hip.stream_handle(DEV))
    return out


def _oracle(psi, n, masks):
    """<psi| Z-string(mask) |psi> in float64 on the host, for a sanity bound only."""
    p = psi.cpu().to(torch.complex128).abs().pow(2).numpy()
    x = np.arange(1 << n, dtype=np.uint64)
    out = []
    for mk in masks.cpu().numpy().view(np.uint64):
        par = np.array([bin(int(v)).count("1") & 1 for v in (x & mk)])
        out.append((p * np.where(par, -1.0, 1.0)).sum(1))
    return np.stack(out, 1)


def test_no_alloc_matches_buffered_entry_point_on_two_calls():
    # first call: 64 circuits of 20 qubits, 20 masks (64 x 64 chunks x 20 x 8 B = 655 KB of partial sums)
    print("no-alloc call 1: batch=64 n=20 masks=20", flush=True)
    psi1, m1 = _state(64, 20, 1), _masks(20, 20, 1)
    ref1 = _with_buffer(psi1, 20, m1)
    got1 = _no_alloc(psi1, 20, m1)
    torch.cuda.synchronize()
    assert torch.equal(got1, ref1)
    assert abs(float(got1[:, 0].min()) - 1.0) < 1e-6 and abs(float(got1[:, 0].max()) - 1.0) < 1e-6
    # second call: smaller batch, fewer qubits (the one-chunk kernel) and masks
    print("no-alloc call 2: batch=3 n=11 masks=5", flush=True)
    psi2, m2 = _state(3, 11, 2), _masks(11, 5, 2)
    ref2 = _with_buffer(psi2, 11, m2)
    got2 = _no_alloc(psi2, 11, m2)
    torch.cuda.synchronize()
    assert torch.equal(got2, ref2)
    assert np.max(np.abs(got2.cpu().numpy() - _oracle(psi2, 11, m2))) < 1e-6
    # a third, mid-sized call after the small one (the transformed kernel again, other sizes)
    print("no-alloc call 3: batch=5 n=16 masks=7", flush=True)
    psi3, m3 = _state(5, 16, 3), _masks(16, 7, 3)
    assert torch.equal(_no_alloc(psi3, 16, m3), _with_buffer(psi3, 16, m3))


def test_no_alloc_rejects_inputs_beyond_its_array():
    # 160 circuits of 22 qubits (256 chunks) with 32 masks need 10.5 MB of partial sums: more than the
    # 8 MiB array, so the call fails on the host and launches nothing (the state is still allocated at
    # its full size, 5.4 GB, so that no access could leave it)
    print("no-alloc oversize call: batch=160 n=22 masks=32", flush=True)
    psi = torch.zeros(160, 1 << 22, dtype=torch.complex64, device=DEV)
    masks = _masks(22, 32, 4)
    out = torch.empty(160, 32, dtype=torch.float64, device=DEV)
    lib = hip.load()
    assert 160 * lib.tedq_sv_reduce_chunks(22) * 32 * 8 > LIMIT
    rc = lib.tedq_sv_expval_z_noalloc(hip.ptr(psi), 160, 22, hip.ptr(masks), 32, hip.ptr(out),
                                      hip.stream_handle(DEV))
    assert rc != 0
    assert b"device scratch holds" in lib.tedq_last_error()
    # the array still works after the rejected call
    psi2, m2 = _state(2, 14, 5), _masks(14, 3, 5)
    assert torch.equal(_no_alloc(psi2, 14, m2), _with_buffer(psi2, 14, m2))
