"""The allocation-free modulus step (pty_pmod1_noalloc) against pty_pmod1.

pty_pmod1_noalloc runs the same kernels as pty_pmod1 with its scratch in a
static per-GPU array instead of a context's allocations, so the two must agree
bit for bit: exits after the step and the eMod sum.  Two calls in a row, the
second on a smaller input (fewer frames, another frame side), then a call
without a mask, and the documented rejections.
"""

import ctypes

import numpy as np
import pytest
import torch

from ptychography_amd import _abi
from ptychography_amd.engine import Engine

pytestmark = pytest.mark.gpu


def _case(dev, N, M, K, masked, seed):
    rng = np.random.default_rng(seed)
    I = (rng.random((N, M, K)) * 50.0 + 1.0).astype(np.float32)
    mask_c = None
    if masked:
        mask_c = torch.from_numpy((rng.random((M, K)) > 0.1).astype(np.uint8)).to(dev)
    eng = Engine(N, M, K, M, K, np.zeros((N, 2), np.int32), device=dev)
    amp, I_norm = eng.prepare_amp(torch.from_numpy(I), mask_c)
    mask_s = None
    if masked:
        mask_s = torch.from_numpy(np.ascontiguousarray(
            np.fft.ifftshift(mask_c.cpu().numpy()))).to(dev)
    eng.set_data(amp, mask_s, I_norm)
    x = (rng.standard_normal((N, M, K)) + 1j * rng.standard_normal((N, M, K))).astype(np.complex64)
    return eng, amp, mask_s, torch.from_numpy(x).to(dev)


def _noalloc(x, amp, mask_s, N, M, K, alpha, dev):
    lib = _abi.load()
    emod = torch.full((1,), -1.0, dtype=torch.float64, device=dev)
    stream = torch.cuda.current_stream(dev).cuda_stream
    code = lib.pty_pmod1_noalloc(ctypes.c_void_p(x.data_ptr()), ctypes.c_void_p(amp.data_ptr()),
                                 None if mask_s is None else ctypes.c_void_p(mask_s.data_ptr()),
                                 N, M, K, alpha, ctypes.c_void_p(emod.data_ptr()),
                                 ctypes.c_void_p(stream))
    _abi.check("pty_pmod1_noalloc", code)
    torch.cuda.synchronize(dev)
    return float(emod.item())


@pytest.mark.parametrize("sizes", [[(96, 256, 256, True), (40, 128, 128, True), (7, 64, 32, False)]])
def test_pmod1_noalloc_matches_pmod1(gpu_device, sizes):
    dev = gpu_device
    for k, (N, M, K, masked) in enumerate(sizes):
        print(f"pty_pmod1_noalloc call {k}: N={N} M={M} K={K} masked={masked}", flush=True)
        eng, amp, mask_s, x = _case(dev, N, M, K, masked, seed=50 + k)
        alpha = 1e-10
        y_ref = x.clone()
        _, e_ref = eng.pmod1(y_ref, alpha)                 # the existing entry point
        y = x.clone()
        e = _noalloc(y, amp, mask_s, N, M, K, alpha, dev)
        eng.close()
        assert torch.equal(y, y_ref), f"call {k}: exits differ (max {float((y - y_ref).abs().max()):.3e})"
        assert e == e_ref, f"call {k}: eMod {e!r} vs {e_ref!r}"
        assert e > 0


def test_pmod1_noalloc_rejects_before_queueing(gpu_device):
    lib = _abi.load()
    x = torch.zeros(4, dtype=torch.complex64, device=gpu_device)
    a = torch.zeros(4, dtype=torch.float32, device=gpu_device)
    e = torch.zeros(1, dtype=torch.float64, device=gpu_device)
    p = [ctypes.c_void_p(t.data_ptr()) for t in (x, a, e)]
    assert lib.pty_pmod1_noalloc(p[0], p[1], None, 2, 100, 64, 1e-10, p[2], None) == _abi.PTY_EUNSUP
    assert lib.pty_pmod1_noalloc(p[0], p[1], None, 0, 64, 64, 1e-10, p[2], None) == _abi.PTY_EINVAL
    assert lib.pty_pmod1_noalloc(p[0], p[1], None, 1 << 20, 256, 256, 1e-10, p[2], None) == \
        _abi.PTY_EINVAL
    assert b"2^34" in lib.pty_last_error()
    assert lib.pty_pmod1_noalloc(None, p[1], None, 1, 64, 64, 1e-10, p[2], None) == _abi.PTY_EINVAL
