"""GPU parity of the building blocks against the NumPy oracle (complex128).

Tolerances: complex64 kernels against complex128 references, relative L2
error <= 2e-6 for one 2-D FFT and <= 1e-5 for a projection step.
"""

import numpy as np
import pytest
import torch

import oracle
from ptychography_amd.engine import Engine
from tests.scans import rel_l2, small_scan

pytestmark = pytest.mark.gpu

SIZES = [(16, 16), (32, 32), (64, 64), (96, 96), (128, 128), (192, 192), (256, 256),
         (64, 128), (128, 32), (96, 256)]


def _engine_for(M, K, N=3, dev=None):
    U, V = M + 4, K + 6
    R = -np.array([[0, 0], [2, 3], [4, 6]][:N], dtype=np.int32)
    return Engine(N, M, K, U, V, R, device=dev), U, V, R


@pytest.mark.parametrize("M,K", SIZES)
def test_fft2_forward_matches_numpy(M, K, gpu_device):
    print(f"fft2 forward {M}x{K}", flush=True)
    eng, *_ = _engine_for(M, K, dev=gpu_device)
    rng = np.random.default_rng(M * 1000 + K)
    x = (rng.standard_normal((5, M, K)) + 1j * rng.standard_normal((5, M, K))).astype(np.complex64)
    d = torch.from_numpy(x).to(gpu_device)
    eng.fft2(d)
    ref = np.fft.fftn(x.astype(np.complex128), axes=(-2, -1))
    assert rel_l2(d.cpu().numpy(), ref) < 2e-6
    eng.fft2(d, inverse=True)
    assert rel_l2(d.cpu().numpy(), x) < 3e-6
    eng.close()


@pytest.mark.parametrize("M,K", [(16, 16), (64, 64), (256, 256), (96, 192)])
def test_fft2_inverse_matches_numpy(M, K, gpu_device):
    print(f"fft2 inverse {M}x{K}", flush=True)
    eng, *_ = _engine_for(M, K, dev=gpu_device)
    rng = np.random.default_rng(7)
    x = (rng.standard_normal((3, M, K)) + 1j * rng.standard_normal((3, M, K))).astype(np.complex64)
    d = torch.from_numpy(x).to(gpu_device)
    eng.fft2(d, inverse=True)
    ref = np.fft.ifftn(x.astype(np.complex128), axes=(-2, -1))
    assert rel_l2(d.cpu().numpy(), ref) < 2e-6
    eng.close()


def test_fft2_impulse_and_single_frequency(gpu_device):
    """Closed forms: delta -> constant; a pure frequency -> a single spike."""
    eng, *_ = _engine_for(64, 64, dev=gpu_device)
    x = np.zeros((2, 64, 64), np.complex64)
    x[0, 0, 0] = 1.0
    i, j = np.meshgrid(np.arange(64), np.arange(64), indexing="ij")
    x[1] = np.exp(2j * np.pi * (5 * i + 9 * j) / 64)
    d = torch.from_numpy(x).to(gpu_device)
    eng.fft2(d)
    y = d.cpu().numpy()
    assert np.allclose(y[0], 1.0, atol=1e-6)
    spike = np.zeros((64, 64))
    spike[5, 9] = 64 * 64
    assert np.abs(y[1] - spike).max() < 64 * 64 * 1e-5
    eng.close()


def test_make_exits_matches_oracle(gpu_device):
    s = small_scan()
    N, M, K = s.I.shape
    R = s.R.astype(np.int32)
    eng = Engine(N, M, K, *s.O.shape, R, device=gpu_device)
    O = torch.from_numpy(s.O.astype(np.complex64)).to(gpu_device)
    P = torch.from_numpy(s.P.astype(np.complex64)).to(gpu_device)
    ex = eng.make_exits(O, P).cpu().numpy()
    ref = oracle.make_exits(s.O, s.P, s.R)
    assert rel_l2(ex, ref) < 1e-6
    eng.close()


def _bound_engine(s, mask, dev):
    N, M, K = s.I.shape
    eng = Engine(N, M, K, *s.O.shape, s.R.astype(np.int32), device=dev)
    mc = None if mask is None else torch.from_numpy(mask.astype(np.uint8)).to(dev)
    ms = None if mask is None else torch.from_numpy(np.fft.ifftshift(mask).astype(np.uint8)).to(dev)
    amp, I_norm = eng.prepare_amp(torch.from_numpy(s.I.astype(np.float32)), mc)
    eng.set_data(amp, ms, I_norm)
    return eng, amp, I_norm


@pytest.mark.parametrize("masked", [False, True])
def test_prepare_amp_and_pmod1_match_oracle(masked, gpu_device):
    s = small_scan(frame=64, grid=(3, 3), step=16)
    N, M, K = s.I.shape
    mask = None
    if masked:
        mask = np.random.default_rng(1).random((M, K)) > 0.1
    eng, amp, I_norm = _bound_engine(s, mask, gpu_device)
    mref = 1 if mask is None else mask
    assert abs(I_norm - np.sum(mref * s.I)) <= 1e-6 * np.sum(s.I)
    amp_ref = np.fft.ifftshift(np.sqrt(s.I), axes=(-2, -1))
    assert rel_l2(amp.cpu().numpy(), amp_ref) < 1e-7
    ex0 = oracle.make_exits(s.O * 0.9, s.P, s.R)
    d = torch.from_numpy(ex0.astype(np.complex64)).to(gpu_device)
    _, emod = eng.pmod1(d, alpha=1e-10)
    ms = 1 if mask is None else np.fft.ifftshift(mask)
    ref, emod_ref = oracle.pmod_1(amp_ref, ex0.copy(), ms, alpha=1e-10, eMod_calc=True)
    assert rel_l2(d.cpu().numpy(), ref) < 1e-5
    assert abs(emod - emod_ref) <= 1e-5 * abs(emod_ref)
    eng.close()


def test_psup_O_and_P_match_oracle(gpu_device):
    s = small_scan(frame=64, grid=(4, 5), step=12, jitter=2)
    N, M, K = s.I.shape
    eng = Engine(N, M, K, *s.O.shape, s.R.astype(np.int32), device=gpu_device)
    rng = np.random.default_rng(5)
    ex = (rng.standard_normal((N, M, K)) + 1j * rng.standard_normal((N, M, K)))
    ex_t = torch.from_numpy(ex.astype(np.complex64)).to(gpu_device)
    P_t = torch.from_numpy(s.P.astype(np.complex64)).to(gpu_device)
    O_gpu, H = eng.psup_O(ex_t, P_t)
    O_ref, H_ref = oracle.psup_O(ex, s.P, s.R, s.O.shape, None, 1e-10)
    cov = H_ref > 1e-6 * H_ref.max()
    assert rel_l2(H.cpu().numpy(), H_ref) < 1e-6
    assert rel_l2(O_gpu.cpu().numpy()[cov], O_ref[cov]) < 1e-5
    # cached heatmap path gives the same object
    O_gpu2, _ = eng.psup_O(ex_t, P_t, H=H)
    assert torch.equal(O_gpu2, O_gpu)
    O_t = torch.from_numpy(s.O.astype(np.complex64)).to(gpu_device)
    P_gpu, G = eng.psup_P(ex_t, O_t)
    P_ref, G_ref = oracle.psup_P(ex, s.O, s.R, None, 1e-10)
    assert rel_l2(G.cpu().numpy(), G_ref) < 1e-6
    assert rel_l2(P_gpu.cpu().numpy(), P_ref) < 1e-5
    eng.close()
