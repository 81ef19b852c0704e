"""GPU HIO (ptychography_amd/hio.py) against the HIO oracle (oracle/hio.py).

Bars as for the ERA / DM solvers (tests/test_gpu_solvers.py): iterates within
max(1e-5, 10 x the complex64 oracle's drift from the complex128 oracle), eMod
per iteration within max(1e-4, 10 x that drift), final eMod within 1 %.  HIO
with beta near 1 feeds round-off back like DM does, so the drift bound is the
one that binds after a few iterations.  Also on the GPU path: Fienup's classic
rule for a 0/1 support projection (binary probe, frames that do not overlap),
within float32 round-off (rel L2 <= 1e-5), and the exact-data fixed point.
"""

import numpy as np
import pytest
import torch

import oracle
import ptychography_amd as pty
from oracle.era import pmod_1
from tests.scans import rel_l2, small_scan
from tests.test_gpu_solvers import _O0, _check_metrics, _check_traj, _illuminated, _run_both

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("update,beta,frame,masked", [
    ("sample", 0.9, 32, False), ("sample", 1.0, 64, True), ("probe", 1.0, 32, False),
    ("both", 0.7, 32, False)])
def test_hio_against_oracle(update, beta, frame, masked, gpu_device):
    print(f"HIO update={update} beta={beta} frame={frame} masked={masked}", flush=True)
    s = small_scan(frame=frame, grid=(5, 5), step=frame * 3 // 16, jitter=1,
                   beamstop_radius=1.5 if masked else 0.0)
    O0 = s.O if update == "probe" else _O0(s)
    P0 = s.P if update == "sample" else np.roll(s.P, (1, 2), axis=(0, 1)) * 0.9
    kw = dict(beta=beta, update=update)
    if masked:
        kw["mask"] = s.mask
    out_g, out_r, tg, tr, single = _run_both(pty.HIO, oracle.HIO, (s.I, s.R, P0, O0), kw, 10)
    region = _illuminated(s.P, s.R, s.O.shape)
    _check_traj(tg, tr, region, {"sample": "O", "probe": "P", "both": "OP"}[update], single)
    _check_metrics(out_g[-1], out_r[-1], single[0][-1], keys=("eMod",))


def test_hio_support_projection_is_fienups_rule(gpu_device):
    print("HIO: classic input-output rule (binary probe, separate frames)", flush=True)
    rng = np.random.default_rng(0)
    M = 32
    off = np.array([[0, 0], [0, 32], [32, 7], [32, 40]])
    R = -off
    P = (rng.random((M, M)) > 0.4).astype(np.complex64)
    O = (rng.standard_normal((64, 72)) + 1j * rng.standard_normal((64, 72))).astype(np.complex64)
    I = (rng.random((4, M, M)) * 4.0 + 0.5).astype(np.float32)
    beta = 0.8
    _, _, info = pty.HIO(I, R.copy(), P, O, 1, beta=beta, update="sample")
    f0 = oracle.make_exits(O.astype(np.complex128), P.astype(np.complex128), R)
    amp = np.fft.ifftshift(np.sqrt(I.astype(np.float64)), axes=(-2, -1))
    m = pmod_1(amp, f0, 1, alpha=1e-10)
    want = np.where((P != 0)[None], m, f0 - beta * m)
    err = rel_l2(info["exits"], want)
    print(f"exits rel L2 against the classic rule: {err:.2e}")
    assert err <= 1e-5


def test_hio_exact_data_fixed_point(gpu_device):
    print("HIO: exact-data fixed point", flush=True)
    s = small_scan(frame=64, grid=(4, 4), step=12, jitter=1)
    ex = oracle.make_exits(s.O, s.P, s.R - s.R.max(axis=0))
    I = np.abs(np.fft.fftshift(np.fft.fftn(ex, axes=(-2, -1)), axes=(-2, -1))) ** 2
    O, P, info = pty.HIO(I.astype(np.float32), s.R.copy(), s.P, s.O, 4, beta=0.9)
    print("eMod:", info["eMod"])
    assert max(info["eMod"]) <= 2e-5
    region = _illuminated(s.P, s.R, s.O.shape)
    assert rel_l2(O[region], s.O[region]) <= 1e-5


def test_hio_argument_errors(gpu_device):
    s = small_scan(frame=32, grid=(2, 2), step=8)
    with pytest.raises(ValueError):
        pty.HIO(s.I, s.R.copy(), s.P, None, 1, update="object")
    with pytest.raises(ValueError):
        pty.HIO(s.I, s.R.copy(), s.P, None, 1, beta=0.0)


def test_hio_entry_points_reject_bad_arguments(gpu_device):
    # the C entry points check their arguments before any device work
    import ctypes
    from ptychography_amd._abi import PtyError, call
    from ptychography_amd.engine import Engine, _ptr
    print("HIO entry points: argument checks", flush=True)
    s = small_scan(frame=32, grid=(2, 2), step=8)
    N, M, K = s.I.shape
    eng = Engine(N, M, K, *s.O.shape, s.R.astype(np.int32), device=gpu_device)
    try:
        amp, I_norm = eng.prepare_amp(torch.from_numpy(s.I.astype(np.float32)), None)
        eng.set_data(amp, None, I_norm)
        O = torch.from_numpy(s.O.astype(np.complex64)).to(gpu_device)
        P = torch.from_numpy(s.P.astype(np.complex64)).to(gpu_device)
        F = eng.make_exits(O, P)
        m = torch.empty_like(F)
        e = ctypes.c_double()
        bad = [(F, m, O, P, 0.0, 0.0),            # beta = 0
               (F, m, O, P, 0.0, float("nan")),   # beta not finite
               (F, F, O, P, 0.0, 1.0),            # f aliases m
               (F, m, None, P, 0.0, 1.0),         # no object
               (F, m, O, P, float("inf"), 1.0)]   # w not finite
        for f_, m_, O_, P_, w, beta in bad:
            with pytest.raises(PtyError) as err:
                call("pty_hio_modulus", eng._ctx, _ptr(f_), _ptr(m_), _ptr(O_), _ptr(P_), w, beta,
                     1e-10, ctypes.byref(e))
            assert err.value.code == -1
        with pytest.raises(PtyError):
            call("pty_add_exits", eng._ctx, _ptr(O), None, 1.0, _ptr(F))
        # a good call after the rejected ones still works
        call("pty_hio_modulus", eng._ctx, _ptr(F), _ptr(m), _ptr(O), _ptr(P), 0.0, 1.0, 1e-10,
             ctypes.byref(e))
        assert np.isfinite(e.value)
    finally:
        eng.close()
