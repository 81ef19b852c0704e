"""hipGraph replay of iterations gives bit-identical results to eager launches,
across several pty_run calls with eager torch work and synchronisations in
between (each call replays cached graphs)."""

import numpy as np
import pytest
import torch

from ptychography_amd.engine import Engine
from tests.scans import broadband_start, flat_start, small_scan

pytestmark = pytest.mark.gpu


def _run(s, dev, algorithm, method, graphs, P0, chunks=(3, 4, 5), bg=None, **kw):
    N, M, K = s.I.shape
    eng = Engine(N, M, K, *s.O.shape, s.R.astype(np.int32), device=dev)
    eng.set_graphs(graphs)
    amp, I_norm = eng.prepare_amp(torch.from_numpy(s.I.astype(np.float32)), None)
    eng.set_data(amp, None, I_norm)
    O_t = torch.from_numpy(broadband_start(s.O.shape).astype(np.complex64)).to(dev)
    P_t = torch.from_numpy(P0.astype(np.complex64)).to(dev)
    eng.set_state(O_t, P_t)
    if bg is not None:
        eng.set_background(torch.from_numpy(bg.astype(np.float32)), per_frame=False)
    rows, it = [], 0
    for ci, n in enumerate(chunks):
        rows.append(eng.run(algorithm, method, n, iter0=it, begin=(ci == 0), **kw))
        it += n
        junk = torch.randn(1000, device=dev).sum()   # eager work between replays
        torch.cuda.synchronize()
        del junk
    out = (np.concatenate(rows), O_t.cpu().numpy(), P_t.cpu().numpy())
    eng.close()
    return out


@pytest.mark.parametrize("frame", [64, 128])
def test_era_graphs_bit_identical(frame, gpu_device):
    print(f"graphs ERA frame {frame}", flush=True)
    s = small_scan(frame=frame, grid=(4, 4), step=frame // 5, jitter=1)
    P0 = flat_start(s)
    kw = dict(OP_iters=(1, 2), Pmod_probe_until=5)
    a = _run(s, gpu_device, 0, 3, False, P0, **kw)
    b = _run(s, gpu_device, 0, 3, True, P0, **kw)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)


@pytest.mark.parametrize("frame", [32, 64])
def test_dm_background_graphs_bit_identical(frame, gpu_device):
    # at 64^2 the leading projection runs fused into its inverse row FFT
    # (k_rows_gather writing f) inside the replayed graphs
    print(f"graphs DM method 4 frame {frame}", flush=True)
    s = small_scan(frame=frame, grid=(4, 4), step=frame * 3 // 16, jitter=1, background_frac=0.02)
    bg = np.sqrt(np.random.default_rng(2).random((frame, frame)) * 0.5)
    a = _run(s, gpu_device, 1, 4, False, s.P, bg=bg)
    b = _run(s, gpu_device, 1, 4, True, s.P, bg=bg)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)


def test_dm_centering_graphs_bit_identical(gpu_device):
    print("graphs DM method 3 with centring", flush=True)
    s = small_scan(frame=32, grid=(4, 4), step=6, jitter=1)
    P0 = np.roll(flat_start(s), (2, 3), axis=(0, 1))
    kw = dict(OP_iters=(1, 1), probe_centering=True)
    a = _run(s, gpu_device, 1, 3, False, P0, **kw)
    b = _run(s, gpu_device, 1, 3, True, P0, **kw)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)


def test_failed_capture_falls_back_to_eager(gpu_device, monkeypatch):
    """A capture that fails (forced from the host) leaves nothing half-run:
    the run finishes eagerly with the same results and reports the fallback."""
    import ptychography_amd as pty
    s = small_scan(frame=64, grid=(4, 4), step=16)
    O0 = broadband_start(s.O.shape)
    O_e, P_e, i_e = pty.ERA(s.I, s.R.copy(), s.P, O0, 6, method=3, device=gpu_device)
    monkeypatch.setenv("PTY_FORCE_CAPTURE_FAILURE", "1")
    O_g, P_g, i_g = pty.ERA(s.I, s.R.copy(), s.P, O0, 6, method=3, device=gpu_device, graphs=True)
    res = pty.ERA_many([dict(I=s.I, R=s.R.copy(), P=s.P, O=O0, method=3)], 6, graphs=True,
                       device=gpu_device)
    monkeypatch.delenv("PTY_FORCE_CAPTURE_FAILURE")
    assert i_g["graphs"] == {"enabled": False, "fell_back": True}
    assert np.array_equal(O_g, O_e) and np.array_equal(P_g, P_e) and i_g["eMod"] == i_e["eMod"]
    O_b, P_b, i_b = res[0]
    assert i_b["graphs"] == {"enabled": False, "fell_back": True}
    assert np.array_equal(O_b, O_e) and i_b["eMod"] == i_e["eMod"]


def test_non_finite_data_raise(gpu_device):
    import ptychography_amd as pty
    s = small_scan(frame=32, grid=(3, 3), step=8)
    I = s.I.copy()
    I[2, 5, 7] = np.nan
    with pytest.raises(FloatingPointError):
        pty.ERA(I, s.R.copy(), s.P, None, 3, method=1, device=gpu_device)
    O, info = pty.ERA(I, s.R.copy(), s.P, None, 3, method=1, device=gpu_device, check_finite=False)
    assert not np.isfinite(info["eMod"][0])


@pytest.mark.parametrize("graphs", [False, True])
@pytest.mark.parametrize("method,op,centre,pmodp", [
    (1, (1, 1), False, 0), (3, (1, 2), False, 0),
    # two non-OP iterations in a row (the reuse happens), with the centring
    # of the OP iterations resetting the heatmaps and the far-field probe
    # constraint changing the metric probe only
    (3, (1, 3), True, 5), (3, (2, 3), True, 5), (1, (1, 1), False, 5),
    # object + probe every iteration: each OP projection is the previous
    # iteration's metric projection (heatmaps included), also around the end
    # of a Pmod_probe window
    (3, (1, 1), False, 0), (3, (1, 1), False, 4)])
def test_dm_projection_reuse_is_exact(method, op, centre, pmodp, graphs, gpu_device):
    # one run of 8 iterations starts each object-only projection from the
    # previous iteration's metric projection; 8 runs of one iteration never do
    # (every run clears it): the iterates and metrics must agree bit for bit
    print(f"DM method {method} OP_iters {op} centring {centre} Pmod_probe {pmodp} "
          f"graphs={graphs}: projection reuse", flush=True)
    s = small_scan(frame=64, grid=(4, 4), step=12, jitter=1)
    P0 = s.P if method == 1 else np.roll(flat_start(s), (2, 3), axis=(0, 1))
    kw = dict(OP_iters=op, probe_centering=centre, Pmod_probe_until=pmodp)
    a = _run(s, gpu_device, 1, method, graphs, P0, chunks=(8,), **kw)
    b = _run(s, gpu_device, 1, method, graphs, P0, chunks=(1,) * 8, **kw)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)
